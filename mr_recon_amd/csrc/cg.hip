// Conjugate-gradient vector kernels with FP64 inner products.
//
// The reference CG (algs.py:779-872) runs x0 = AHb, r = AHb - (AHA x0 + lam x0),
// p = r and per iteration: Ap; pAp = Re<p, Ap>; alpha = <r,r>/pAp; x += alpha p;
// r -= alpha Ap; beta = <r_new,r_new>/<r,r>; p = r + beta p, with complex64 sums
// and a host sync per step. Here the sums accumulate in FP64 and every scalar
// stays on the device, so an iteration needs no host round trip and can be
// captured in a graph. Vector updates stay in FP32 (alpha, beta cast to fp32).
//
// Device scalar block `sc` (double[8]):
//   sc[0] rz (= <r,r>), sc[1] pAp, sc[2] rz of the previous iteration, sc[3] beta.
// Once the recursive residual underflows to exactly zero the step sizes are set
// to 0 (the iterate stays put) rather than 0/0.
#include "common.h"
#include "capi_util.h"

namespace mra {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double block_sum(double v) {
  __shared__ double sh[kRedThreads / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kRedThreads / 64; ++w) t += sh[w];
  return t;  // valid in thread 0
}

// part[blk] = Re sum conj(a) * (b + lam*a)       (b may equal a)
__global__ __launch_bounds__(kRedThreads) void cdot_partial_kernel(const float2* __restrict__ a,
                                                                   const float2* __restrict__ b, long long n,
                                                                   float lam, double* __restrict__ part) {
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 av = a[i], bv = b[i];
    const double bx = (double)bv.x + (double)lam * av.x, by = (double)bv.y + (double)lam * av.y;
    s += (double)av.x * bx + (double)av.y * by;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out = sum part  (deterministic order); optional post-processing:
//   mode 0: sc[slot] = sum
//   mode 1: rz update: sc[3] = sum / sc[0]; sc[2] = sc[0]; sc[0] = sum
__global__ __launch_bounds__(kRedThreads) void reduce_final_kernel(const double* __restrict__ part, int nparts,
                                                                   double* __restrict__ sc, int slot, int mode) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  const double t = block_sum(s);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      sc[slot] = t;
    } else {
      // a residual that has reached exactly zero gives beta = 0 instead of 0/0
      sc[3] = sc[0] > 0.0 ? t / sc[0] : 0.0;
      sc[2] = sc[0];
      sc[0] = t;
    }
  }
}

// alpha = sc[0]/sc[1]; x += alpha p; r -= alpha (Ap + lam p); part = |r|^2 partials
__global__ __launch_bounds__(kRedThreads) void cg_xr_kernel(float2* __restrict__ x, float2* __restrict__ r,
                                                            const float2* __restrict__ p,
                                                            const float2* __restrict__ ap, long long n, float lam,
                                                            const double* __restrict__ sc,
                                                            double* __restrict__ part) {
  const float alpha = sc[1] > 0.0 ? (float)(sc[0] / sc[1]) : 0.0f;  // p = 0 once converged exactly
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 pv = p[i], av = ap[i];
    float2 xv = x[i], rv = r[i];
    xv.x += alpha * pv.x;
    xv.y += alpha * pv.y;
    rv.x -= alpha * (av.x + lam * pv.x);
    rv.y -= alpha * (av.y + lam * pv.y);
    x[i] = xv;
    r[i] = rv;
    s += (double)rv.x * rv.x + (double)rv.y * rv.y;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// p = r + beta p with beta = sc[3]
__global__ __launch_bounds__(kRedThreads) void cg_p_kernel(float2* __restrict__ p, const float2* __restrict__ r,
                                                           long long n, const double* __restrict__ sc) {
  const float beta = (float)sc[3];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 rv = r[i], pv = p[i];
    p[i] = make_float2(rv.x + beta * pv.x, rv.y + beta * pv.y);
  }
}

// r = b - (Ax + lam x); p = r; part = |r|^2 partials
__global__ __launch_bounds__(kRedThreads) void cg_init_kernel(const float2* __restrict__ b,
                                                              const float2* __restrict__ ax,
                                                              const float2* __restrict__ x, float2* __restrict__ r,
                                                              float2* __restrict__ p, long long n, float lam,
                                                              double* __restrict__ part) {
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 bv = b[i], av = ax[i], xv = x[i];
    const float2 rv = make_float2(bv.x - (av.x + lam * xv.x), bv.y - (av.y + lam * xv.y));
    r[i] = rv;
    p[i] = rv;
    s += (double)rv.x * rv.x + (double)rv.y * rv.y;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// ---- segmented variants: nseg independent systems of seg_len elements each (batched
// problems of one image size). blockIdx.y = segment; partials part[seg][kSegBlocks];
// scalars sc[seg][8] with the same slots as above.
constexpr int kSegBlocks = 32;

__global__ __launch_bounds__(kRedThreads) void seg_cdot_partial_kernel(const float2* __restrict__ a,
                                                                       const float2* __restrict__ b,
                                                                       long long seg_len, float lam,
                                                                       double* __restrict__ part) {
  const long long base = (long long)blockIdx.y * seg_len;
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < seg_len;
       i += (long long)gridDim.x * blockDim.x) {
    const float2 av = a[base + i], bv = b[base + i];
    const double bx = (double)bv.x + (double)lam * av.x, by = (double)bv.y + (double)lam * av.y;
    s += (double)av.x * bx + (double)av.y * by;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.y * kSegBlocks + blockIdx.x] = t;
}

__global__ __launch_bounds__(kRedThreads) void seg_reduce_final_kernel(const double* __restrict__ part, int nparts,
                                                                       double* __restrict__ sc, int slot, int mode) {
  const int seg = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[seg * kSegBlocks + i];
  const double t = block_sum(s);
  if (threadIdx.x == 0) {
    double* c = sc + seg * 8;
    if (mode == 0) {
      c[slot] = t;
    } else {
      c[3] = c[0] > 0.0 ? t / c[0] : 0.0;
      c[2] = c[0];
      c[0] = t;
    }
  }
}

__global__ __launch_bounds__(kRedThreads) void seg_cg_xr_kernel(float2* __restrict__ x, float2* __restrict__ r,
                                                                const float2* __restrict__ p,
                                                                const float2* __restrict__ ap, long long seg_len,
                                                                float lam, const double* __restrict__ sc,
                                                                double* __restrict__ part) {
  const double* c = sc + blockIdx.y * 8;
  const float alpha = c[1] > 0.0 ? (float)(c[0] / c[1]) : 0.0f;
  const long long base = (long long)blockIdx.y * seg_len;
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < seg_len;
       i += (long long)gridDim.x * blockDim.x) {
    const long long j = base + i;
    const float2 pv = p[j], av = ap[j];
    float2 xv = x[j], rv = r[j];
    xv.x += alpha * pv.x;
    xv.y += alpha * pv.y;
    rv.x -= alpha * (av.x + lam * pv.x);
    rv.y -= alpha * (av.y + lam * pv.y);
    x[j] = xv;
    r[j] = rv;
    s += (double)rv.x * rv.x + (double)rv.y * rv.y;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.y * kSegBlocks + blockIdx.x] = t;
}

__global__ __launch_bounds__(kRedThreads) void seg_cg_p_kernel(float2* __restrict__ p, const float2* __restrict__ r,
                                                               long long seg_len, const double* __restrict__ sc) {
  const float beta = (float)sc[blockIdx.y * 8 + 3];
  const long long base = (long long)blockIdx.y * seg_len;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < seg_len;
       i += (long long)gridDim.x * blockDim.x) {
    const float2 rv = r[base + i], pv = p[base + i];
    p[base + i] = make_float2(rv.x + beta * pv.x, rv.y + beta * pv.y);
  }
}

__global__ __launch_bounds__(kRedThreads) void seg_cg_init_kernel(const float2* __restrict__ b,
                                                                  const float2* __restrict__ ax,
                                                                  const float2* __restrict__ x,
                                                                  float2* __restrict__ r, float2* __restrict__ p,
                                                                  long long seg_len, float lam,
                                                                  double* __restrict__ part) {
  const long long base = (long long)blockIdx.y * seg_len;
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < seg_len;
       i += (long long)gridDim.x * blockDim.x) {
    const long long j = base + i;
    const float2 bv = b[j], av = ax[j], xv = x[j];
    const float2 rv = make_float2(bv.x - (av.x + lam * xv.x), bv.y - (av.y + lam * xv.y));
    r[j] = rv;
    p[j] = rv;
    s += (double)rv.x * rv.x + (double)rv.y * rv.y;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.y * kSegBlocks + blockIdx.x] = t;
}

// ---- one CG step as a differentiable function (the unrolled CG's iteration) --------------
// Forward, out of place: ap = inv * nrm + lam * p (nrm = the SENSE normal operator applied to
// p), pAp = Re<p, ap>, alpha = rz / pAp, x1 = x + alpha p, r1 = r - alpha ap, rz1 = <r1, r1>,
// beta = rz1 / rz, p1 = r1 + beta p; dots in FP64, alpha and beta cast to fp32 for the vector
// updates. Scalars sd (double[8]): sd[0] rz, sd[1] pAp, sd[2] rz1, sd[3] beta, sd[4] alpha.
// Backward: the adjoint of these operations for the cotangents (gx1, gr1, gp1, grz1), with
// gbeta = Re<gp1, p>, grz1' = grz1 + gbeta / rz, g = gr1 + gp1 + 2 grz1' r1,
// galpha = Re<gx1, p> - Re<g, ap>, gpAp = -galpha rz / pAp^2:
//   gr = g, gap = gpAp p - alpha g, gp = beta gp1 + alpha gx1 + gpAp ap + lam gap,
//   gnrm = inv gap, grz = -gbeta rz1 / rz^2 + galpha / pAp   (gx = gx1).
// A step whose pAp (rz) is not positive has alpha (beta) = 0 and passes no gradient through it.

__device__ __forceinline__ float2 cg_ap(const float2* __restrict__ nrm, const float2* __restrict__ p, long long i,
                                        float inv, float lam) {
  const float2 nv = nrm[i], pv = p[i];
  return make_float2(nv.x * inv + lam * pv.x, nv.y * inv + lam * pv.y);
}

__device__ __forceinline__ float2 ld_or0(const float2* __restrict__ a, long long i) {
  return a ? a[i] : make_float2(0.f, 0.f);
}

// four block sums at once (valid in thread 0)
__device__ __forceinline__ void block_sum4(double (&v)[4]) {
  __shared__ double sh4[4][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_down(v[k], o);
    if (lane == 0) sh4[k][wid] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double t = 0.0;
      for (int w = 0; w < kRedThreads / 64; ++w) t += sh4[k][w];
      v[k] = t;
    }
}

__global__ __launch_bounds__(kRedThreads) void cgd_pap_kernel(const float2* __restrict__ p,
                                                              const float2* __restrict__ nrm, long long n,
                                                              float inv, float lam, double* __restrict__ part) {
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 pv = p[i], av = cg_ap(nrm, p, i, inv, lam);
    s += (double)pv.x * av.x + (double)pv.y * av.y;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kRedThreads) void cgd_alpha_kernel(const double* __restrict__ part, int nparts,
                                                                const double* __restrict__ rz,
                                                                double* __restrict__ sd) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  const double t = block_sum(s);
  if (threadIdx.x == 0) {
    const double z = rz[0];
    sd[0] = z;
    sd[1] = t;
    sd[4] = t > 0.0 ? z / t : 0.0;
  }
}

__global__ __launch_bounds__(kRedThreads) void cgd_xr_kernel(const float2* __restrict__ x,
                                                             const float2* __restrict__ r,
                                                             const float2* __restrict__ p,
                                                             const float2* __restrict__ nrm, long long n, float inv,
                                                             float lam, const double* __restrict__ sd,
                                                             float2* __restrict__ x1, float2* __restrict__ r1,
                                                             double* __restrict__ part) {
  const float alpha = (float)sd[4];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 pv = p[i], av = cg_ap(nrm, p, i, inv, lam), xv = x[i], rv = r[i];
    const float2 xo = make_float2(xv.x + alpha * pv.x, xv.y + alpha * pv.y);
    const float2 ro = make_float2(rv.x - alpha * av.x, rv.y - alpha * av.y);
    x1[i] = xo;
    r1[i] = ro;
    s += (double)ro.x * ro.x + (double)ro.y * ro.y;
  }
  const double t = block_sum(s);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kRedThreads) void cgd_beta_kernel(const double* __restrict__ part, int nparts,
                                                               double* __restrict__ sd, double* __restrict__ rz1) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) s += part[i];
  const double t = block_sum(s);
  if (threadIdx.x == 0) {
    sd[2] = t;
    sd[3] = sd[0] > 0.0 ? t / sd[0] : 0.0;
    rz1[0] = t;
  }
}

__global__ __launch_bounds__(kRedThreads) void cgd_p_kernel(const float2* __restrict__ r1,
                                                            const float2* __restrict__ p, long long n,
                                                            const double* __restrict__ sd,
                                                            float2* __restrict__ p1) {
  const float beta = (float)sd[3];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 rv = r1[i], pv = p[i];
    p1[i] = make_float2(rv.x + beta * pv.x, rv.y + beta * pv.y);
  }
}

// part[k * nblk + blk]: k = 0 Re<gp1, p>, 1 Re<gr1 + gp1, ap>, 2 Re<r1, ap>, 3 Re<gx1, p>
__global__ __launch_bounds__(kRedThreads) void cgd_bwd_dots_kernel(
    const float2* __restrict__ p, const float2* __restrict__ nrm, const float2* __restrict__ r1,
    const float2* __restrict__ gx1, const float2* __restrict__ gr1, const float2* __restrict__ gp1, long long n,
    float inv, float lam, double* __restrict__ part) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 pv = p[i], av = cg_ap(nrm, p, i, inv, lam), rv = r1[i];
    const float2 gx = ld_or0(gx1, i), gr = ld_or0(gr1, i), gp = ld_or0(gp1, i);
    v[0] += (double)gp.x * pv.x + (double)gp.y * pv.y;
    v[1] += ((double)gr.x + gp.x) * av.x + ((double)gr.y + gp.y) * av.y;
    v[2] += (double)rv.x * av.x + (double)rv.y * av.y;
    v[3] += (double)gx.x * pv.x + (double)gx.y * pv.y;
  }
  block_sum4(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 4; ++k) part[k * gridDim.x + blockIdx.x] = v[k];
}

// bs[0] grz1' (= grz1 + gbeta / rz), bs[1] galpha, bs[2] gpAp; grz[0] = gradient of rz
__global__ __launch_bounds__(kRedThreads) void cgd_bwd_scalars_kernel(const double* __restrict__ part, int nparts,
                                                                      const double* __restrict__ sd,
                                                                      const double* __restrict__ grz1,
                                                                      double* __restrict__ bs,
                                                                      double* __restrict__ grz) {
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = threadIdx.x; i < nparts; i += blockDim.x)
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] += part[k * nparts + i];
  block_sum4(v);
  if (threadIdx.x == 0) {
    const double rz = sd[0], pap = sd[1], rz1 = sd[2];
    const double gbeta = v[0];
    const double g1 = (grz1 ? grz1[0] : 0.0) + (rz > 0.0 ? gbeta / rz : 0.0);
    const double galpha = pap > 0.0 ? v[3] - (v[1] + 2.0 * g1 * v[2]) : 0.0;
    bs[0] = g1;
    bs[1] = galpha;
    bs[2] = pap > 0.0 ? -galpha * rz / (pap * pap) : 0.0;
    grz[0] = (rz > 0.0 ? -gbeta * rz1 / (rz * rz) : 0.0) + (pap > 0.0 ? galpha / pap : 0.0);
  }
}

__global__ __launch_bounds__(kRedThreads) void cgd_bwd_vec_kernel(
    const float2* __restrict__ p, const float2* __restrict__ nrm, const float2* __restrict__ r1,
    const float2* __restrict__ gx1, const float2* __restrict__ gr1, const float2* __restrict__ gp1, long long n,
    float inv, float lam, const double* __restrict__ sd, const double* __restrict__ bs, float2* __restrict__ gr,
    float2* __restrict__ gp, float2* __restrict__ gnrm) {
  const float alpha = (float)sd[4], beta = (float)sd[3];
  const float two_g1 = (float)(2.0 * bs[0]), gpap = (float)bs[2];
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 pv = p[i], av = cg_ap(nrm, p, i, inv, lam), rv = r1[i];
    const float2 gx = ld_or0(gx1, i), g_r = ld_or0(gr1, i), g_p = ld_or0(gp1, i);
    const float2 g = make_float2(g_r.x + g_p.x + two_g1 * rv.x, g_r.y + g_p.y + two_g1 * rv.y);
    const float2 gap = make_float2(gpap * pv.x - alpha * g.x, gpap * pv.y - alpha * g.y);
    gr[i] = g;
    gnrm[i] = make_float2(inv * gap.x, inv * gap.y);
    gp[i] = make_float2(beta * g_p.x + alpha * gx.x + gpap * av.x + lam * gap.x,
                        beta * g_p.y + alpha * gx.y + gpap * av.y + lam * gap.y);
  }
}

static unsigned seg_blocks(long long seg_len) {
  long long b = (seg_len + kRedThreads - 1) / kRedThreads;
  if (b > kSegBlocks) b = kSegBlocks;
  if (b < 1) b = 1;
  return (unsigned)b;
}

static unsigned red_blocks(long long n) {
  long long b = (n + kRedThreads - 1) / kRedThreads;
  if (b > kRedBlocks) b = kRedBlocks;
  if (b < 1) b = 1;
  return (unsigned)b;
}

}  // namespace mra

using namespace mra;

extern "C" {

int mra_cg_partials() { return kRedBlocks; }
int mra_cg_seg_blocks() { return kSegBlocks; }

// Segmented: part holds nseg * mra_cg_seg_blocks() doubles, sc nseg * 8.
int mra_cg_init_seg(const void* b, const void* ax, const void* x, void* r, void* p, long long seg_len, int nseg,
                    float lam, double* part, double* sc, void* stream) {
  MRA_TRY_BEGIN
  if (seg_len <= 0 || nseg <= 0 || nseg > 65535) MRA_FAIL("cg_init_seg: bad segment sizes");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = seg_blocks(seg_len);
  seg_cg_init_kernel<<<dim3(nb, nseg), kRedThreads, 0, st>>>((const float2*)b, (const float2*)ax,
                                                               (const float2*)x, (float2*)r, (float2*)p, seg_len,
                                                               lam, part);
  seg_reduce_final_kernel<<<nseg, kRedThreads, 0, st>>>(part, (int)nb, sc, 0, 0);
  MRA_CHECK_LAUNCH("cg_init_seg");
  MRA_TRY_END
}

int mra_cg_step_seg(void* x, void* r, void* p, const void* ap, long long seg_len, int nseg, float lam,
                    double* part, double* sc, void* stream) {
  MRA_TRY_BEGIN
  if (seg_len <= 0 || nseg <= 0 || nseg > 65535) MRA_FAIL("cg_step_seg: bad segment sizes");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = seg_blocks(seg_len);
  const dim3 grd(nb, nseg);
  seg_cdot_partial_kernel<<<grd, kRedThreads, 0, st>>>((const float2*)p, (const float2*)ap, seg_len, lam, part);
  seg_reduce_final_kernel<<<nseg, kRedThreads, 0, st>>>(part, (int)nb, sc, 1, 0);
  seg_cg_xr_kernel<<<grd, kRedThreads, 0, st>>>((float2*)x, (float2*)r, (const float2*)p, (const float2*)ap,
                                                seg_len, lam, sc, part);
  seg_reduce_final_kernel<<<nseg, kRedThreads, 0, st>>>(part, (int)nb, sc, 0, 1);
  seg_cg_p_kernel<<<grd, kRedThreads, 0, st>>>((float2*)p, (const float2*)r, seg_len, sc);
  MRA_CHECK_LAUNCH("cg_step_seg");
  MRA_TRY_END
}

// sc[slot] = Re <a, b + lam a>  (FP64)
int mra_cdot_re(const void* a, const void* b, long long n, float lam, double* part, double* sc, int slot,
                void* stream) {
  MRA_TRY_BEGIN
  if (n <= 0) MRA_FAIL("cdot: n must be positive");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = red_blocks(n);
  cdot_partial_kernel<<<nb, kRedThreads, 0, st>>>((const float2*)a, (const float2*)b, n, lam, part);
  reduce_final_kernel<<<1, kRedThreads, 0, st>>>(part, (int)nb, sc, slot, 0);
  MRA_CHECK_LAUNCH("cdot");
  MRA_TRY_END
}

// r = b - (Ax + lam x); p = r; sc[0] = <r, r>
int mra_cg_init(const void* b, const void* ax, const void* x, void* r, void* p, long long n, float lam,
                double* part, double* sc, void* stream) {
  MRA_TRY_BEGIN
  if (n <= 0) MRA_FAIL("cg_init: n must be positive");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = red_blocks(n);
  cg_init_kernel<<<nb, kRedThreads, 0, st>>>((const float2*)b, (const float2*)ax, (const float2*)x, (float2*)r,
                                            (float2*)p, n, lam, part);
  reduce_final_kernel<<<1, kRedThreads, 0, st>>>(part, (int)nb, sc, 0, 0);
  MRA_CHECK_LAUNCH("cg_init");
  MRA_TRY_END
}

// One CG step after Ap has been formed: pAp, x/r update, rz/beta, p update.
int mra_cg_step(void* x, void* r, void* p, const void* ap, long long n, float lam, double* part, double* sc,
                void* stream) {
  MRA_TRY_BEGIN
  if (n <= 0) MRA_FAIL("cg_step: n must be positive");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = red_blocks(n);
  cdot_partial_kernel<<<nb, kRedThreads, 0, st>>>((const float2*)p, (const float2*)ap, n, lam, part);
  reduce_final_kernel<<<1, kRedThreads, 0, st>>>(part, (int)nb, sc, 1, 0);
  cg_xr_kernel<<<nb, kRedThreads, 0, st>>>((float2*)x, (float2*)r, (const float2*)p, (const float2*)ap, n, lam, sc,
                                          part);
  reduce_final_kernel<<<1, kRedThreads, 0, st>>>(part, (int)nb, sc, 0, 1);
  cg_p_kernel<<<nb, kRedThreads, 0, st>>>((float2*)p, (const float2*)r, n, sc);
  MRA_CHECK_LAUNCH("cg_step");
  MRA_TRY_END
}

// Differentiable CG step, forward (see cgd_* above); part holds mra_cg_partials() doubles.
int mra_cg_step_fwd(const void* x, const void* r, const void* p, const void* nrm, const double* rz, long long n,
                    float inv, float lam, void* x1, void* r1, void* p1, double* rz1, double* sd, double* part,
                    void* stream) {
  MRA_TRY_BEGIN
  if (n <= 0) MRA_FAIL("cg_step_fwd: n must be positive");
  if (!x || !r || !p || !nrm || !rz || !x1 || !r1 || !p1 || !rz1 || !sd || !part)
    MRA_FAIL("cg_step_fwd: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = red_blocks(n);
  cgd_pap_kernel<<<nb, kRedThreads, 0, st>>>((const float2*)p, (const float2*)nrm, n, inv, lam, part);
  cgd_alpha_kernel<<<1, kRedThreads, 0, st>>>(part, (int)nb, rz, sd);
  cgd_xr_kernel<<<nb, kRedThreads, 0, st>>>((const float2*)x, (const float2*)r, (const float2*)p,
                                           (const float2*)nrm, n, inv, lam, sd, (float2*)x1, (float2*)r1, part);
  cgd_beta_kernel<<<1, kRedThreads, 0, st>>>(part, (int)nb, sd, rz1);
  cgd_p_kernel<<<nb, kRedThreads, 0, st>>>((const float2*)r1, (const float2*)p, n, sd, (float2*)p1);
  MRA_CHECK_LAUNCH("cg_step_fwd");
  MRA_TRY_END
}

// Its adjoint; gx1, gr1, gp1, grz1 may be null (zero cotangents); part holds
// 4 * mra_cg_partials() doubles, bs 8.
int mra_cg_step_bwd(const void* p, const void* nrm, const void* r1, const double* sd, const void* gx1,
                    const void* gr1, const void* gp1, const double* grz1, long long n, float inv, float lam,
                    void* gr, void* gp, void* gnrm, double* grz, double* part, double* bs, void* stream) {
  MRA_TRY_BEGIN
  if (n <= 0) MRA_FAIL("cg_step_bwd: n must be positive");
  if (!p || !nrm || !r1 || !sd || !gr || !gp || !gnrm || !grz || !part || !bs) MRA_FAIL("cg_step_bwd: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = red_blocks(n);
  cgd_bwd_dots_kernel<<<nb, kRedThreads, 0, st>>>((const float2*)p, (const float2*)nrm, (const float2*)r1,
                                                 (const float2*)gx1, (const float2*)gr1, (const float2*)gp1, n,
                                                 inv, lam, part);
  cgd_bwd_scalars_kernel<<<1, kRedThreads, 0, st>>>(part, (int)nb, sd, grz1, bs, grz);
  cgd_bwd_vec_kernel<<<nb, kRedThreads, 0, st>>>((const float2*)p, (const float2*)nrm, (const float2*)r1,
                                                (const float2*)gx1, (const float2*)gr1, (const float2*)gp1, n,
                                                inv, lam, sd, bs, (float2*)gr, (float2*)gp, (float2*)gnrm);
  MRA_CHECK_LAUNCH("cg_step_bwd");
  MRA_TRY_END
}

}  // extern "C"
