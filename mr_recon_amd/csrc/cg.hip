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

}  // extern "C"
