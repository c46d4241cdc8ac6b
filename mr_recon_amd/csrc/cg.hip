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
