// Pipe-Menon density compensation: the elementwise step of one iteration and the
// final normalisation.
//
// The reference (algs.py:20-106) calls sigpy.mri.pipe_menon_dcf, which starts at
// w = 1 and repeats w /= |G^H G w|, G the Kaiser-Bessel gridding of the samples
// onto the image grid; the result is divided by its maximum. The gridding and
// interpolation run in the NUFFT kernels; these kernels do the division by
// |G^H G w| and the max-normalisation, both in fp32 like the reference.
#include <math.h>

#include "common.h"
#include "capi_util.h"

namespace mra {
namespace dcf {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 1024;

static unsigned blocks_for(long long n) {
  long long b = (n + kThreads - 1) / kThreads;
  if (b > kMaxBlocks) b = kMaxBlocks;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// w[i] /= |ghg[i]|, ghg = G^H G w as complex64 (its imaginary part is exactly zero
// when the gridded values are real, so |ghg| = |Re ghg|; hypotf keeps |.| exact).
__global__ __launch_bounds__(kThreads) void update_kernel(const float2* __restrict__ ghg, float* __restrict__ w,
                                                          long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const float2 g = ghg[i];
    w[i] = w[i] / hypotf(g.x, g.y);
  }
}

__device__ __forceinline__ float wave_max(float m) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_down(m, o));
  return m;
}

// part[blk] = max of the block's grid-strided elements
__global__ __launch_bounds__(kThreads) void max_partial_kernel(const float* __restrict__ w, long long n,
                                                               float* __restrict__ part) {
  __shared__ float sh[kThreads / 64];
  float m = -INFINITY;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    m = fmaxf(m, w[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = sh[0];
    for (int k = 1; k < kThreads / 64; ++k) t = fmaxf(t, sh[k]);
    part[blockIdx.x] = t;
  }
}

// out[perm ? perm[i] : i] = w[i] / max(part[0..nb)); each block folds the nb partials
// itself (nb <= 1024)
// (w and out may alias when perm is null: each element is read and written by one thread)
__global__ __launch_bounds__(kThreads) void scale_kernel(const float* w, const int* __restrict__ perm, float* out,
                                                         long long n,
                                                         const float* __restrict__ part, int nb) {
  __shared__ float mx;
  if (threadIdx.x < 64) {
    float m = -INFINITY;
    for (int j = threadIdx.x; j < nb; j += 64) m = fmaxf(m, part[j]);
    m = wave_max(m);
    if (threadIdx.x == 0) mx = m;
  }
  __syncthreads();
  const float d = mx;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[perm ? perm[i] : i] = w[i] / d;
}

}  // namespace dcf
}  // namespace mra

using namespace mra;

extern "C" {

int mra_dcf_partials() { return dcf::kMaxBlocks; }

int mra_dcf_update(const void* ghg, float* w, long long n, void* stream) {
  MRA_TRY_BEGIN
  if (n <= 0) MRA_FAIL("dcf_update: n must be positive");
  if (!ghg || !w) MRA_FAIL("dcf_update: null pointer");
  dcf::update_kernel<<<dcf::blocks_for(n), dcf::kThreads, 0, (hipStream_t)stream>>>((const float2*)ghg, w, n);
  MRA_CHECK_LAUNCH("dcf_update");
  MRA_TRY_END
}

// out[perm ? perm[i] : i] = w[i] / max(w); perm (a permutation of 0..n-1) may be null,
// out may equal w only when perm is null. part: mra_dcf_partials() floats of workspace.
int mra_dcf_normalize(const float* w, const int* perm, float* out, long long n, float* part, void* stream) {
  MRA_TRY_BEGIN
  if (n <= 0) MRA_FAIL("dcf_normalize: n must be positive");
  if (!w || !out || !part) MRA_FAIL("dcf_normalize: null pointer");
  if (perm && out == w) MRA_FAIL("dcf_normalize: a permuted result cannot overwrite its input");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = dcf::blocks_for(n);
  dcf::max_partial_kernel<<<nb, dcf::kThreads, 0, st>>>(w, n, part);
  dcf::scale_kernel<<<nb, dcf::kThreads, 0, st>>>(w, perm, out, n, part, (int)nb);
  MRA_CHECK_LAUNCH("dcf_normalize");
  MRA_TRY_END
}

}  // extern "C"
