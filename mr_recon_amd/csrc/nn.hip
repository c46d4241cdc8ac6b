// Nearest-grid sampling: the gather and scatter of the gridded NUFFT.
//
// The reference's gridded_nufft (fourier.py:1213-1326) rounds each sample to the
// nearest point of the oversampled grid; forward_interp_only reads that point
// (indexing.multi_index, indexing.py:29-51) and adjoint_grid_only adds the samples
// into their points (indexing.multi_grid, index_add_, indexing.py:53-79). Here each
// sample carries one flat grid index; the scatter adds with float atomics, like the
// reference's index_add_ on the GPU (summation order is not fixed).
#include "common.h"
#include "capi_util.h"

namespace mra {
namespace nn {

constexpr int kThreads = 256;

static unsigned blocks_for(long long n) {
  long long b = (n + kThreads - 1) / kThreads;
  if (b > 8192) b = 8192;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// out[b, m] = scale * grid[b, idx[m]]
__global__ __launch_bounds__(kThreads) void gather_kernel(const float2* __restrict__ grid, long long gstride,
                                                          const int* __restrict__ idx, long long M, int B,
                                                          float scale, float2* __restrict__ out) {
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (long long)gridDim.x * blockDim.x) {
    const long long i = idx[m];
    for (int b = 0; b < B; ++b) {
      const float2 v = grid[b * gstride + i];
      out[b * M + m] = make_float2(v.x * scale, v.y * scale);
    }
  }
}

// grid[b, idx[m]] += scale * ksp[b, m]
__global__ __launch_bounds__(kThreads) void scatter_kernel(const float2* __restrict__ ksp, const int* __restrict__ idx,
                                                           long long M, int B, float scale, float2* __restrict__ grid,
                                                           long long gstride) {
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (long long)gridDim.x * blockDim.x) {
    const long long i = idx[m];
    for (int b = 0; b < B; ++b) {
      const float2 v = ksp[b * M + m];
      float* g = reinterpret_cast<float*>(grid + b * gstride + i);
      atomicAdd(g, v.x * scale);
      atomicAdd(g + 1, v.y * scale);
    }
  }
}

}  // namespace nn
}  // namespace mra

using namespace mra;

extern "C" {

// idx: M flat indices in [0, gstride); the caller checks the range.
int mra_nn_gather(const void* grid, long long gstride, const int* idx, long long M, int B, float scale, void* out,
                  void* stream) {
  MRA_TRY_BEGIN
  if (M <= 0 || B <= 0 || gstride <= 0) MRA_FAIL("nn_gather: sizes must be positive");
  if (!grid || !idx || !out) MRA_FAIL("nn_gather: null pointer");
  nn::gather_kernel<<<nn::blocks_for(M), nn::kThreads, 0, (hipStream_t)stream>>>(
      (const float2*)grid, gstride, idx, M, B, scale, (float2*)out);
  MRA_CHECK_LAUNCH("nn_gather");
  MRA_TRY_END
}

// Adds into grid (B grids of gstride points): the caller zeroes it first.
int mra_nn_scatter(const void* ksp, const int* idx, long long M, int B, float scale, void* grid, long long gstride,
                   void* stream) {
  MRA_TRY_BEGIN
  if (M <= 0 || B <= 0 || gstride <= 0) MRA_FAIL("nn_scatter: sizes must be positive");
  if (!grid || !idx || !ksp) MRA_FAIL("nn_scatter: null pointer");
  nn::scatter_kernel<<<nn::blocks_for(M), nn::kThreads, 0, (hipStream_t)stream>>>(
      (const float2*)ksp, idx, M, B, scale, (float2*)grid, gstride);
  MRA_CHECK_LAUNCH("nn_scatter");
  MRA_TRY_END
}

}  // extern "C"
