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

// Segmented form of the scatter: the samples sorted by grid index (perm: sorted ->
// sample) and cut into chunks of at most one run's samples (a run = all samples of one
// grid point), chunk c = sorted positions cstart[c] .. cend[c]-1 at point cidx[c]. One
// thread per (chunk, batch entry) sums its chunk in sorted order and adds the sum with
// one atomic: a point hit by many samples (the centre of a radial trajectory) takes a
// few chunk sums instead of one atomic per sample.
__global__ __launch_bounds__(kThreads) void scatter_chunks_kernel(const float2* __restrict__ ksp,
                                                                  const int* __restrict__ perm,
                                                                  const int* __restrict__ cstart,
                                                                  const int* __restrict__ cend,
                                                                  const int* __restrict__ cidx, long long NC,
                                                                  long long M, int B, float scale,
                                                                  float2* __restrict__ grid, long long gstride) {
  const long long n = NC * B;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const long long c = t / B;
    const int b = (int)(t - c * B);
    const float2* kb = ksp + b * M;
    float sx = 0.f, sy = 0.f;
    const int j1 = cend[c];
#pragma unroll 4
    for (int j = cstart[c]; j < j1; ++j) {
      const float2 v = kb[perm[j]];
      sx += v.x;
      sy += v.y;
    }
    float* g = reinterpret_cast<float*>(grid + b * gstride + cidx[c]);
    atomicAdd(g, sx * scale);
    atomicAdd(g + 1, sy * scale);
  }
}

// Low-rank nearest-grid forms (Chebyshev NUFFT, fourier.py:1574-1774): L spectra per batch
// entry, each sample reads its point in all L and sums them with per-sample weights h:
//   out[b, m] = scale * sum_l h[l, m] * grid[b * L + l, idx[m]]
__global__ __launch_bounds__(kThreads) void gather_sum_kernel(const float2* __restrict__ grid, long long gstride,
                                                              const int* __restrict__ idx,
                                                              const float2* __restrict__ h, int L, long long M,
                                                              int B, float scale, float2* __restrict__ out) {
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (long long)gridDim.x * blockDim.x) {
    const long long i = idx[m];
    for (int b = 0; b < B; ++b) {
      const float2* gb = grid + (long long)b * L * gstride + i;
      float sx = 0.f, sy = 0.f;
      for (int l = 0; l < L; ++l) {
        const float2 w = h[(long long)l * M + m];
        const float2 v = gb[(long long)l * gstride];
        sx += w.x * v.x - w.y * v.y;
        sy += w.x * v.y + w.y * v.x;
      }
      out[(long long)b * M + m] = make_float2(sx * scale, sy * scale);
    }
  }
}

// Adjoint, segmented like scatter_chunks_kernel: one thread per (chunk, l, batch entry) sums
// conj(h[l, s]) * ksp[b, s] over its chunk in sorted order and adds the sum into
// grid[b * L + l, cidx[c]] with one atomic.
__global__ __launch_bounds__(kThreads) void scatter_sum_chunks_kernel(
    const float2* __restrict__ ksp, const float2* __restrict__ h, int L, const int* __restrict__ perm,
    const int* __restrict__ cstart, const int* __restrict__ cend, const int* __restrict__ cidx, long long NC,
    long long M, int B, float scale, float2* __restrict__ grid, long long gstride) {
  const long long n = NC * L * B;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const long long c = t / ((long long)L * B);
    const int r = (int)(t - c * L * B);
    const int l = r / B, b = r - l * B;
    const float2* kb = ksp + (long long)b * M;
    const float2* hl = h + (long long)l * M;
    float sx = 0.f, sy = 0.f;
    const int j1 = cend[c];
    for (int j = cstart[c]; j < j1; ++j) {
      const int s = perm[j];
      const float2 v = kb[s];
      const float2 w = hl[s];
      sx += w.x * v.x + w.y * v.y;  // conj(w) * v
      sy += w.x * v.y - w.y * v.x;
    }
    float* g = reinterpret_cast<float*>(grid + ((long long)b * L + l) * gstride + cidx[c]);
    atomicAdd(g, sx * scale);
    atomicAdd(g + 1, sy * scale);
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

// Adds scale * (sum of chunk c's samples) into grid[b, cidx[c]] (the caller zeroes the
// grid). perm / cstart / cend / cidx come from sorting the samples' flat indices and
// cutting each run of equal indices into chunks.
int mra_nn_scatter_chunks(const void* ksp, const int* perm, const int* cstart, const int* cend, const int* cidx,
                          long long NC, long long M, int B, float scale, void* grid, long long gstride,
                          void* stream) {
  MRA_TRY_BEGIN
  if (NC <= 0 || M <= 0 || B <= 0 || gstride <= 0) MRA_FAIL("nn_scatter_chunks: sizes must be positive");
  if (!ksp || !perm || !cstart || !cend || !cidx || !grid) MRA_FAIL("nn_scatter_chunks: null pointer");
  nn::scatter_chunks_kernel<<<nn::blocks_for(NC * B), nn::kThreads, 0, (hipStream_t)stream>>>(
      (const float2*)ksp, perm, cstart, cend, cidx, NC, M, B, scale, (float2*)grid, gstride);
  MRA_CHECK_LAUNCH("nn_scatter_chunks");
  MRA_TRY_END
}

// out[b, m] = scale * sum_l h[l, m] * grid[b * L + l, idx[m]]; grid holds B * L grids of
// gstride points, h is (L, M) complex64, idx M flat indices in [0, gstride).
int mra_nn_gather_sum(const void* grid, long long gstride, const int* idx, const void* h, int L, long long M, int B,
                      float scale, void* out, void* stream) {
  MRA_TRY_BEGIN
  if (M <= 0 || B <= 0 || L <= 0 || gstride <= 0) MRA_FAIL("nn_gather_sum: sizes must be positive");
  if (!grid || !idx || !h || !out) MRA_FAIL("nn_gather_sum: null pointer");
  nn::gather_sum_kernel<<<nn::blocks_for(M), nn::kThreads, 0, (hipStream_t)stream>>>(
      (const float2*)grid, gstride, idx, (const float2*)h, L, M, B, scale, (float2*)out);
  MRA_CHECK_LAUNCH("nn_gather_sum");
  MRA_TRY_END
}

// Adjoint of mra_nn_gather_sum over the chunk tables of mra_nn_scatter_chunks: adds
// scale * sum_{s in chunk c} conj(h[l, s]) ksp[b, s] into grid[b * L + l, cidx[c]] (the caller
// zeroes the B * L grids).
int mra_nn_scatter_sum_chunks(const void* ksp, const void* h, int L, const int* perm, const int* cstart,
                              const int* cend, const int* cidx, long long NC, long long M, int B, float scale,
                              void* grid, long long gstride, void* stream) {
  MRA_TRY_BEGIN
  if (NC <= 0 || M <= 0 || B <= 0 || L <= 0 || gstride <= 0) MRA_FAIL("nn_scatter_sum_chunks: sizes must be positive");
  if (!ksp || !h || !perm || !cstart || !cend || !cidx || !grid) MRA_FAIL("nn_scatter_sum_chunks: null pointer");
  nn::scatter_sum_chunks_kernel<<<nn::blocks_for(NC * L * B), nn::kThreads, 0, (hipStream_t)stream>>>(
      (const float2*)ksp, (const float2*)h, L, perm, cstart, cend, cidx, NC, M, B, scale, (float2*)grid, gstride);
  MRA_CHECK_LAUNCH("nn_scatter_sum_chunks");
  MRA_TRY_END
}

}  // extern "C"
