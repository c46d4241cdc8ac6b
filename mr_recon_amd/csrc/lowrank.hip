// Low-rank coil / imperfection mixing of imperf_coil_lowrank (linops.py:525-759): the
// k-space of coil c is sum_l h[l, c] * (F B_l x), so after one NUFFT of the L images the
// forward contracts the L transforms with h, and the adjoint forms the L adjoint inputs
// sum_c conj(h[l, c]) * dcf * y_c. Both are one pass over h (L x C x M complex64, the largest
// operand), one thread per sample with coalesced loads along m.
#include "common.h"
#include "capi_util.h"

namespace mra {
namespace lowrank {

constexpr int kThreads = 256;

static unsigned blocks_for(long long n) {
  long long b = (n + kThreads - 1) / kThreads;
  if (b > 16384) b = 16384;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// out[c, m] = sum_l h[l, c, m] * y[l, m]
__global__ __launch_bounds__(kThreads) void expand_kernel(const float2* __restrict__ y, const float2* __restrict__ h,
                                                          int L, int C, long long M, float2* __restrict__ out) {
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (long long)gridDim.x * blockDim.x) {
    for (int c = 0; c < C; ++c) {
      float sx = 0.f, sy = 0.f;
      for (int l = 0; l < L; ++l) {
        const float2 w = h[((long long)l * C + c) * M + m];
        const float2 v = y[(long long)l * M + m];
        sx += w.x * v.x - w.y * v.y;
        sy += w.x * v.y + w.y * v.x;
      }
      out[(long long)c * M + m] = make_float2(sx, sy);
    }
  }
}

// out[l, m] = dcf[m] * sum_c conj(h[l, c, m]) * ksp[c, m]   (dcf may be null: 1)
__global__ __launch_bounds__(kThreads) void reduce_kernel(const float2* __restrict__ ksp, const float2* __restrict__ h,
                                                          const float* __restrict__ dcf, int L, int C, long long M,
                                                          float2* __restrict__ out) {
  for (long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x; m < M; m += (long long)gridDim.x * blockDim.x) {
    const float d = dcf ? dcf[m] : 1.0f;
    for (int l = 0; l < L; ++l) {
      float sx = 0.f, sy = 0.f;
      for (int c = 0; c < C; ++c) {
        const float2 w = h[((long long)l * C + c) * M + m];
        const float2 v = ksp[(long long)c * M + m];
        sx += w.x * v.x + w.y * v.y;
        sy += w.x * v.y - w.y * v.x;
      }
      out[(long long)l * M + m] = make_float2(sx * d, sy * d);
    }
  }
}

}  // namespace lowrank
}  // namespace mra

using namespace mra;

extern "C" {

int mra_lowrank_expand(const void* y, const void* h, int L, int C, long long M, void* out, void* stream) {
  MRA_TRY_BEGIN
  if (L <= 0 || C <= 0 || M <= 0) MRA_FAIL("lowrank_expand: sizes must be positive");
  if (!y || !h || !out) MRA_FAIL("lowrank_expand: null pointer");
  lowrank::expand_kernel<<<lowrank::blocks_for(M), lowrank::kThreads, 0, (hipStream_t)stream>>>(
      (const float2*)y, (const float2*)h, L, C, M, (float2*)out);
  MRA_CHECK_LAUNCH("lowrank_expand");
  MRA_TRY_END
}

int mra_lowrank_reduce(const void* ksp, const void* h, const float* dcf, int L, int C, long long M, void* out,
                       void* stream) {
  MRA_TRY_BEGIN
  if (L <= 0 || C <= 0 || M <= 0) MRA_FAIL("lowrank_reduce: sizes must be positive");
  if (!ksp || !h || !out) MRA_FAIL("lowrank_reduce: null pointer");
  lowrank::reduce_kernel<<<lowrank::blocks_for(M), lowrank::kThreads, 0, (hipStream_t)stream>>>(
      (const float2*)ksp, (const float2*)h, dcf, L, C, M, (float2*)out);
  MRA_CHECK_LAUNCH("lowrank_reduce");
  MRA_TRY_END
}

}  // extern "C"
