// Kaiser-Bessel interpolation and gridding in 4 and 5 dimensions.
//
// The reference interpolates 4D / 5D grids with the Triton kernels _ungrid4d / _ungrid5d
// (triton_interp/ungrid.py:498-845) and grids them with its pure-torch fallback
// (grid_torch, grid.py:86-127 and :556-624: index_put_ with accumulate). Both cover the
// (floor(w)+1)^D patch of grid points g = ceil(t - w/2) + j around each sample with
// weight prod_d KB(|g_d - t_d| / (w/2)) where |g_d - t_d| <= w/2, indices wrapping mod G.
// At these dimensions the patch is large (5^5 = 3125 points for w = 4) and the grids are
// small, so one thread takes one sample: it evaluates its D x P weights once, then walks
// the patch with an odometer over the D indices. Interpolation reads the grid directly
// (L2-resident at these sizes) for up to kB batch entries per pass; gridding adds with
// float atomics, as the reference's index_put_ does on the GPU (summation order not fixed).
#include "common.h"
#include "capi_util.h"

namespace mra {
namespace direct {

constexpr int kThreads = 256;
constexpr int kB = 4;  // batch entries accumulated per pass over the patch

struct DGeom {
  int G[5];
  float beta[5];
  float r;          // half width
  float out_scale;  // multiplies every weight
};

template <int D, int P>
__device__ __forceinline__ bool sample_weights(const float* __restrict__ t, long long m, const DGeom& g,
                                               int (&lower)[D], float (&w)[D][P]) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const float td = t[m * D + d];
    if (!isfinite(td)) return false;
    lower[d] = kb_weights<P>(td, g.r, g.beta[d], d == 0 ? g.out_scale : 1.0f, w[d]);
  }
  return true;
}

template <int D, int P>
__global__ __launch_bounds__(kThreads) void interp_kernel(const float2* __restrict__ grid, long long gstride,
                                                          const float* __restrict__ t, long long M, int B,
                                                          DGeom g, float2* __restrict__ out) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  int lower[D];
  float w[D][P];
  const bool ok = sample_weights<D, P>(t, m, g, lower, w);
  for (int b0 = 0; b0 < B; b0 += kB) {
    float2 acc[kB];
#pragma unroll
    for (int u = 0; u < kB; ++u) acc[u] = make_float2(0.f, 0.f);
    if (ok) {
      int j[D];
#pragma unroll
      for (int d = 0; d < D; ++d) j[d] = 0;
      for (int n = 0; n < ipow(P, D); ++n) {
        long long idx = 0;
        float wt = 1.0f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          idx = idx * g.G[d] + pmod(lower[d] + j[d], g.G[d]);
          wt *= w[d][j[d]];
        }
        if (wt != 0.0f) {
#pragma unroll
          for (int u = 0; u < kB; ++u)
            if (b0 + u < B) {
              const float2 v = grid[(long long)(b0 + u) * gstride + idx];
              acc[u].x += wt * v.x;
              acc[u].y += wt * v.y;
            }
        }
#pragma unroll
        for (int d = D - 1; d >= 0; --d) {  // odometer, last dim fastest
          if (++j[d] < P) break;
          j[d] = 0;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kB; ++u)
      if (b0 + u < B) out[(long long)(b0 + u) * M + m] = acc[u];
  }
}

template <int D, int P>
__global__ __launch_bounds__(kThreads) void grid_kernel(const float2* __restrict__ ksp, long long M, int B,
                                                        const float* __restrict__ t, DGeom g,
                                                        float2* __restrict__ grid, long long gstride) {
  const long long m = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  int lower[D];
  float w[D][P];
  if (!sample_weights<D, P>(t, m, g, lower, w)) return;
  int j[D];
#pragma unroll
  for (int d = 0; d < D; ++d) j[d] = 0;
  for (int n = 0; n < ipow(P, D); ++n) {
    long long idx = 0;
    float wt = 1.0f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      idx = idx * g.G[d] + pmod(lower[d] + j[d], g.G[d]);
      wt *= w[d][j[d]];
    }
    if (wt != 0.0f) {
      for (int b = 0; b < B; ++b) {
        const float2 v = ksp[(long long)b * M + m];
        float* dst = reinterpret_cast<float*>(grid + (long long)b * gstride + idx);
        atomicAdd(dst, wt * v.x);
        atomicAdd(dst + 1, wt * v.y);
      }
    }
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      if (++j[d] < P) break;
      j[d] = 0;
    }
  }
}

template <int D, int P>
static void launch(bool adjoint, const void* in, void* outp, const float* t, long long M, int B, long long gstride,
                   const DGeom& g, hipStream_t st) {
  const unsigned nb = (unsigned)((M + kThreads - 1) / kThreads);
  if (adjoint)
    grid_kernel<D, P><<<nb, kThreads, 0, st>>>((const float2*)in, M, B, t, g, (float2*)outp, gstride);
  else
    interp_kernel<D, P><<<nb, kThreads, 0, st>>>((const float2*)in, gstride, t, M, B, g, (float2*)outp);
}

template <int D>
static int launch_p(int P, bool adjoint, const void* in, void* outp, const float* t, long long M, int B,
                    long long gstride, const DGeom& g, hipStream_t st) {
  switch (P) {
    case 2: launch<D, 2>(adjoint, in, outp, t, M, B, gstride, g, st); return 0;
    case 3: launch<D, 3>(adjoint, in, outp, t, M, B, gstride, g, st); return 0;
    case 4: launch<D, 4>(adjoint, in, outp, t, M, B, gstride, g, st); return 0;
    case 5: launch<D, 5>(adjoint, in, outp, t, M, B, gstride, g, st); return 0;
    case 6: launch<D, 6>(adjoint, in, outp, t, M, B, gstride, g, st); return 0;
    case 7: launch<D, 7>(adjoint, in, outp, t, M, B, gstride, g, st); return 0;
    default: return 1;
  }
}

static int run(int ndim, int P, bool adjoint, const void* in, void* outp, const float* t, long long M, int B,
               const int* G, const float* betas, float width, float out_scale, void* stream) {
  if (ndim != 4 && ndim != 5) MRA_FAIL("direct interp/grid: ndim must be 4 or 5");
  if (P < 2 || P > 7) MRA_FAIL("direct interp/grid: floor(width)+1 must be in 2..7");
  if (M < 0 || B < 1) MRA_FAIL("direct interp/grid: bad sizes");
  if (!in || !outp || !t || !G || !betas) MRA_FAIL("direct interp/grid: null pointer");
  DGeom g{};
  long long gstride = 1;
  for (int d = 0; d < ndim; ++d) {
    if (G[d] < 1) MRA_FAIL("direct interp/grid: grid sizes must be positive");
    g.G[d] = G[d];
    g.beta[d] = betas[d];
    gstride *= G[d];
  }
  g.r = 0.5f * width;
  g.out_scale = out_scale;
  if (M == 0) return 0;
  const int rc = ndim == 4 ? launch_p<4>(P, adjoint, in, outp, t, M, B, gstride, g, (hipStream_t)stream)
                           : launch_p<5>(P, adjoint, in, outp, t, M, B, gstride, g, (hipStream_t)stream);
  if (rc) MRA_FAIL("direct interp/grid: unsupported kernel width");
  MRA_CHECK_LAUNCH("direct interp/grid");
  return 0;
}

}  // namespace direct
}  // namespace mra

using namespace mra;

extern "C" {

int mra_interp_direct(int ndim, int P, const void* grid, const float* t, long long M, int B, const int* G,
                      const float* betas, float width, float out_scale, void* out, void* stream) {
  MRA_TRY_BEGIN
  return direct::run(ndim, P, false, grid, out, t, M, B, G, betas, width, out_scale, stream);
  MRA_TRY_END
}

int mra_grid_direct(int ndim, int P, const void* ksp, const float* t, long long M, int B, const int* G,
                    const float* betas, float width, float out_scale, void* grid, void* stream) {
  MRA_TRY_BEGIN
  return direct::run(ndim, P, true, ksp, grid, t, M, B, G, betas, width, out_scale, stream);
  MRA_TRY_END
}

}  // extern "C"
