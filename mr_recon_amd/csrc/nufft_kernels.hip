// Kaiser-Bessel interpolation (gather), gridding (scatter) and the trajectory
// gradient of interpolation, on tile-sorted samples.
//
// Work decomposition (all three kernels): one 256-thread block per work item
// = one spatial tile's slice of at most 256*K tile-sorted samples. Each thread
// owns K samples, computes their separable KB weights once and keeps them in
// registers for all batch entries (coils). The tile's grid neighbourhood
// (T + P - 1 points per dim, wrapped mod G) is staged in LDS one batch entry
// at a time; every patch read is an LDS read at a compile-time offset from the
// sample's base.
//
//  interp:  LDS double-buffered; the next coil's subgrid is fetched into
//           registers while the current one is consumed.
//  grid:    samples accumulate into an LDS subgrid with LDS float atomics; the
//           subgrid is then flushed to HBM with one global float atomic per
//           non-zero (re, im) component and cleared in the same pass.
//  grad:    like interp, with derivative weights, reduced over coils.
//
// Reference semantics (ungrid.py:356-495, grid.py:361-502, kernels.py:180-393):
// out[b, p] = sum_{g in nbhd(t_p)} prod_d KB(|g_d - t_d| / r) * grid[b, g mod G]
// with KB(u) = I0(beta sqrt(1 - u^2)), |g_d - t_d| <= r, r = w/2.
#include "common.h"
#include "capi_util.h"

namespace mra {

template <int D, int P, int T>
struct TileShape {
  static constexpr int SUB = T + P - 1;            // subgrid edge
  static constexpr int SUBN = ipow(SUB, D);         // subgrid points
  static constexpr int SUBD1 = ipow(SUB, D - 1);   // stride of dim 0 in the subgrid
  static constexpr int PR = ipow(P, D - 1);        // patch points in dims 1..D-1
  static constexpr int NQ = (SUBN + kThreads - 1) / kThreads;
};

// Offset in the subgrid of patch point q of the trailing D-1 dims.
template <int D, int P, int T>
__device__ __forceinline__ constexpr int rest_off(int q) {
  constexpr int SUB = T + P - 1;
  if constexpr (D == 3) return (q / P) * SUB + (q % P);
  else if constexpr (D == 2) return q;
  else return 0;
}

template <int D, int T>
__device__ __forceinline__ void tile_origin(int tile, const Geom& g, int (&org)[D]) {
  int rem = tile;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    org[d] = (rem % g.ntile[d]) * T;
    rem /= g.ntile[d];
  }
}

// Global (wrapped) linear offsets of the subgrid points this thread moves.
template <int D, int P, int T>
__device__ __forceinline__ void subgrid_offsets(const Geom& g, const int (&org)[D],
                                                int (&goff)[TileShape<D, P, T>::NQ]) {
  using TS = TileShape<D, P, T>;
#pragma unroll
  for (int q = 0; q < TS::NQ; ++q) {
    int e = threadIdx.x + q * kThreads;
    if (e < TS::SUBN) {
      int c[D];
#pragma unroll
      for (int d = D - 1; d >= 0; --d) {
        c[d] = e % TS::SUB;
        e /= TS::SUB;
      }
      int lin = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) lin = lin * g.G[d] + (org[d] + c[d]) % g.G[d];
      goff[q] = lin;
    } else {
      goff[q] = -1;
    }
  }
}

// Weights and LDS base of one sample. Returns the LDS base offset.
template <int D, int P, int T>
__device__ __forceinline__ int sample_weights(const float* __restrict__ tsort, long long s, const Geom& g,
                                              const int (&org)[D], float (&w)[D][P]) {
  constexpr int SUB = T + P - 1;
  int base = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const float t = tsort[s * D + d];
    const int lower = kb_weights<P>(t, g.r, g.beta, d == 0 ? g.out_scale : 1.0f, w[d]);
    const int lw = pmod(lower + g.wrap[d], g.G[d]) - org[d];
    base = base * SUB + lw;
  }
  return base;
}

template <int D, int P>
__device__ __forceinline__ void rest_products(const float (&w)[D][P], float (&wr)[ipow(P, D - 1)]) {
  if constexpr (D == 3) {
#pragma unroll
    for (int j = 0; j < P; ++j)
#pragma unroll
      for (int l = 0; l < P; ++l) wr[j * P + l] = w[1][j] * w[2][l];
  } else if constexpr (D == 2) {
#pragma unroll
    for (int l = 0; l < P; ++l) wr[l] = w[1][l];
  } else {
    wr[0] = 1.0f;
  }
}

// --------------------------------------------------------------------------------------------
template <int D, int P, int T, int K>
__global__ __launch_bounds__(kThreads) void interp_kernel(const float2* __restrict__ grid,
                                                          float2* __restrict__ out,
                                                          const float* __restrict__ tsort,
                                                          const int* __restrict__ perm,
                                                          const Geom* __restrict__ geoms,
                                                          const Item* __restrict__ items, int nitems) {
  using TS = TileShape<D, P, T>;
  extern __shared__ float2 lds2[];
  const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
  if (bid >= nitems) return;
  const Item it = items[bid];
  if (it.w <= 0) return;
  const Geom g = geoms[it.x];
  const int tid = threadIdx.x;
  float2* odst = out + g.ksp_off;

  if (it.y < 0) {  // samples with non-finite coordinates: defined as 0
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int s = it.z + tid + k * kThreads;
      if (s < it.z + it.w) {
        const int o = perm[g.samp_off + s];
        for (int b = 0; b < g.nbatch; ++b) odst[(long long)b * g.ksp_stride + o] = make_float2(0.f, 0.f);
      }
    }
    return;
  }

  int org[D];
  tile_origin<D, T>(it.y, g, org);
  int goff[TS::NQ];
  subgrid_offsets<D, P, T>(g, org, goff);

  float wx[K][P];
  float wr[K][TS::PR];
  int base[K], oidx[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = it.z + tid + k * kThreads;
    if (s < it.z + it.w) {
      float w[D][P];
      base[k] = sample_weights<D, P, T>(tsort, g.samp_off + s, g, org, w);
#pragma unroll
      for (int i = 0; i < P; ++i) wx[k][i] = w[0][i];
      rest_products<D, P>(w, wr[k]);
      oidx[k] = perm[g.samp_off + s];
    } else {
#pragma unroll
      for (int i = 0; i < P; ++i) wx[k][i] = 0.f;
#pragma unroll
      for (int q = 0; q < TS::PR; ++q) wr[k][q] = 0.f;
      base[k] = 0;
      oidx[k] = -1;
    }
  }

  const float2* gsrc = grid + g.grid_off;
  float2 pre[TS::NQ];
#pragma unroll
  for (int q = 0; q < TS::NQ; ++q)
    if (goff[q] >= 0) lds2[tid + q * kThreads] = gsrc[goff[q]];
  __syncthreads();

  for (int b = 0; b < g.nbatch; ++b) {
    const int cur = (b & 1) * TS::SUBN;
    const bool more = b + 1 < g.nbatch;
    if (more) {
      const float2* src = gsrc + (long long)(b + 1) * g.grid_stride;
#pragma unroll
      for (int q = 0; q < TS::NQ; ++q)
        if (goff[q] >= 0) pre[q] = src[goff[q]];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const float2* sb = lds2 + cur + base[k];
      float ax = 0.f, ay = 0.f;
#pragma unroll
      for (int i = 0; i < P; ++i) {
        float bx = 0.f, by = 0.f;
#pragma unroll
        for (int q = 0; q < TS::PR; ++q) {
          const float2 v = sb[i * TS::SUBD1 + rest_off<D, P, T>(q)];
          bx += wr[k][q] * v.x;
          by += wr[k][q] * v.y;
        }
        ax += wx[k][i] * bx;
        ay += wx[k][i] * by;
      }
      if (oidx[k] >= 0) odst[(long long)b * g.ksp_stride + oidx[k]] = make_float2(ax, ay);
    }
    if (more) {
      const int nxt = TS::SUBN - cur;
#pragma unroll
      for (int q = 0; q < TS::NQ; ++q)
        if (goff[q] >= 0) lds2[nxt + tid + q * kThreads] = pre[q];
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------------------------
template <int D, int P, int T, int K>
__global__ __launch_bounds__(kThreads) void grid_kernel(float2* __restrict__ grid, const float2* __restrict__ ksp,
                                                        const float* __restrict__ dcf,
                                                        const float* __restrict__ tsort,
                                                        const int* __restrict__ perm,
                                                        const Geom* __restrict__ geoms,
                                                        const Item* __restrict__ items, int nitems) {
  using TS = TileShape<D, P, T>;
  extern __shared__ float ldsf[];
  const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
  if (bid >= nitems) return;
  const Item it = items[bid];
  if (it.w <= 0 || it.y < 0) return;  // non-finite samples contribute nothing
  const Geom g = geoms[it.x];
  const int tid = threadIdx.x;

  int org[D];
  tile_origin<D, T>(it.y, g, org);
  int goff[TS::NQ];
  subgrid_offsets<D, P, T>(g, org, goff);

  float wx[K][P];
  float wr[K][TS::PR];
  float dk[K];
  int base[K], oidx[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = it.z + tid + k * kThreads;
    if (s < it.z + it.w) {
      float w[D][P];
      base[k] = sample_weights<D, P, T>(tsort, g.samp_off + s, g, org, w);
#pragma unroll
      for (int i = 0; i < P; ++i) wx[k][i] = w[0][i];
      rest_products<D, P>(w, wr[k]);
      oidx[k] = perm[g.samp_off + s];
      dk[k] = dcf ? dcf[g.samp_off + oidx[k]] : 1.0f;
    } else {
      base[k] = 0;
      oidx[k] = -1;
      dk[k] = 0.f;
    }
  }

  for (int e = tid; e < 2 * TS::SUBN; e += kThreads) ldsf[e] = 0.f;
  __syncthreads();

  const float2* ksrc = ksp + g.ksp_off;
  float2* gdst = grid + g.grid_off;
  for (int b = 0; b < g.nbatch; ++b) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (oidx[k] < 0) continue;
      float2 v = ksrc[(long long)b * g.ksp_stride + oidx[k]];
      v.x *= dk[k];
      v.y *= dk[k];
      float* sb = ldsf + 2 * base[k];
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const float vx = wx[k][i] * v.x, vy = wx[k][i] * v.y;
#pragma unroll
        for (int q = 0; q < TS::PR; ++q) {
          const int o = 2 * (i * TS::SUBD1 + rest_off<D, P, T>(q));
          __hip_atomic_fetch_add(sb + o, wr[k][q] * vx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(sb + o + 1, wr[k][q] * vy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
    __syncthreads();
    float2* dst = gdst + (long long)b * g.grid_stride;
#pragma unroll
    for (int q = 0; q < TS::NQ; ++q) {
      const int e = tid + q * kThreads;
      if (goff[q] >= 0) {
        const float2 v = reinterpret_cast<float2*>(ldsf)[e];
        reinterpret_cast<float2*>(ldsf)[e] = make_float2(0.f, 0.f);
        if (v.x != 0.f) unsafeAtomicAdd(&dst[goff[q]].x, v.x);
        if (v.y != 0.f) unsafeAtomicAdd(&dst[goff[q]].y, v.y);
      }
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------------------------
// Trajectory gradient of interpolation, reduced over batch entries:
//   gt[o, d] = sum_b Re( conj(a[b, o]) * sum_p dW_p/dt_d * grid[b, g_p] )
template <int D, int P, int T, int K>
__global__ __launch_bounds__(kThreads) void interp_grad_kernel(const float2* __restrict__ grid,
                                                               const float2* __restrict__ a,
                                                               float* __restrict__ gt,
                                                               const float* __restrict__ tsort,
                                                               const int* __restrict__ perm,
                                                               const Geom* __restrict__ geoms,
                                                               const Item* __restrict__ items, int nitems) {
  using TS = TileShape<D, P, T>;
  extern __shared__ float2 lds2[];
  const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
  if (bid >= nitems) return;
  const Item it = items[bid];
  if (it.w <= 0) return;
  const Geom g = geoms[it.x];
  const int tid = threadIdx.x;
  if (it.y < 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int s = it.z + tid + k * kThreads;
      if (s < it.z + it.w) {
        const int o = perm[g.samp_off + s];
#pragma unroll
        for (int d = 0; d < D; ++d) gt[(g.samp_off + o) * D + d] = 0.f;
      }
    }
    return;
  }

  int org[D];
  tile_origin<D, T>(it.y, g, org);
  int goff[TS::NQ];
  subgrid_offsets<D, P, T>(g, org, goff);

  float w[K][D][P], dw[K][D][P];
  int base[K], oidx[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = it.z + tid + k * kThreads;
    if (s < it.z + it.w) {
      constexpr int SUB = T + P - 1;
      int bs = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const float t = tsort[(g.samp_off + s) * D + d];
        const float sc = d == 0 ? g.out_scale : 1.0f;
        const int lower = kb_weights<P>(t, g.r, g.beta, sc, w[k][d]);
        kb_dweights<P>(t, g.r, g.beta, sc, lower, dw[k][d]);
        bs = bs * SUB + (pmod(lower + g.wrap[d], g.G[d]) - org[d]);
      }
      base[k] = bs;
      oidx[k] = perm[g.samp_off + s];
    } else {
#pragma unroll
      for (int d = 0; d < D; ++d)
#pragma unroll
        for (int j = 0; j < P; ++j) w[k][d][j] = dw[k][d][j] = 0.f;
      base[k] = 0;
      oidx[k] = -1;
    }
  }

  float acc[K][D];
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d) acc[k][d] = 0.f;

  const float2* gsrc = grid + g.grid_off;
  const float2* asrc = a + g.ksp_off;
  float2 pre[TS::NQ];
#pragma unroll
  for (int q = 0; q < TS::NQ; ++q)
    if (goff[q] >= 0) lds2[tid + q * kThreads] = gsrc[goff[q]];
  __syncthreads();

  constexpr int SUB = T + P - 1;
  for (int b = 0; b < g.nbatch; ++b) {
    const int cur = (b & 1) * TS::SUBN;
    const bool more = b + 1 < g.nbatch;
    if (more) {
      const float2* src = gsrc + (long long)(b + 1) * g.grid_stride;
#pragma unroll
      for (int q = 0; q < TS::NQ; ++q)
        if (goff[q] >= 0) pre[q] = src[goff[q]];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (oidx[k] < 0) continue;
      const float2 av = asrc[(long long)b * g.ksp_stride + oidx[k]];
      const float2* sb = lds2 + cur + base[k];
      float2 sd[D];
#pragma unroll
      for (int d = 0; d < D; ++d) sd[d] = make_float2(0.f, 0.f);
      if constexpr (D == 3) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
          float2 cA = {0.f, 0.f}, cB = {0.f, 0.f}, cC = {0.f, 0.f};
#pragma unroll
          for (int j = 0; j < P; ++j) {
            float2 rA = {0.f, 0.f}, rB = {0.f, 0.f};
#pragma unroll
            for (int l = 0; l < P; ++l) {
              const float2 v = sb[(i * SUB + j) * SUB + l];
              rA.x += w[k][2][l] * v.x;  rA.y += w[k][2][l] * v.y;
              rB.x += dw[k][2][l] * v.x; rB.y += dw[k][2][l] * v.y;
            }
            cA.x += w[k][1][j] * rA.x;  cA.y += w[k][1][j] * rA.y;
            cB.x += dw[k][1][j] * rA.x; cB.y += dw[k][1][j] * rA.y;
            cC.x += w[k][1][j] * rB.x;  cC.y += w[k][1][j] * rB.y;
          }
          sd[0].x += dw[k][0][i] * cA.x; sd[0].y += dw[k][0][i] * cA.y;
          sd[1].x += w[k][0][i] * cB.x;  sd[1].y += w[k][0][i] * cB.y;
          sd[2].x += w[k][0][i] * cC.x;  sd[2].y += w[k][0][i] * cC.y;
        }
      } else if constexpr (D == 2) {
#pragma unroll
        for (int i = 0; i < P; ++i) {
          float2 rA = {0.f, 0.f}, rB = {0.f, 0.f};
#pragma unroll
          for (int l = 0; l < P; ++l) {
            const float2 v = sb[i * SUB + l];
            rA.x += w[k][1][l] * v.x;  rA.y += w[k][1][l] * v.y;
            rB.x += dw[k][1][l] * v.x; rB.y += dw[k][1][l] * v.y;
          }
          sd[0].x += dw[k][0][i] * rA.x; sd[0].y += dw[k][0][i] * rA.y;
          sd[1].x += w[k][0][i] * rB.x;  sd[1].y += w[k][0][i] * rB.y;
        }
      } else {
#pragma unroll
        for (int l = 0; l < P; ++l) {
          const float2 v = sb[l];
          sd[0].x += dw[k][0][l] * v.x; sd[0].y += dw[k][0][l] * v.y;
        }
      }
#pragma unroll
      for (int d = 0; d < D; ++d) acc[k][d] += av.x * sd[d].x + av.y * sd[d].y;
    }
    if (more) {
      const int nxt = TS::SUBN - cur;
#pragma unroll
      for (int q = 0; q < TS::NQ; ++q)
        if (goff[q] >= 0) lds2[nxt + tid + q * kThreads] = pre[q];
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (oidx[k] >= 0)
#pragma unroll
      for (int d = 0; d < D; ++d) gt[(g.samp_off + oidx[k]) * D + d] = acc[k][d];
}

// Tile edge per dimensionality (grid points) and samples per thread.
template <int D> struct TileCfg;
template <> struct TileCfg<1> { static constexpr int T = 256; };
template <> struct TileCfg<2> { static constexpr int T = 32; };
template <> struct TileCfg<3> { static constexpr int T = 8; };
constexpr int kK = 2;

enum class Op { Interp, Grid, Grad };

struct LaunchArgs {
  const void* src;     // grid (interp, grad) or ksp (grid)
  void* dst;           // ksp (interp), grid (grid), gt (grad)
  const void* aux;     // dcf (grid) or a (grad)
  const float* tsort;
  const int* perm;
  const Geom* geoms;
  const Item* items;
  int nitems;
  hipStream_t stream;
};

template <int D, int P>
static int launch_op(Op op, const LaunchArgs& a) {
  constexpr int T = TileCfg<D>::T;
  using TS = TileShape<D, P, T>;
  if (a.nitems <= 0) return 0;
  const int nblk = (a.nitems + 7) / 8 * 8;
  switch (op) {
    case Op::Interp:
      interp_kernel<D, P, T, kK><<<nblk, kThreads, 2 * TS::SUBN * sizeof(float2), a.stream>>>(
          (const float2*)a.src, (float2*)a.dst, a.tsort, a.perm, a.geoms, a.items, a.nitems);
      break;
    case Op::Grid:
      grid_kernel<D, P, T, kK><<<nblk, kThreads, TS::SUBN * sizeof(float2), a.stream>>>(
          (float2*)a.dst, (const float2*)a.src, (const float*)a.aux, a.tsort, a.perm, a.geoms, a.items,
          a.nitems);
      break;
    case Op::Grad:
      interp_grad_kernel<D, P, T, kK><<<nblk, kThreads, 2 * TS::SUBN * sizeof(float2), a.stream>>>(
          (const float2*)a.src, (const float2*)a.aux, (float*)a.dst, a.tsort, a.perm, a.geoms, a.items,
          a.nitems);
      break;
  }
  MRA_CHECK_LAUNCH("nufft kernel");
  return 0;
}

template <int D>
static int dispatch_p(int P, Op op, const LaunchArgs& a) {
  switch (P) {
    case 2: return launch_op<D, 2>(op, a);
    case 3: return launch_op<D, 3>(op, a);
    case 4: return launch_op<D, 4>(op, a);
    case 5: return launch_op<D, 5>(op, a);
    case 6: return launch_op<D, 6>(op, a);
    case 7: return launch_op<D, 7>(op, a);
    case 8: return launch_op<D, 8>(op, a);
    case 9: return launch_op<D, 9>(op, a);
    default:
      set_last_error("kernel width unsupported: floor(width)+1 must be in 2..9");
      return 1;
  }
}

static int dispatch(int ndim, int P, Op op, const LaunchArgs& a) {
  switch (ndim) {
    case 1: return dispatch_p<1>(P, op, a);
    case 2: return dispatch_p<2>(P, op, a);
    case 3: return dispatch_p<3>(P, op, a);
    default:
      set_last_error("ndim must be 1, 2 or 3");
      return 1;
  }
}

}  // namespace mra

using namespace mra;

extern "C" {

int mra_geom_size() { return (int)sizeof(Geom); }

int mra_default_tile(int ndim) {
  switch (ndim) {
    case 1: return TileCfg<1>::T;
    case 2: return TileCfg<2>::T;
    case 3: return TileCfg<3>::T;
    default: return 0;
  }
}

int mra_samples_per_item() { return kThreads * kK; }

// Fill one Geom record (host memory) for a problem.
int mra_geom_fill(void* dst, long long grid_off, long long grid_stride, long long ksp_off, long long ksp_stride,
                  long long samp_off, int nbatch, int npts, int ndim, const int* G, int tile, const int* wrap,
                  float beta, float width, float out_scale) {
  MRA_TRY_BEGIN
  if (ndim < 1 || ndim > kMaxDim) MRA_FAIL("geom: ndim must be 1..3");
  Geom g{};
  g.grid_off = grid_off;
  g.grid_stride = grid_stride;
  g.ksp_off = ksp_off;
  g.ksp_stride = ksp_stride;
  g.samp_off = samp_off;
  g.nbatch = nbatch;
  g.npts = npts;
  long long gtot = 1;
  for (int d = 0; d < kMaxDim; ++d) {
    g.G[d] = d < ndim ? G[d] : 1;
    g.ntile[d] = d < ndim ? (G[d] + tile - 1) / tile : 1;
    g.wrap[d] = (d < ndim && wrap) ? wrap[d] : 0;
    gtot *= g.G[d];
  }
  if (gtot > 0x7fffffffLL) MRA_FAIL("geom: grid larger than 2^31-1 points");
  g.beta = beta;
  g.r = 0.5f * width;
  g.out_scale = out_scale;
  *(Geom*)dst = g;
  MRA_TRY_END
}

int mra_interp(int ndim, int P, const void* grid, void* out, const float* tsort, const int* perm,
               const void* geoms, const void* items, int nitems, void* stream) {
  LaunchArgs a{grid, out, nullptr, tsort, perm, (const Geom*)geoms, (const Item*)items, nitems,
               (hipStream_t)stream};
  return dispatch(ndim, P, Op::Interp, a);
}

int mra_grid(int ndim, int P, void* grid, const void* ksp, const float* dcf, const float* tsort, const int* perm,
             const void* geoms, const void* items, int nitems, void* stream) {
  LaunchArgs a{ksp, grid, dcf, tsort, perm, (const Geom*)geoms, (const Item*)items, nitems, (hipStream_t)stream};
  return dispatch(ndim, P, Op::Grid, a);
}

int mra_interp_grad(int ndim, int P, const void* grid, const void* a, float* grad_t, const float* tsort,
                    const int* perm, const void* geoms, const void* items, int nitems, void* stream) {
  LaunchArgs la{grid, grad_t, a, tsort, perm, (const Geom*)geoms, (const Item*)items, nitems,
                (hipStream_t)stream};
  return dispatch(ndim, P, Op::Grad, la);
}

}  // extern "C"
