// Kaiser-Bessel interpolation (gather), gridding (scatter) and the trajectory
// gradient of interpolation, on tile-sorted samples.
//
// Reference semantics (ungrid.py:356-495, grid.py:361-502, kernels.py:180-393):
//   out[b, p] = sum_{g in nbhd(t_p)} prod_d KB(|g_d - t_d| / r) * grid[b, g mod G]
//   KB(u) = I0(beta sqrt(1 - u^2)),  |g_d - t_d| <= r = w/2,  g_d = ceil(t_d - r) + j.
//
// Decomposition. Samples are binned by the tile (T^D grid points) holding the
// lower corner of their patch; a tile's neighbourhood ("subgrid", T + P - 1
// points per dim, wrapped mod G) holds every grid point its samples touch.
// A work item is one tile's slice of at most S sorted samples; a block handles
// one (item, coil chunk) pair, so hot tiles split over samples and every tile
// splits over coils.
//
//  plan_weights  once per trajectory: the D x P separable KB weights and the
//                subgrid offset of each sorted sample.
//  interp        the block stages its coils' subgrids in LDS, then each thread
//                walks samples: 25 weight products, P^D LDS reads at
//                compile-time offsets, one output per coil.
//  grid_partial  one wave accumulates its samples into LDS subgrids with plain
//                (atomic-free) read-modify-writes, summing runs of samples in
//                register windows, and stores each subgrid to a scratch slot with
//                plain coalesced stores -- no global atomics.
//  grid_reduce   one block per (output tile, coil) sums, in a fixed order, the
//                scratch subgrids of every item whose subgrid covers the tile
//                and writes the grid once. Device-scope float atomics run at a
//                few G/s on MI355X, so gridding never uses them.
//  interp_grad   derivative weights, reduced over coils per sample.
#include <mutex>
#include <type_traits>

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
  static constexpr int TN = ipow(T, D);             // tile points
};

// Tile edge per dimensionality (grid points): the default, and the edges the kernels are
// compiled for (2D: 8, 16 or 32, chosen per plan; 1D and 3D: one edge).
template <int D> struct TileCfg;
template <> struct TileCfg<1> { static constexpr int T = 256; };
template <> struct TileCfg<2> { static constexpr int T = 16; };
template <> struct TileCfg<3> { static constexpr int T = 8; };
constexpr bool tile_compiled(int D, int T) {
  return D == 1 ? T == 256 : D == 2 ? (T == 8 || T == 16 || T == 32) : D == 3 ? T == 8 : false;
}
// Tile edge the calls of this thread use, per dimensionality (set by mra_set_tile). Plans,
// Geom records and kernel calls on one trajectory must use the same edge.
static thread_local int g_tile[4] = {0, TileCfg<1>::T, TileCfg<2>::T, TileCfg<3>::T};

#ifndef MRA_INTERP_LDS_KB
#define MRA_INTERP_LDS_KB 32
#endif
constexpr int kLdsBudget = MRA_INTERP_LDS_KB * 1024;  // bytes of LDS subgrids per interp block (more, smaller blocks hide latency better)
// Waves per SIMD that the LDS budget leaves room for (160 KiB LDS per CU, 4 SIMDs). Telling
// the compiler this bound (amdgpu_waves_per_eu) lets it spend the registers that the
// unreachable occupancies would have needed on more LDS reads in flight: at 3D P = 5 the
// inner loop issues 8 128-bit reads before its first wait instead of 2.
constexpr int kInterpWavesPerEU = (160 * 1024 / kLdsBudget) * (kThreads / 64) / 4;
static_assert(kInterpWavesPerEU >= 1 && kInterpWavesPerEU <= 8, "interp LDS budget out of range");

// Offset in the subgrid of patch point q of the trailing D-1 dims.
template <int D, int P, int T>
__device__ __forceinline__ constexpr int rest_off(int q) {
  constexpr int SUB = T + P - 1;
  if constexpr (D == 3) return (q / P) * SUB + (q % P);
  else if constexpr (D == 2) return q;
  else return 0;
}

template <int D, int T>
__device__ __forceinline__ void tile_coords(int tile, const Geom& g, int (&tc)[D]) {
  int rem = tile;
#pragma unroll
  for (int d = D - 1; d >= 0; --d) {
    tc[d] = rem % g.ntile[d];
    rem /= g.ntile[d];
  }
}

// Global (wrapped) linear offsets of the subgrid points this thread moves.
template <int D, int P, int T>
__device__ __forceinline__ void subgrid_offsets(const Geom& g, const int (&tc)[D],
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
      for (int d = 0; d < D; ++d) lin = lin * g.G[d] + (tc[d] * T + c[d]) % g.G[d];
      goff[q] = lin;
    } else {
      goff[q] = -1;
    }
  }
}

// Returns whether the last patch point of some dimension carries weight. A point is in the
// support when |g - t| <= w/2, so a sample reaches P = floor(w) + 1 points along a dimension
// only where t - w/2 falls on a grid integer (for an integer width; the k-space centre of a
// radial or kooshball trajectory, say); everywhere else that point's weight is exactly 0.
template <int D, int P>
__device__ __forceinline__ bool load_weights(const float* __restrict__ wsort, long long s, float scale,
                                             float (&wx)[P], float (&wr)[ipow(P, D - 1)]) {
  const float* w = wsort + s * (D * P);
  float wd[D][P];
  bool edge = false;
#pragma unroll
  for (int d = 0; d < D; ++d) {
#pragma unroll
    for (int j = 0; j < P; ++j) wd[d][j] = w[d * P + j];
    edge = edge || wd[d][P - 1] != 0.0f;
  }
#pragma unroll
  for (int i = 0; i < P; ++i) wx[i] = wd[0][i] * scale;
  if constexpr (D == 3) {
#pragma unroll
    for (int j = 0; j < P; ++j)
#pragma unroll
      for (int l = 0; l < P; ++l) wr[j * P + l] = wd[1][j] * wd[2][l];
  } else if constexpr (D == 2) {
#pragma unroll
    for (int l = 0; l < P; ++l) wr[l] = wd[1][l];
  } else {
    wr[0] = 1.0f;
  }
  return edge;
}

// Index in wr (patch points of dims 1..D-1, P per dim) of point u of the leading PE x PE
// block (PE <= P points per dim).
template <int D, int P, int PE>
__device__ __forceinline__ constexpr int sub_q(int u) {
  if constexpr (D == 3) return (u / PE) * P + (u % PE);
  else if constexpr (D == 2) return u;
  else return 0;
}

// sum over the first PE patch points of every dimension of w * v (one coil, complex).
// With PE = P - 1 and zero weights on the last points this is the full sum bit for bit:
// each skipped term is a fused multiply-add of 0 into the running sum.
template <int D, int P, int T, int PE>
__device__ __forceinline__ float2 patch_sum(const float2* sb, const float (&wx)[P],
                                            const float (&wr)[ipow(P, D - 1)]) {
  using TS = TileShape<D, P, T>;
  constexpr int NR = ipow(PE, D - 1);
  constexpr int UI = NR <= 36 ? PE : 1;  // unroll dim 0 unless the registers would cap occupancy
  float ax = 0.f, ay = 0.f;
#pragma unroll UI
  for (int i = 0; i < PE; ++i) {
    float bx = 0.f, by = 0.f;
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const int q = sub_q<D, P, PE>(u);
      const float2 v = sb[i * TS::SUBD1 + rest_off<D, P, T>(q)];
      bx += wr[q] * v.x;
      by += wr[q] * v.y;
    }
    ax += wx[i] * bx;
    ay += wx[i] * by;
  }
  return make_float2(ax, ay);
}

// Same for a coil pair stored interleaved (float4 = two coils' values of a point).
template <int D, int P, int T, int PE>
__device__ __forceinline__ float4 patch_sum_pair(const float4* sb, const float (&wx)[P],
                                                 const float (&wr)[ipow(P, D - 1)]) {
  using TS = TileShape<D, P, T>;
  constexpr int NR = ipow(PE, D - 1);
  float ax = 0.f, ay = 0.f, az = 0.f, aw = 0.f;
#ifdef MRA_INTERP_PAIR_UNROLL
#pragma unroll
#else
#pragma unroll 1
#endif
  for (int i = 0; i < PE; ++i) {
    float bx = 0.f, by = 0.f, bz = 0.f, bw = 0.f;
#pragma unroll
    for (int u = 0; u < NR; ++u) {
      const int q = sub_q<D, P, PE>(u);
      const float4 v = sb[i * TS::SUBD1 + rest_off<D, P, T>(q)];
      bx += wr[q] * v.x;
      by += wr[q] * v.y;
      bz += wr[q] * v.z;
      bw += wr[q] * v.w;
    }
    ax += wx[i] * bx;
    ay += wx[i] * by;
    az += wx[i] * bz;
    aw += wx[i] * bw;
  }
  return make_float4(ax, ay, az, aw);
}

// --------------------------------------------------------------------------------------------
// Kaiser-Bessel shape per dimension (all equal except for sigpy-style per-dim kernels).
struct DimBetas {
  float b[kMaxDim];
};

// Per sorted sample: KB weights along each dim (unscaled) and the subgrid base offset.
template <int D, int P, int T>
__global__ __launch_bounds__(256) void plan_weights_kernel(const float* __restrict__ tsort, int npts, Geom g,
                                                           DimBetas bt, float* __restrict__ wsort,
                                                           int* __restrict__ lbase) {
  constexpr int SUB = T + P - 1;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= npts) return;
  float w[P];
  int base = 0;
  bool ok = true;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const float t = tsort[(long long)s * D + d];
    ok = ok && isfinite(t);
    const int lower = ok ? kb_weights<P>(t, g.r, bt.b[d], 1.0f, w) : 0;
    const int lw = ok ? pmod(lower + g.wrap[d], g.G[d]) % T : 0;
    base = base * SUB + lw;
#pragma unroll
    for (int j = 0; j < P; ++j) wsort[(long long)s * D * P + d * P + j] = ok ? w[j] : 0.0f;
  }
  lbase[s] = ok ? base : 0;
}

// --------------------------------------------------------------------------------------------
template <int D, int P, int T>
__global__ __launch_bounds__(kThreads) void interp_kernel(const float2* __restrict__ grid,
                                                          float2* __restrict__ out,
                                                          const float* __restrict__ wsort,
                                                          const int* __restrict__ lbase,
                                                          const int* __restrict__ perm,
                                                          const Geom* __restrict__ geoms,
                                                          const Item* __restrict__ items, int nitems, int ccoils,
                                                          int nchunks) {
  using TS = TileShape<D, P, T>;
  extern __shared__ float2 lds2[];
  int iidx, cc;
  item_of_block(blockIdx.x, nchunks, iidx, cc);
  if (iidx >= nitems) return;
  const Item it = items[iidx];
  if (it.w <= 0) return;
  const Geom g = geoms[it.x];
  const int c0 = cc * ccoils;
  if (c0 >= g.nbatch) return;
  const int nc = min(ccoils, g.nbatch - c0);
  const int tid = threadIdx.x;
  float2* odst = out + g.ksp_off + (long long)c0 * g.ksp_stride;

  if (it.y < 0) {  // samples with non-finite coordinates: defined as 0
    for (int s = it.z + tid; s < it.z + it.w; s += kThreads) {
      const int o = sample_slot(perm, g, g.samp_off + s);
      for (int cl = 0; cl < nc; ++cl) odst[(long long)cl * g.ksp_stride + o] = make_float2(0.f, 0.f);
    }
    return;
  }

  int tc[D];
  tile_coords<D, T>(it.y, g, tc);
  {
    int goff[TS::NQ];
    subgrid_offsets<D, P, T>(g, tc, goff);
    const float2* gsrc = grid + g.grid_off + (long long)c0 * g.grid_stride;
    for (int cl = 0; cl < nc; ++cl) {
#pragma unroll
      for (int q = 0; q < TS::NQ; ++q)
        if (goff[q] >= 0) lds2[cl * TS::SUBN + tid + q * kThreads] = gsrc[(long long)cl * g.grid_stride + goff[q]];
    }
  }
  __syncthreads();

  for (int s = it.z + tid; s < it.z + it.w; s += kThreads) {
    const long long gs = g.samp_off + s;
    float wx[P], wr[TS::PR];
    const bool edge = load_weights<D, P>(wsort, gs, g.out_scale, wx, wr);
    const bool full = __any(edge);  // wave-uniform: P points per dim only where a lane needs them
    const int base = lbase[gs];
    const int o = sample_slot(perm, g, gs);
    for (int cl = 0; cl < nc; ++cl) {
      const float2* sb = lds2 + cl * TS::SUBN + base;
      odst[(long long)cl * g.ksp_stride + o] =
          full ? patch_sum<D, P, T, P>(sb, wx, wr) : patch_sum<D, P, T, P - 1>(sb, wx, wr);
    }
  }
}

// --------------------------------------------------------------------------------------------
// Interpolation with the coils of a chunk in pairs: the LDS subgrids are stored coil-pair
// interleaved (one float4 = two coils' values of a grid point), so every patch point is
// one 128-bit LDS read serving both coils (gfx950 moves ~2.7x the bytes per clock with
// 128-bit reads than with 64-bit ones). Same work decomposition as interp_kernel; an odd
// coil count pads its last pair with zeros.
template <int D, int P, int T>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, kInterpWavesPerEU))) void interp_pair_kernel(const float2* __restrict__ grid,
                                                               float2* __restrict__ out,
                                                               const float* __restrict__ wsort,
                                                               const int* __restrict__ lbase,
                                                               const int* __restrict__ perm,
                                                               const Geom* __restrict__ geoms,
                                                               const Item* __restrict__ items, int nitems,
                                                               int ccoils, int nchunks) {
  using TS = TileShape<D, P, T>;
  extern __shared__ float4 lds4[];
  int iidx, cc;
  item_of_block(blockIdx.x, nchunks, iidx, cc);
  if (iidx >= nitems) return;
  const Item it = items[iidx];
  if (it.w <= 0) return;
  const Geom g = geoms[it.x];
  const int c0 = cc * ccoils;
  if (c0 >= g.nbatch) return;
  const int nc = min(ccoils, g.nbatch - c0);
  const int npair = (nc + 1) / 2;
  const int tid = threadIdx.x;
  float2* odst = out + g.ksp_off + (long long)c0 * g.ksp_stride;

  if (it.y < 0) {  // samples with non-finite coordinates: defined as 0
    for (int s = it.z + tid; s < it.z + it.w; s += kThreads) {
      const int o = sample_slot(perm, g, g.samp_off + s);
      for (int cl = 0; cl < nc; ++cl) odst[(long long)cl * g.ksp_stride + o] = make_float2(0.f, 0.f);
    }
    return;
  }

  int tc[D];
  tile_coords<D, T>(it.y, g, tc);
  {
    int goff[TS::NQ];
    subgrid_offsets<D, P, T>(g, tc, goff);
    const float2* gsrc = grid + g.grid_off + (long long)c0 * g.grid_stride;
    for (int cp = 0; cp < npair; ++cp) {
      const bool two = 2 * cp + 1 < nc;
#pragma unroll
      for (int q = 0; q < TS::NQ; ++q)
        if (goff[q] >= 0) {
          const float2 a = gsrc[(long long)(2 * cp) * g.grid_stride + goff[q]];
          const float2 b = two ? gsrc[(long long)(2 * cp + 1) * g.grid_stride + goff[q]] : make_float2(0.f, 0.f);
          lds4[cp * TS::SUBN + tid + q * kThreads] = make_float4(a.x, a.y, b.x, b.y);
        }
    }
  }
  __syncthreads();

  for (int s = it.z + tid; s < it.z + it.w; s += kThreads) {
    const long long gs = g.samp_off + s;
    float wx[P], wr[TS::PR];
    const bool edge = load_weights<D, P>(wsort, gs, g.out_scale, wx, wr);
    const bool full = __any(edge);  // wave-uniform: P points per dim only where a lane needs them
    const int base = lbase[gs];
    const int o = sample_slot(perm, g, gs);
    for (int cp = 0; cp < npair; ++cp) {
      const float4* sb = lds4 + cp * TS::SUBN + base;
      const float4 a = full ? patch_sum_pair<D, P, T, P>(sb, wx, wr) : patch_sum_pair<D, P, T, P - 1>(sb, wx, wr);
      odst[(long long)(2 * cp) * g.ksp_stride + o] = make_float2(a.x, a.y);
      if (2 * cp + 1 < nc) odst[(long long)(2 * cp + 1) * g.ksp_stride + o] = make_float2(a.z, a.w);
    }
  }
}

// --------------------------------------------------------------------------------------------
// Gridding phase 1, atomic-free. One wave per (item, coil chunk).
//
// Lanes are laid out as NS coil slots of PR lanes; lane (cs, r) owns coil c0+cs and
// patch position r of the trailing D-1 dims and loops over the P positions of dim 0.
// Every LDS read-modify-write of one instruction hits a distinct address, so plain LDS
// loads and stores replace LDS float atomics (which run ~20x slower than plain LDS
// traffic on MI355X). Consecutive samples may touch the same addresses, which is safe
// because one wave's LDS operations complete in order.
//
// Plane ring. The samples of an item are sorted by patch corner, dim 0 slowest, so the
// corner's dim-0 coordinate (its plane) never decreases along the item and a sample
// touches planes bi .. bi+P-1 only. Each coil therefore keeps a ring of P subgrid
// planes in LDS instead of the whole subgrid: when the next run starts at a higher
// plane, the planes below it are final -- they are stored to the item's scratch slot
// and their ring slots cleared for the planes entering at the top. This cuts LDS per
// coil from SUB^D to P*SUB^(D-1) points and raises the waves per CU accordingly.
//
// Runs. Samples with the same corner are summed in registers; each run costs one
// read-modify-write of its patch, and the read half of the next run's read-modify-
// write is issued right after the current run's writes (LDS order makes it see them),
// so by the time that run ends its old values are already in registers.
//
// Stage. Sample data reach the serial loop through an LDS stage of CH samples: lane l
// gathers sample s0+l (corner, D*P weights, its coils' k-space values times dcf), and
// the next chunk's global loads are issued before the current chunk is accumulated.
// Within a chunk a ring of LA register sets is loaded ahead of use, and each sample's
// corner / run-end word comes from v_readlane.
//
// Scratch slot of (item, batch entry): scratch + scratch_off + ((item - item_base) * nbatch + b) * SUBN,
// dense row-major subgrid.

#ifndef MRA_GRID_STAGE
#define MRA_GRID_STAGE 64
#endif
#ifndef MRA_GRID_RING
#define MRA_GRID_RING 6
#endif
#ifndef MRA_GRID_CPL
#define MRA_GRID_CPL 1
#endif
#ifndef MRA_GRID_NSMAX
#define MRA_GRID_NSMAX 12  // most coil slots per wave (12 over 8: batched 2D +14 %, same box)
#endif

// LDS padding (row pitch of the last dim += E in 3D, coil-slot stride += SP), from a
// bank model of the run read-modify-write (64 banks, 8-byte lanes) per (D, P).
template <int D, int P> struct GridPad { static constexpr int E = 0, SP = 0; };
template <> struct GridPad<1, 8> { static constexpr int E = 0, SP = 1; };
template <> struct GridPad<2, 3> { static constexpr int E = 0, SP = 3; };
template <> struct GridPad<2, 6> { static constexpr int E = 0, SP = 5; };
template <> struct GridPad<2, 7> { static constexpr int E = 0, SP = 2; };
template <> struct GridPad<2, 9> { static constexpr int E = 0, SP = 17; };
template <> struct GridPad<3, 2> { static constexpr int E = 1, SP = 0; };
template <> struct GridPad<3, 3> { static constexpr int E = 1, SP = 19; };
template <> struct GridPad<3, 4> { static constexpr int E = 1, SP = 0; };
template <> struct GridPad<3, 5> { static constexpr int E = 0, SP = 9; };
template <> struct GridPad<3, 9> { static constexpr int E = 2, SP = 0; };

template <int D, int P, int T>
struct RowLanes {
  static constexpr int SUB = T + P - 1;
  static constexpr int PR = ipow(P, D - 1);
  static constexpr int SUBN = TileShape<D, P, T>::SUBN;
  static constexpr int NR = (PR + 63) / 64;                        // rest positions per lane
  static constexpr int LPS = PR < 64 ? PR : 64;                    // lanes per coil slot
  static constexpr int NS = (64 / LPS) < MRA_GRID_NSMAX ? (64 / LPS) : MRA_GRID_NSMAX;  // lane groups per wave
  static constexpr int CPL = MRA_GRID_CPL;                         // coils per lane
  static constexpr int NC = NS * CPL;                              // coils per wave (lane group cs
                                                                   // owns coils cs + u*NS, u < CPL)
  static constexpr int PP = ipow(SUB, D - 1);                      // dense points per plane
  // padded LDS layout: row pitch PJ (3D), plane pitch PL, ring of P planes per slot
  static constexpr int PJ = D == 3 ? SUB + GridPad<D, P>::E : 1;
  static constexpr int PL = D == 3 ? SUB * PJ : (D == 2 ? SUB : 1);
  static constexpr int SLOT = P * PL;
  static constexpr int SSTR = NS > 1 ? SLOT + GridPad<D, P>::SP : SLOT;
  static constexpr int LDS_FLOAT2 = (NC - 1) * SSTR + SLOT;
  // LDS stage of CH samples: rec[.][RECW] = the P dim-0 weights, the weights of dims
  // 1..D-1 as [(D-1)*P][WS] and the k-space values as float2 [NC][WS]. WS is odd so the
  // P weights one lane may pick for a sample fall in distinct banks, and covers the
  // LA-1 loads the ring issues past the end of a stage.
  static constexpr int CH = MRA_GRID_STAGE;                        // samples per stage
  static constexpr int LA = MRA_GRID_RING;                         // register ring depth
  static constexpr int RECW = (P + 3) / 4 * 4;
  static constexpr int WS = CH + LA + 1;
  static constexpr int NWL = (D - 1) * P > 0 ? (D - 1) * P : 1;
  // dense in-plane index -> padded in-plane offset
  static __device__ __forceinline__ int plane_off(int r) {
    if constexpr (D == 3) return (r / SUB) * PJ + r % SUB;
    else return r;
  }
};

template <int D, int P, int NS>
struct StageRegs {
  int base;
  float w[D * P];
  float2 v[NS];
};

template <int D, int P, int T>
__device__ __forceinline__ void stage_load(StageRegs<D, P, RowLanes<D, P, T>::NC>& r, bool ok, long long gs, int o,
                                           const float2* __restrict__ ksp, const float* __restrict__ dcf,
                                           const float* __restrict__ wsort, const int* __restrict__ lbase,
                                           const Geom& g, int c0, int nc) {
  using RL = RowLanes<D, P, T>;
  r.base = ok ? lbase[gs] : 0;
  const float* w = wsort + (ok ? gs : 0) * (D * P);
#pragma unroll
  for (int e = 0; e < D * P; ++e) r.w[e] = ok ? w[e] : 0.0f;
  const float sc = (ok && dcf) ? dcf[g.dcf_plan_order ? gs : g.samp_off + o] * g.out_scale : g.out_scale;
  const float2* kc = ksp + g.ksp_off + (long long)c0 * g.ksp_stride + (ok ? o : 0);
#pragma unroll
  for (int cs = 0; cs < RL::NC; ++cs) {
    float2 v = (ok && cs < nc) ? kc[(long long)cs * g.ksp_stride] : make_float2(0.f, 0.f);
    r.v[cs] = make_float2(v.x * sc, v.y * sc);
  }
}

// Stage values of one sample as loaded (no arithmetic on them until they are used,
// so later samples' loads stay in flight while this one is accumulated).
template <int D, int P, int NR, int CPL>
struct SampleView {
  float wx[P];           // dim-0 weights
  float wa[NR], wb[NR];  // weights of dims 1 and 2 at this lane's rest positions
  float2 v[CPL];         // k-space values of this lane's coils
};

template <int D, int P, int T>
__device__ __forceinline__ void read_view(SampleView<D, P, RowLanes<D, P, T>::NR, RowLanes<D, P, T>::CPL>& sv,
                                          int t, const float* rec,
                                          const float* wl, const float2* vmine,
                                          const int (&ridx)[RowLanes<D, P, T>::NR][D > 1 ? D - 1 : 1]) {
  using RL = RowLanes<D, P, T>;
  const float* rt = rec + t * RL::RECW;
#pragma unroll
  for (int i = 0; i < P; ++i) sv.wx[i] = rt[i];
#pragma unroll
  for (int u = 0; u < RL::CPL; ++u) sv.v[u] = vmine[u * RL::NS * RL::WS + t];
#pragma unroll
  for (int q = 0; q < RL::NR; ++q) {
    if constexpr (D >= 2) sv.wa[q] = wl[ridx[q][0] * RL::WS + t];
    if constexpr (D == 3) sv.wb[q] = wl[(P + ridx[q][1]) * RL::WS + t];
  }
}

// Run word of a stage sample (published by the lane that gathered it): bits 0..15 the
// padded in-plane offset of its corner, bits 16..23 its plane, bit 30 = last of its run.
__device__ __forceinline__ int word_rest(int w) { return w & 0xffff; }
__device__ __forceinline__ int word_plane(int w) { return (w >> 16) & 0xff; }

// Ring slot offsets of the P planes of a patch whose corner plane is bi.
template <int D, int P, int T>
__device__ __forceinline__ void patch_planes(int bi, int (&off)[P]) {
  using RL = RowLanes<D, P, T>;
  int s = bi % P;
#pragma unroll
  for (int i = 0; i < P; ++i) {
    off[i] = s * RL::PL;
    s = s + 1 == P ? 0 : s + 1;
  }
}

// Old subgrid values at the patch of the run whose word is w.
template <int D, int P, int T, int PE = P>
__device__ __forceinline__ void read_run(int w, const bool (&rok)[RowLanes<D, P, T>::NR],
                                         const int (&roff)[RowLanes<D, P, T>::NR], const float2* sgrid,
                                         float2 (&pre)[RowLanes<D, P, T>::CPL][RowLanes<D, P, T>::NR][P]) {
  using RL = RowLanes<D, P, T>;
  int off[P];
  patch_planes<D, P, T>(word_plane(w), off);
#pragma unroll
  for (int u = 0; u < RL::CPL; ++u)
#pragma unroll
    for (int q = 0; q < RL::NR; ++q) {
      const float2* pt = sgrid + u * RL::NS * RL::SSTR + word_rest(w) + (rok[q] ? roff[q] : 0);
#pragma unroll
      for (int i = 0; i < PE; ++i) pre[u][q][i] = pt[off[i]];
    }
}

// Planes lo .. hi-1 are final: store them to the scratch slots (dense) and clear their
// ring slots, which then hold planes lo+P .. hi+P-1. Planes never held by the ring
// (p >= lo0 + P, with lo0 the first plane not yet stored) are stored as zeros.
template <int D, int P, int T>
__device__ __forceinline__ void retire_planes(int lo, int hi, int nc, float2* lds2, float2* __restrict__ dst) {
  using RL = RowLanes<D, P, T>;
  constexpr int NE = RL::NC * RL::PP;
  const int lane = threadIdx.x;
  for (int p = lo; p < hi; ++p) {
    const bool held = p < lo + P;
    const int slot = (p % P) * RL::PL;
#pragma unroll
    for (int e0 = 0; e0 < NE; e0 += 64) {
      const int e = e0 + lane;
      const int c = e / RL::PP, r = e - c * RL::PP;
      if (e < NE && c < nc) {
        float2* src = lds2 + c * RL::SSTR + slot + RL::plane_off(r);
        float2 v = make_float2(0.f, 0.f);
        if (held) {
          v = *src;
          *src = make_float2(0.f, 0.f);
        }
        dst[(long long)c * RL::SUBN + (long long)p * RL::PP + r] = v;
      }
    }
  }
}

// Accumulate stage sample t; at the end of its run (bit 30 of its word) write old + sum
// back and clear the sums; then, before the next run's old values are read, store the
// planes the next run has moved past. PE = P - 1 when no sample of the item reaches the
// last patch plane (its dim-0 weight there is exactly 0): that plane is left untouched.
template <int D, int P, int T, int PE = P>
__device__ __forceinline__ void accumulate_sample(
    const SampleView<D, P, RowLanes<D, P, T>::NR, RowLanes<D, P, T>::CPL>& sv, int t, int n, int pub,
    const bool (&rok)[RowLanes<D, P, T>::CPL][RowLanes<D, P, T>::NR], const int (&roff)[RowLanes<D, P, T>::NR],
    float2* sgrid, float2* lds2, float2* __restrict__ dst, int nc, int& lo,
    float2 (&acc)[RowLanes<D, P, T>::CPL][RowLanes<D, P, T>::NR][P],
    float2 (&pre)[RowLanes<D, P, T>::CPL][RowLanes<D, P, T>::NR][P]) {
  using RL = RowLanes<D, P, T>;
  const int raw = __builtin_amdgcn_readlane(pub, t);  // scalar, no LDS round trip
#pragma unroll
  for (int q = 0; q < RL::NR; ++q) {
    float wr = 1.0f;
    if constexpr (D == 3) wr = sv.wa[q] * sv.wb[q];
    else if constexpr (D == 2) wr = sv.wa[q];
#pragma unroll
    for (int u = 0; u < RL::CPL; ++u) {
      const float vx = sv.v[u].x * wr, vy = sv.v[u].y * wr;
#pragma unroll
      for (int i = 0; i < PE; ++i) {
        acc[u][q][i].x += sv.wx[i] * vx;
        acc[u][q][i].y += sv.wx[i] * vy;
      }
    }
  }
  if (raw >> 30) {  // last sample of its run (uniform)
    int off[P];
    patch_planes<D, P, T>(word_plane(raw), off);
#pragma unroll
    for (int u = 0; u < RL::CPL; ++u)
#pragma unroll
      for (int q = 0; q < RL::NR; ++q) {
        if (rok[u][q]) {
          float2* pt = sgrid + u * RL::NS * RL::SSTR + word_rest(raw) + roff[q];
#pragma unroll
          for (int i = 0; i < PE; ++i)
            pt[off[i]] = make_float2(pre[u][q][i].x + acc[u][q][i].x, pre[u][q][i].y + acc[u][q][i].y);
        }
#pragma unroll
        for (int i = 0; i < PE; ++i) acc[u][q][i] = make_float2(0.f, 0.f);
      }
    if (t + 1 < n) {
      const int nw = __builtin_amdgcn_readlane(pub, t + 1);
      const int bn = word_plane(nw);
      if (bn > lo) {
        retire_planes<D, P, T>(lo, bn, nc, lds2, dst);
        lo = bn;
      }
      read_run<D, P, T, PE>(nw, rok[0], roff, sgrid, pre);
    }
  }
}

// One (item, coil chunk) of grid_partial_kernel with PE rows per patch (P, or P - 1 when no
// sample of the item reaches the last patch plane).
template <int D, int P, int T, int PE>
__device__ __forceinline__ void grid_partial_item(const float2* __restrict__ ksp, const float* __restrict__ dcf,
                                                  const float* __restrict__ wsort, const int* __restrict__ lbase,
                                                  const int* __restrict__ perm, const Geom& g, const Item& it,
                                                  int iidx, int c0, int nc, float2* __restrict__ scratch,
                                                  float2* lds2, float* rec, float* wl, float2* vst) {
  using RL = RowLanes<D, P, T>;
  constexpr int SUB = RL::SUB;
  const int lane = threadIdx.x;
  const int cnt = it.w;
  const long long gs0 = g.samp_off + it.z;
  constexpr int CH = RL::CH;
  float2* dst = scratch + g.scratch_off + ((long long)(iidx - g.item_base) * g.nbatch + c0) * RL::SUBN;

  // first chunk's gathers go out before the ring is cleared
  StageRegs<D, P, RL::NC> sr;
  {
    const bool ok = lane < CH && lane < cnt;
    const int o = ok ? sample_slot(perm, g, gs0 + lane) : 0;
    stage_load<D, P, T>(sr, ok, gs0 + lane, o, ksp, dcf, wsort, lbase, g, c0, nc);
  }
  int onext = (lane < CH && CH + lane < cnt) ? sample_slot(perm, g, gs0 + CH + lane) : 0;

  for (int e = lane; e < RL::LDS_FLOAT2; e += 64) lds2[e] = make_float2(0.f, 0.f);

  // this lane's coil slot and rest positions
  const int cs = lane / RL::LPS;
  const int rl = lane - cs * RL::LPS;
  const bool active = cs < nc;
  int roff[RL::NR], ridx[RL::NR][D > 1 ? D - 1 : 1];
  bool rok[RL::CPL][RL::NR];
#pragma unroll
  for (int q = 0; q < RL::NR; ++q) {
    const int r = rl + q * 64;
#pragma unroll
    for (int u = 0; u < RL::CPL; ++u) rok[u][q] = cs + u * RL::NS < nc && cs < RL::NS && r < RL::PR;
    const int rr = rok[0][q] ? r : 0;
    if constexpr (D == 3) {
      ridx[q][0] = rr / P;
      ridx[q][1] = rr % P;
      roff[q] = ridx[q][0] * RL::PJ + ridx[q][1];
    } else if constexpr (D == 2) {
      ridx[q][0] = rr;
      roff[q] = rr;
    } else {
      ridx[q][0] = 0;
      roff[q] = 0;
    }
  }
  float2* sgrid = lds2 + (active ? cs : 0) * RL::SSTR;
  const float2* vmine = vst + (active ? cs : 0) * RL::WS;
  float2 acc[RL::CPL][RL::NR][P], pre[RL::CPL][RL::NR][P];
#pragma unroll
  for (int u = 0; u < RL::CPL; ++u)
#pragma unroll
    for (int q = 0; q < RL::NR; ++q)
#pragma unroll
      for (int i = 0; i < P; ++i) acc[u][q][i] = make_float2(0.f, 0.f);

  int lo = 0;   // lowest plane not yet stored to scratch (the ring holds lo .. lo+P-1)
  int pub = 0;  // lane l: run word of stage sample l
  for (int s0 = 0; s0 < cnt; s0 += CH) {
    // publish the gathered chunk; a run ends where the next sample's corner differs,
    // and at the end of the stage
    {
      int plane, rest;
      if constexpr (D == 1) {
        plane = sr.base;
        rest = 0;
      } else {
        plane = sr.base / RL::PP;
        rest = RL::plane_off(sr.base - plane * RL::PP);
      }
      const int key = rest | (plane << 16);
      const int nkey = __shfl_down(key, 1);
      const bool last = lane + 1 >= CH || s0 + lane + 1 >= cnt || nkey != key;
      pub = key | (last ? (1 << 30) : 0);
    }
    if (lane < CH) {
      float* rr = rec + lane * RL::RECW;
#pragma unroll
      for (int i = 0; i < P; ++i) rr[i] = sr.w[i];
#pragma unroll
      for (int e = 0; e < (D - 1) * P; ++e) wl[e * RL::WS + lane] = sr.w[P + e];
#pragma unroll
      for (int c = 0; c < RL::NC; ++c) vst[c * RL::WS + lane] = sr.v[c];
    }
    __syncthreads();
    // issue the next chunk's gathers; they land while this chunk is accumulated
    if (s0 + CH < cnt) {
      const int sl = s0 + CH + lane;
      stage_load<D, P, T>(sr, lane < CH && sl < cnt, gs0 + sl, onext, ksp, dcf, wsort, lbase, g, c0, nc);
      onext = (lane < CH && sl + CH < cnt) ? sample_slot(perm, g, gs0 + sl + CH) : 0;
    }
    const int n = min(CH, cnt - s0);
    {
      const int w0 = __builtin_amdgcn_readlane(pub, 0);
      const int b0 = word_plane(w0);
      if (b0 > lo) {
        retire_planes<D, P, T>(lo, b0, nc, lds2, dst);
        lo = b0;
      }
      read_run<D, P, T, PE>(w0, rok[0], roff, sgrid, pre);
    }
    // ring of LA register sets, loaded LA-2..LA-1 samples ahead; the stage arrays are
    // padded so loads past the last sample stay in bounds (their values are unused)
    SampleView<D, P, RL::NR, RL::CPL> ring[RL::LA];
#pragma unroll
    for (int u = 0; u < RL::LA - 1; ++u) read_view<D, P, T>(ring[u], u, rec, wl, vmine, ridx);
    for (int t = 0; t < n; t += RL::LA) {
#pragma unroll
      for (int u = 0; u < RL::LA; ++u) {
        if (t + u < n)
          accumulate_sample<D, P, T, PE>(ring[u], t + u, n, pub, rok, roff, sgrid, lds2, dst, nc, lo, acc, pre);
        read_view<D, P, T>(ring[(u + RL::LA - 1) % RL::LA], t + u + RL::LA - 1, rec, wl, vmine, ridx);
      }
    }
    __syncthreads();
  }
  // the remaining planes (held or never reached) are final
  retire_planes<D, P, T>(lo, SUB, nc, lds2, dst);
}


template <int D, int P, int T>
__global__ __launch_bounds__(64) void grid_partial_kernel(const float2* __restrict__ ksp,
                                                          const float* __restrict__ dcf,
                                                          const float* __restrict__ wsort,
                                                          const int* __restrict__ lbase,
                                                          const int* __restrict__ perm,
                                                          const Geom* __restrict__ geoms,
                                                          const Item* __restrict__ items, int nitems,
                                                          int nchunks, float2* __restrict__ scratch) {
  using RL = RowLanes<D, P, T>;
  extern __shared__ float2 lds2[];
  __shared__ float rec[RL::WS * RL::RECW];
  __shared__ float wl[RL::NWL * RL::WS];
  __shared__ float2 vst[RL::NC * RL::WS];
  int iidx, cc;
  item_of_block(blockIdx.x, nchunks, iidx, cc);
  if (iidx >= nitems) return;
  const Item it = items[iidx];
  if (it.w <= 0 || it.y < 0) return;  // non-finite samples contribute nothing
  const Geom g = geoms[it.x];
  const int c0 = cc * RL::NC;
  if (c0 >= g.nbatch) return;
  const int nc = min(RL::NC, g.nbatch - c0);
  // does a sample of the item reach the last patch plane (dim-0 weight P-1 nonzero)?
  // (one pass over those weights; the item's coil chunks run back to back on one XCD)
  int edge = 0;
  if constexpr (P >= 3) {
    const float* w0 = wsort + (g.samp_off + it.z) * (D * P);
    for (int s = threadIdx.x; s < it.w; s += 64) edge |= w0[(long long)s * (D * P) + P - 1] != 0.0f;
  }
  if (P < 3 || __any(edge))
    grid_partial_item<D, P, T, P>(ksp, dcf, wsort, lbase, perm, g, it, iidx, c0, nc, scratch, lds2, rec, wl, vst);
  else
    grid_partial_item<D, P, T, (P >= 3 ? P - 1 : P)>(ksp, dcf, wsort, lbase, perm, g, it, iidx, c0, nc, scratch,
                                                      lds2, rec, wl, vst);
}

}  // namespace mra

#include "grid_mfma.h"

namespace mra {

// Gridding phase 2. One block per (problem, output tile, group of RB batch entries):
// sums, in a fixed order, the scratch subgrids of every item whose subgrid covers the
// tile and writes each grid point once. blockIdx.x = (prob - prob0) * ntiles + tile,
// blockIdx.y = batch group; problems of one launch share the tile count. A thread owns
// NQ points of the tile for all RB batch entries, so each covering item gives it
// NQ * RB independent loads.
#ifndef MRA_REDUCE_BATCH
#define MRA_REDUCE_BATCH 4
#endif
constexpr int kReduceBatch = MRA_REDUCE_BATCH;
#ifndef MRA_REDUCE_AHEAD
#define MRA_REDUCE_AHEAD 4
#endif
constexpr int kReduceAhead = MRA_REDUCE_AHEAD;  // items whose loads are issued together (2D)
#ifndef MRA_REDUCE_AHEAD3
#define MRA_REDUCE_AHEAD3 2
#endif
constexpr int kReduceAhead3 = MRA_REDUCE_AHEAD3;  // further items of one tile issued together (3D)

// grid_reduce for 2D with every G >= SUB and even G[1]: the T x T tile has T*T/2 point
// pairs (y, x), (y, x + 1), x even; thread t owns pair t % (T*T/2) for batch entries
// 2h, 2h+1 of the block's four, h = t / (T*T/2), one 16-byte load per covering item each.
template <int P, int T>
__device__ __forceinline__ void reduce_pairs_2d(float2* __restrict__ grid, const float2* __restrict__ sbase,
                                                const Geom& g, const int (&tc)[2], const int (&cand)[2][3],
                                                const int (&ncand)[2], const int* __restrict__ titems, int b0,
                                                int nb, int accumulate) {
  using TS = TileShape<2, P, T>;
  constexpr int NPAIR = T * T / 2;
  const int t = threadIdx.x;
  const int pr = t % NPAIR, h = t / NPAIR;
  const int gc[2] = {tc[0] * T + pr / (T / 2), tc[1] * T + 2 * (pr % (T / 2))};
  const bool inside = gc[0] < g.G[0] && gc[1] < g.G[1];
  float4 acc[2] = {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
  if (ncand[0] <= 2 && ncand[1] <= 2) {
    // at most 2 x 2 covering tiles: item ranges first, then the first item of every tile
    // in one unrolled batch of loads, then any further items (as in reduce_pairs_3d)
    const float2* sp[4];
    int nxt[4], end[4], ex[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int k0 = c >> 1, k1 = c & 1;
      const bool valid = k0 < ncand[0] && k1 < ncand[1];
      const int c0 = cand[0][valid ? k0 : 0], c1 = cand[1][valid ? k1 : 0];
      const int l0 = pmod(gc[0] - c0 * T, g.G[0]);
      const int l1 = pmod(gc[1] - c1 * T, g.G[1]);  // even, so l1 + 1 < SUB when l1 < SUB
      const bool ok = valid && inside && l0 < TS::SUB && l1 < TS::SUB;
      ex[c] = l0 * TS::SUB + l1;
      const int stile = c0 * g.ntile[1] + c1;
      const int i0 = ok ? titems[stile] : 0, i1 = ok ? titems[stile + 1] : 0;
      sp[c] = i1 > i0 ? sbase + ((long long)(i0 - g.item_base) * g.nbatch + b0) * TS::SUBN + ex[c] : nullptr;
      nxt[c] = i0 + 1;
      end[c] = i1;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c)
      if (sp[c]) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (2 * h + u < nb) {
            const float4 v = *(const float4*)(sp[c] + (long long)(2 * h + u) * TS::SUBN);
            acc[u].x += v.x;
            acc[u].y += v.y;
            acc[u].z += v.z;
            acc[u].w += v.w;
          }
      }
    // further items (tiles split into several: the dense k-space centre) kReduceAhead at a
    // time, loads first, so that the hot tiles' blocks do not wait out one load per item;
    // the sums keep item order
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      int it = nxt[c];
      for (; it + kReduceAhead <= end[c]; it += kReduceAhead) {
        float4 v[kReduceAhead][2];
#pragma unroll
        for (int j = 0; j < kReduceAhead; ++j) {
          const float2* slot = sbase + ((long long)(it + j - g.item_base) * g.nbatch + b0) * TS::SUBN + ex[c];
#pragma unroll
          for (int u = 0; u < 2; ++u)
            v[j][u] = 2 * h + u < nb ? *(const float4*)(slot + (long long)(2 * h + u) * TS::SUBN)
                                     : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int j = 0; j < kReduceAhead; ++j)
#pragma unroll
          for (int u = 0; u < 2; ++u) {
            acc[u].x += v[j][u].x;
            acc[u].y += v[j][u].y;
            acc[u].z += v[j][u].z;
            acc[u].w += v[j][u].w;
          }
      }
      for (; it < end[c]; ++it) {
        const float2* slot = sbase + ((long long)(it - g.item_base) * g.nbatch + b0) * TS::SUBN + ex[c];
#pragma unroll
        for (int u = 0; u < 2; ++u)
          if (2 * h + u < nb) {
            const float4 v = *(const float4*)(slot + (long long)(2 * h + u) * TS::SUBN);
            acc[u].x += v.x;
            acc[u].y += v.y;
            acc[u].z += v.z;
            acc[u].w += v.w;
          }
      }
    }
  } else
  for (int k0 = 0; k0 < ncand[0]; ++k0)
    for (int k1 = 0; k1 < ncand[1]; ++k1) {
      const int c0 = cand[0][k0], c1 = cand[1][k1];
      const int l0 = pmod(gc[0] - c0 * T, g.G[0]);
      const int l1 = pmod(gc[1] - c1 * T, g.G[1]);  // even, so l1 + 1 < SUB when l1 < SUB
      const bool ok = inside && l0 < TS::SUB && l1 < TS::SUB;
      const int e = l0 * TS::SUB + l1;
      const int stile = c0 * g.ntile[1] + c1;
      const int i1 = titems[stile + 1];
      for (int it = titems[stile]; it < i1; ++it) {
        const float2* slot = sbase + ((long long)(it - g.item_base) * g.nbatch + b0) * TS::SUBN + e;
        if (ok) {
#pragma unroll
          for (int u = 0; u < 2; ++u)
            if (2 * h + u < nb) {
              const float4 v = *(const float4*)(slot + (long long)(2 * h + u) * TS::SUBN);
              acc[u].x += v.x;
              acc[u].y += v.y;
              acc[u].z += v.z;
              acc[u].w += v.w;
            }
        }
      }
    }
  if (!inside) return;
  const long long lin = (long long)gc[0] * g.G[1] + gc[1];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int b = 2 * h + u;
    if (b >= nb) break;
    float4* o = (float4*)(grid + g.grid_off + (long long)(b0 + b) * g.grid_stride + lin);
    if (accumulate) {
      const float4 v = *o;
      acc[u].x += v.x;
      acc[u].y += v.y;
      acc[u].z += v.z;
      acc[u].w += v.w;
    }
    *o = acc[u];
  }
}

// grid_reduce for 3D with every G >= SUB and even G[2]: thread t owns the points
// (z, y, x), (z, y, x + 1) of the output tile, x even, for kReduceBatch batch entries;
// each covering item's pair is one 16-byte load per entry.
template <int P, int T>
__device__ __forceinline__ void reduce_pairs_3d(float2* __restrict__ grid, const float2* __restrict__ sbase,
                                                const Geom& g, const int (&tc)[3], const int (&cand)[3][3],
                                                const int (&ncand)[3], const int* __restrict__ titems, int b0,
                                                int nb, int accumulate) {
  using TS = TileShape<3, P, T>;
  constexpr int RB = kReduceBatch;
  const int t = threadIdx.x;
  const int xp = t % (T / 2), rest = t / (T / 2);
  const int gc[3] = {tc[0] * T + rest / T, tc[1] * T + rest % T, tc[2] * T + 2 * xp};
  const bool inside = gc[0] < g.G[0] && gc[1] < g.G[1] && gc[2] < g.G[2];
  float4 acc[RB];
#pragma unroll
  for (int b = 0; b < RB; ++b) acc[b] = make_float4(0.f, 0.f, 0.f, 0.f);
  auto add = [&](const float4 (&v)[RB]) {
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      acc[b].x += v[b].x;
      acc[b].y += v[b].y;
      acc[b].z += v[b].z;
      acc[b].w += v[b].w;
    }
  };
  auto load = [&](int it, int e, float4 (&v)[RB]) {
    const float2* slot = sbase + ((long long)(it - g.item_base) * g.nbatch + b0) * TS::SUBN + e;
#pragma unroll
    for (int b = 0; b < RB; ++b)
      v[b] = b < nb ? *(const float4*)(slot + (long long)b * TS::SUBN) : make_float4(0.f, 0.f, 0.f, 0.f);
  };
  // Covering tiles in row-major order, each tile's items in order (the sums' order). The
  // first items of two neighbouring tiles along dim 2 are loaded together, further items
  // (split hot tiles) two at a time, so that a block does not wait out one load per item.
  for (int k0 = 0; k0 < ncand[0]; ++k0)
    for (int k1 = 0; k1 < ncand[1]; ++k1)
      for (int k2 = 0; k2 < ncand[2]; k2 += 2) {
        int e[2], i0[2], i1[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const bool valid = k2 + j < ncand[2];
          const int c0 = cand[0][k0], c1 = cand[1][k1], c2 = cand[2][valid ? k2 + j : k2];
          const int l0 = pmod(gc[0] - c0 * T, g.G[0]), l1 = pmod(gc[1] - c1 * T, g.G[1]);
          const int l2 = pmod(gc[2] - c2 * T, g.G[2]);  // even, so l2 + 1 < SUB when l2 < SUB
          const bool ok = valid && inside && l0 < TS::SUB && l1 < TS::SUB && l2 < TS::SUB;
          e[j] = (l0 * TS::SUB + l1) * TS::SUB + l2;
          const int stile = (c0 * g.ntile[1] + c1) * g.ntile[2] + c2;
          i0[j] = ok ? titems[stile] : 0;
          i1[j] = ok ? titems[stile + 1] : 0;
        }
        float4 first[2][RB];
#pragma unroll
        for (int j = 0; j < 2; ++j)
          if (i0[j] < i1[j]) load(i0[j], e[j], first[j]);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (i0[j] >= i1[j]) continue;
          add(first[j]);
          int it = i0[j] + 1;
          if constexpr (kReduceAhead3 >= 2)
            for (; it + 2 <= i1[j]; it += 2) {
              float4 v0[RB], v1[RB];
              load(it, e[j], v0);
              load(it + 1, e[j], v1);
              add(v0);
              add(v1);
            }
          for (; it < i1[j]; ++it) {
            float4 v0[RB];
            load(it, e[j], v0);
            add(v0);
          }
        }
      }
  if (!inside) return;
  const long long lin = ((long long)gc[0] * g.G[1] + gc[1]) * g.G[2] + gc[2];
#pragma unroll
  for (int b = 0; b < RB; ++b) {
    if (b >= nb) break;
    float4* o = (float4*)(grid + g.grid_off + (long long)(b0 + b) * g.grid_stride + lin);
    if (accumulate) {
      const float4 v = *o;
      acc[b].x += v.x;
      acc[b].y += v.y;
      acc[b].z += v.z;
      acc[b].w += v.w;
    }
    *o = acc[b];
  }
}

template <int D, int P, int T>
__global__ __launch_bounds__(kThreads) void grid_reduce_kernel(float2* __restrict__ grid,
                                                               const float2* __restrict__ scratch,
                                                               const Geom* __restrict__ geoms, int prob0,
                                                               int ntiles, const int* __restrict__ tile_items,
                                                               int accumulate) {
  using TS = TileShape<D, P, T>;
  constexpr int RB = kReduceBatch;
  constexpr int NQ = (TS::TN + kThreads - 1) / kThreads;
  const int prob = prob0 + blockIdx.x / ntiles;
  const int tile = blockIdx.x - (prob - prob0) * ntiles;
  const Geom g = geoms[prob];
  const int b0 = blockIdx.y * RB;
  if (b0 >= g.nbatch) return;
  const int nb = min(RB, g.nbatch - b0);
  int tc[D];
  tile_coords<D, T>(tile, g, tc);
  // Candidate source tiles per dim: t, t-1, t-2 (mod ntile), those whose subgrid
  // interval [org', org' + SUB) mod G can reach this tile.
  int cand[D][3], ncand[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    ncand[d] = 0;
    const int kmax = min(3, g.ntile[d]);
    for (int k = 0; k < kmax; ++k) {
      const int tp = pmod(tc[d] - k, g.ntile[d]);
      // [org', org' + SUB) mod G meets [org, org + T) iff the tile starts inside
      // the subgrid (dist < SUB) or wraps past org' + G (dist + T - 1 >= G).
      const int dist = pmod(tc[d] * T - tp * T, g.G[d]);
      if (dist < TS::SUB || dist + T - 1 >= g.G[d]) cand[d][ncand[d]++] = tp;
    }
  }
  // this thread's points
  int gc[NQ][D];
  bool inside[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int p = threadIdx.x + q * kThreads;
    int rem = p;
    inside[q] = p < TS::TN;
#pragma unroll
    for (int d = D - 1; d >= 0; --d) {
      gc[q][d] = tc[d] * T + rem % T;
      rem /= T;
      inside[q] = inside[q] && gc[q][d] < g.G[d];
    }
  }
  float2 acc[NQ][RB];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int b = 0; b < RB; ++b) acc[q][b] = make_float2(0.f, 0.f);

  const int* titems = tile_items + g.tile_items_off;
  const float2* sbase = scratch + g.scratch_off;
  const bool simple = g.G[0] >= TS::SUB && (D < 2 || g.G[1] >= TS::SUB) && (D < 3 || g.G[D - 1] >= TS::SUB);
  if constexpr (D == 2 && T % 2 == 0 && TS::SUB % 2 == 0 && TS::TN == kThreads && kReduceBatch == 4) {
    const bool pairs = simple && g.G[1] % 2 == 0 && g.scratch_off % 2 == 0 && g.grid_off % 2 == 0 &&
                       g.grid_stride % 2 == 0 && (uintptr_t)grid % 16 == 0 && (uintptr_t)scratch % 16 == 0;
    if (pairs) {
      reduce_pairs_2d<P, T>(grid, sbase, g, tc, cand, ncand, titems, b0, nb, accumulate);
      return;
    }
  }
  if constexpr (D == 3 && T % 2 == 0 && TS::SUB % 2 == 0 && TS::TN == 2 * kThreads) {
    // two adjacent points along the last dim per thread, 16-byte loads and stores; needs
    // even offsets so that a pair never straddles the wrap and stays 16-byte aligned
    const bool pairs = simple && g.G[2] % 2 == 0 && g.scratch_off % 2 == 0 && g.grid_off % 2 == 0 &&
                       g.grid_stride % 2 == 0 && (uintptr_t)grid % 16 == 0 && (uintptr_t)scratch % 16 == 0;
    if (pairs) {
      reduce_pairs_3d<P, T>(grid, sbase, g, tc, cand, ncand, titems, b0, nb, accumulate);
      return;
    }
  }
  // iterate over candidate tile combinations (row-major over dims; uniform per block)
  int k[D];
#pragma unroll
  for (int d = 0; d < D; ++d) k[d] = 0;
  while (true) {
    int stile = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) stile = stile * g.ntile[d] + cand[d][k[d]];
    int loc0[NQ][D];
    bool ok[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      ok[q] = inside[q];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        loc0[q][d] = pmod(gc[q][d] - cand[d][k[d]] * T, g.G[d]);
        ok[q] = ok[q] && loc0[q][d] < TS::SUB;
      }
    }
    const int i0 = titems[stile], i1 = titems[stile + 1];
    for (int it = i0; it < i1; ++it) {
      const float2* slot = sbase + ((long long)(it - g.item_base) * g.nbatch + b0) * TS::SUBN;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        if (!ok[q]) continue;
        if (simple) {  // every subgrid point lands on a distinct grid point
          int e = 0;
#pragma unroll
          for (int d = 0; d < D; ++d) e = e * TS::SUB + loc0[q][d];
#pragma unroll
          for (int b = 0; b < RB; ++b)
            if (b < nb) {
              const float2 v = slot[(long long)b * TS::SUBN + e];
              acc[q][b].x += v.x;
              acc[q][b].y += v.y;
            }
        } else {  // G < SUB somewhere: a subgrid can wrap onto a grid point more than once
          for (int b = 0; b < nb; ++b) {
            const float2* sl = slot + (long long)b * TS::SUBN;
            float ax = 0.f, ay = 0.f;
            if constexpr (D == 1) {
              for (int e0 = loc0[q][0]; e0 < TS::SUB; e0 += g.G[0]) {
                ax += sl[e0].x; ay += sl[e0].y;
              }
            } else if constexpr (D == 2) {
              for (int e0 = loc0[q][0]; e0 < TS::SUB; e0 += g.G[0])
                for (int e1 = loc0[q][1]; e1 < TS::SUB; e1 += g.G[1]) {
                  const float2 v = sl[e0 * TS::SUB + e1];
                  ax += v.x; ay += v.y;
                }
            } else {
              for (int e0 = loc0[q][0]; e0 < TS::SUB; e0 += g.G[0])
                for (int e1 = loc0[q][1]; e1 < TS::SUB; e1 += g.G[1])
                  for (int e2 = loc0[q][2]; e2 < TS::SUB; e2 += g.G[2]) {
                    const float2 v = sl[(e0 * TS::SUB + e1) * TS::SUB + e2];
                    ax += v.x; ay += v.y;
                  }
            }
#pragma unroll
            for (int bb = 0; bb < RB; ++bb)
              if (bb == b) {
                acc[q][bb].x += ax;
                acc[q][bb].y += ay;
              }
          }
        }
      }
    }
    // next combination
    int d = D - 1;
    while (d >= 0) {
      if (++k[d] < ncand[d]) break;
      k[d] = 0;
      --d;
    }
    if (d < 0) break;
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    if (!inside[q]) continue;
    long long lin = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) lin = lin * g.G[d] + gc[q][d];
#pragma unroll
    for (int b = 0; b < RB; ++b) {
      if (b >= nb) break;
      float2* o = grid + g.grid_off + (long long)(b0 + b) * g.grid_stride + lin;
      if (accumulate) {
        float2 v = *o;
        v.x += acc[q][b].x;
        v.y += acc[q][b].y;
        *o = v;
      } else {
        *o = acc[q][b];
      }
    }
  }
}

// --------------------------------------------------------------------------------------------
// Trajectory gradient of interpolation, reduced over batch entries:
//   gt[o, d] = sum_b Re( conj(a[b, o]) * sum_p dW_p/dt_d * grid[b, g_p] )
// One block per (item, 256*K-sample sub-slice); all coils in the block, the
// coil subgrid double-buffered in LDS; weights and derivative weights are
// computed on the fly from the sorted coordinates.
// Value gather and its per-dimension derivative gathers over the first PE points of every
// dimension: sd[d] = sum_p (prod_{e != d} w_e[p_e]) dw_d[p_d] v[p] (see interp_grad_kernel).
// With PE = P - 1 the skipped terms carry weight and derivative 0 (the last point lies
// outside the support), so the sums are unchanged.
template <int D, int P, int SUB, int PE>
__device__ __forceinline__ void grad_sums(const float2* sb, const float (&w)[D][P], const float (&dw)[D][P],
                                          float2 (&sd)[D]) {
  if constexpr (D == 3) {
#pragma unroll
    for (int i = 0; i < PE; ++i) {
      float2 cA = {0.f, 0.f}, cB = {0.f, 0.f}, cC = {0.f, 0.f};
#pragma unroll
      for (int j = 0; j < PE; ++j) {
        float2 rA = {0.f, 0.f}, rB = {0.f, 0.f};
#pragma unroll
        for (int l = 0; l < PE; ++l) {
          const float2 v = sb[(i * SUB + j) * SUB + l];
          rA.x += w[2][l] * v.x;  rA.y += w[2][l] * v.y;
          rB.x += dw[2][l] * v.x; rB.y += dw[2][l] * v.y;
        }
        cA.x += w[1][j] * rA.x;  cA.y += w[1][j] * rA.y;
        cB.x += dw[1][j] * rA.x; cB.y += dw[1][j] * rA.y;
        cC.x += w[1][j] * rB.x;  cC.y += w[1][j] * rB.y;
      }
      sd[0].x += dw[0][i] * cA.x; sd[0].y += dw[0][i] * cA.y;
      sd[1].x += w[0][i] * cB.x;  sd[1].y += w[0][i] * cB.y;
      sd[2].x += w[0][i] * cC.x;  sd[2].y += w[0][i] * cC.y;
    }
  } else if constexpr (D == 2) {
#pragma unroll
    for (int i = 0; i < PE; ++i) {
      float2 rA = {0.f, 0.f}, rB = {0.f, 0.f};
#pragma unroll
      for (int l = 0; l < PE; ++l) {
        const float2 v = sb[i * SUB + l];
        rA.x += w[1][l] * v.x;  rA.y += w[1][l] * v.y;
        rB.x += dw[1][l] * v.x; rB.y += dw[1][l] * v.y;
      }
      sd[0].x += dw[0][i] * rA.x; sd[0].y += dw[0][i] * rA.y;
      sd[1].x += w[0][i] * rB.x;  sd[1].y += w[0][i] * rB.y;
    }
  } else {
#pragma unroll
    for (int l = 0; l < PE; ++l) {
      const float2 v = sb[l];
      sd[0].x += dw[0][l] * v.x; sd[0].y += dw[0][l] * v.y;
    }
  }
}

template <int D, int P, int T, int K>
__global__ __launch_bounds__(kThreads) void interp_grad_kernel(const float2* __restrict__ grid,
                                                               const float2* __restrict__ a,
                                                               float* __restrict__ gt,
                                                               const float* __restrict__ tsort,
                                                               const int* __restrict__ perm,
                                                               const Geom* __restrict__ geoms,
                                                               const Item* __restrict__ items, int nitems,
                                                               int nsub) {
  using TS = TileShape<D, P, T>;
  extern __shared__ float2 lds2[];
  int iidx, sub;
  item_of_block(blockIdx.x, nsub, iidx, sub);
  if (iidx >= nitems) return;
  const Item it = items[iidx];
  const int s0 = it.z + sub * kThreads * K;
  const int s1 = min(it.z + it.w, s0 + kThreads * K);
  if (s0 >= s1) return;
  const Geom g = geoms[it.x];
  const int tid = threadIdx.x;
  if (it.y < 0) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const int s = s0 + tid + k * kThreads;
      if (s < s1) {
        const int o = perm[g.samp_off + s];
#pragma unroll
        for (int d = 0; d < D; ++d) gt[(g.samp_off + o) * D + d] = 0.f;
      }
    }
    return;
  }

  int tc[D];
  tile_coords<D, T>(it.y, g, tc);
  int goff[TS::NQ];
  subgrid_offsets<D, P, T>(g, tc, goff);

  constexpr int SUB = T + P - 1;
  float w[K][D][P], dw[K][D][P];
  int base[K], oidx[K], aidx[K];  // output slot (sample order); slot of the coefficients a
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int s = s0 + tid + k * kThreads;
    if (s < s1) {
      int bs = 0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const float t = tsort[(g.samp_off + s) * D + d];
        const float sc = d == 0 ? g.out_scale : 1.0f;
        const int lower = kb_weights<P>(t, g.r, g.beta, sc, w[k][d]);
        kb_dweights<P>(t, g.r, g.beta, sc, lower, dw[k][d]);
        bs = bs * SUB + pmod(lower + g.wrap[d], g.G[d]) % T;
      }
      base[k] = bs;
      oidx[k] = perm[g.samp_off + s];
      aidx[k] = g.ksp_plan_order ? s : oidx[k];
    } else {
#pragma unroll
      for (int d = 0; d < D; ++d)
#pragma unroll
        for (int j = 0; j < P; ++j) w[k][d][j] = dw[k][d][j] = 0.f;
      base[k] = 0;
      oidx[k] = -1;
      aidx[k] = 0;
    }
  }

  float acc[K][D];
  bool edge = false;  // some sample reaches the last patch point of a dimension
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int d = 0; d < D; ++d) {
      acc[k][d] = 0.f;
      edge = edge || w[k][d][P - 1] != 0.0f;
    }
  const bool full = __any(edge);  // wave-uniform

  const float2* gsrc = grid + g.grid_off;
  const float2* asrc = a + g.ksp_off;
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
      if (oidx[k] < 0) continue;
      const float2 av = asrc[(long long)b * g.ksp_stride + aidx[k]];
      const float2* sb = lds2 + cur + base[k];
      float2 sd[D];
#pragma unroll
      for (int d = 0; d < D; ++d) sd[d] = make_float2(0.f, 0.f);
      if constexpr (D == 3 && P >= 9) {  // two unrolled copies would not fit the registers
        grad_sums<D, P, SUB, P>(sb, w[k], dw[k], sd);
      } else {
        if (full)
          grad_sums<D, P, SUB, P>(sb, w[k], dw[k], sd);
        else
          grad_sums<D, P, SUB, P - 1>(sb, w[k], dw[k], sd);
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

constexpr int kGradK = 2;

// --------------------------------------------------------------------------------------------
enum class Op { Weights, Interp, GridPartial, GridReduce, Grad };

struct LaunchArgs {
  Op op;
  const void* src;      // grid (interp, grad, reduce: scratch) / ksp (grid_partial) / tsort (weights)
  void* dst;            // out ksp / scratch / grid / gt / wsort
  const void* aux;      // dcf (grid_partial) / a (grad)
  const float* wsort;
  const int* lbase;     // lbase (also written by Weights)
  const float* tsort;
  const int* perm;
  const Geom* geoms;    // device table (Weights: host copy in `geom`)
  Geom geom;
  const Item* items;
  int nitems;
  int ccoils, nchunks;  // coils per block, coil chunks
  int prob, nprob, ntiles, nbatch, accumulate;
  const int* tile_items;
  int npts;
  DimBetas betas;       // Weights: KB shape per dimension
  hipStream_t stream;
};


// Phase-1 gridding kernel: 1 = matrix cores (3D, P <= 7), 0 = per-lane LDS read-modify-write.
static int g_grid_kernel = 1;
// Interpolation reads coil pairs as 128-bit LDS values when the coil chunk is even.
#ifndef MRA_INTERP_PAIRS
#define MRA_INTERP_PAIRS 1
#endif
static int g_interp_pairs = MRA_INTERP_PAIRS;

template <int D, int P, int T>
static int launch_op(const LaunchArgs& a) {
  using TS = TileShape<D, P, T>;
  switch (a.op) {
    case Op::Weights:
      if (a.npts > 0)
        plan_weights_kernel<D, P, T><<<(a.npts + 255) / 256, 256, 0, a.stream>>>(
            (const float*)a.src, a.npts, a.geom, a.betas, (float*)a.dst, (int*)a.lbase);
      break;
    case Op::Interp: {
      if (a.nitems <= 0) break;
      if (g_interp_pairs && a.ccoils % 2 == 0) {
        const size_t sh = (size_t)a.ccoils * TS::SUBN * sizeof(float2);
        interp_pair_kernel<D, P, T><<<item_grid(a.nitems, a.nchunks), kThreads, sh, a.stream>>>(
            (const float2*)a.src, (float2*)a.dst, a.wsort, a.lbase, a.perm, a.geoms, a.items, a.nitems, a.ccoils,
            a.nchunks);
        break;
      }
      const size_t sh = (size_t)a.ccoils * TS::SUBN * sizeof(float2);
      interp_kernel<D, P, T><<<item_grid(a.nitems, a.nchunks), kThreads, sh, a.stream>>>(
          (const float2*)a.src, (float2*)a.dst, a.wsort, a.lbase, a.perm, a.geoms, a.items, a.nitems, a.ccoils,
          a.nchunks);
      break;
    }
    case Op::GridPartial: {
      if (a.nitems <= 0) break;
      if constexpr (D == 3 && P >= 2 && P <= 7) {
        if (g_grid_kernel == 1) {
          // A batch that fits one wave's coils runs one wave per block: a second wave
          // would have no coils but hold half the block's LDS, halving residency.
          if (a.nbatch <= MfmaGrid<P, 1>::NC) {
            grid_mfma_kernel<P, 1><<<item_grid(a.nitems, 1), 64, 0, a.stream>>>(
                (const float2*)a.src, (const float*)a.aux, a.wsort, a.lbase, a.perm, a.geoms, a.items, a.nitems, 1,
                (float2*)a.dst);
            break;
          }
          using M = MfmaGrid<P>;
          const int ng = (a.nbatch + M::NC - 1) / M::NC;
          grid_mfma_kernel<P, M::NW><<<item_grid(a.nitems, ng), 64 * M::NW, 0, a.stream>>>(
              (const float2*)a.src, (const float*)a.aux, a.wsort, a.lbase, a.perm, a.geoms, a.items, a.nitems, ng,
              (float2*)a.dst);
          break;
        }
      }
      using RL = RowLanes<D, P, T>;
      const int nch = (a.nbatch + RL::NC - 1) / RL::NC;
      const size_t sh = (size_t)RL::LDS_FLOAT2 * sizeof(float2);
      grid_partial_kernel<D, P, T><<<item_grid(a.nitems, nch), 64, sh, a.stream>>>(
          (const float2*)a.src, (const float*)a.aux, a.wsort, a.lbase, a.perm, a.geoms, a.items, a.nitems, nch,
          (float2*)a.dst);
      break;
    }
    case Op::GridReduce: {
      if (a.ntiles <= 0 || a.nbatch <= 0 || a.nprob <= 0) break;
      const int ngroups = (a.nbatch + kReduceBatch - 1) / kReduceBatch;
      if (ngroups > 65535) {
        set_last_error("grid_reduce: batch too large");
        return 1;
      }
      dim3 grd((unsigned)((long long)a.ntiles * a.nprob), (unsigned)ngroups);
      grid_reduce_kernel<D, P, T><<<grd, kThreads, 0, a.stream>>>((float2*)a.dst, (const float2*)a.src, a.geoms,
                                                                  a.prob, a.ntiles, a.tile_items, a.accumulate);
      break;
    }
    case Op::Grad: {
      if (a.nitems <= 0) break;
      const int nsub = a.nchunks;  // sub-slices of kThreads*kGradK samples per item
      interp_grad_kernel<D, P, T, kGradK><<<item_grid(a.nitems, nsub), kThreads,
                                            2 * TS::SUBN * sizeof(float2), a.stream>>>(
          (const float2*)a.src, (const float2*)a.aux, (float*)a.dst, a.tsort, a.perm, a.geoms, a.items, a.nitems,
          nsub);
      break;
    }
  }
  MRA_CHECK_LAUNCH("nufft kernel");
  return 0;
}

template <int D, int T>
static int dispatch_p(int P, const LaunchArgs& a) {
  switch (P) {
    case 2: return launch_op<D, 2, T>(a);
    case 3: return launch_op<D, 3, T>(a);
    case 4: return launch_op<D, 4, T>(a);
    case 5: return launch_op<D, 5, T>(a);
    case 6: return launch_op<D, 6, T>(a);
    case 7: return launch_op<D, 7, T>(a);
    case 8: return launch_op<D, 8, T>(a);
    case 9: return launch_op<D, 9, T>(a);
    default:
      set_last_error("kernel width unsupported: floor(width)+1 must be in 2..9");
      return 1;
  }
}

static int dispatch(int ndim, int P, const LaunchArgs& a) {
  switch (ndim) {
    case 1: return dispatch_p<1, TileCfg<1>::T>(P, a);
    case 2:
      switch (g_tile[2]) {
        case 8: return dispatch_p<2, 8>(P, a);
        case 32: return dispatch_p<2, 32>(P, a);
        default: return dispatch_p<2, 16>(P, a);
      }
    case 3: return dispatch_p<3, TileCfg<3>::T>(P, a);
    default:
      set_last_error("ndim must be 1, 2 or 3");
      return 1;
  }
}

static int subgrid_points(int ndim, int P) {
  const int T = ndim >= 1 && ndim <= 3 ? g_tile[ndim] : 1;
  return ipow(T + P - 1, ndim);
}

// ---- gridding without device allocation ----------------------------------------------------
// Callers that cannot allocate device memory (code under HIP graph capture, serving loops)
// grid through mra_grid_noalloc, whose scratch is this file-scope device array: a Geom slot
// followed by the per-(work item, batch entry) subgrids of phase 1. Every GPU the library
// runs on holds its own copy until the process ends. Host code never passes the array's
// name to a kernel: it uses the address hipGetSymbolAddress returns on each device.
constexpr size_t kNoAllocBytes = size_t(64) << 20;  // 64 MiB per GPU
constexpr size_t kNoAllocGeomBytes = 256;           // Geom slot at the start (keeps subgrids 256-byte aligned)
static_assert(sizeof(Geom) <= kNoAllocGeomBytes, "Geom slot");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kNoAllocMaxDevices];

// Device address of `scratch` on the current device (looked up once per device).
static int noalloc_scratch(void** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) {
    set_last_error("grid_noalloc: no current HIP device");
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      char b[256];
      snprintf(b, sizeof(b), "grid_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      set_last_error(b);
      return 1;
    }
    g_noalloc_ptr[dev] = p;
  }
  *out = g_noalloc_ptr[dev];
  return 0;
}

__global__ void store_geom_kernel(Geom g, Geom* __restrict__ dst) { *dst = g; }

}  // namespace mra

using namespace mra;

extern "C" {

long long mra_grid_noalloc_max_bytes() { return (long long)(kNoAllocBytes - kNoAllocGeomBytes); }

// Gridding (phase 1 + reduce, the kernels of mra_grid_partial / mra_grid_reduce) of one
// problem with the library's static scratch instead of a caller buffer. `geom_host` is the
// problem's host Geom record (mra_geom_fill); it is copied into the scratch's Geom slot by a
// kernel, so the call makes no allocation and no host-to-device copy.
int mra_grid_noalloc(int ndim, int P, const void* ksp, const float* dcf, const float* wsort, const int* lbase,
                     const int* perm, const void* geom_host, const void* items, int nitems, int nbatch,
                     const int* tile_items, int ntiles, void* grid, int accumulate, void* stream) {
  MRA_TRY_BEGIN
  if (!geom_host) MRA_FAIL("grid_noalloc: null Geom record");
  const Geom g = *(const Geom*)geom_host;
  if (ndim < 1 || ndim > kMaxDim) MRA_FAIL("grid_noalloc: ndim must be 1..3");
  if (nbatch < 1 || g.nbatch != nbatch) MRA_FAIL("grid_noalloc: nbatch must be positive and match the Geom record");
  if (g.item_base != 0 || g.scratch_off != 0 || g.tile_items_off != 0)
    MRA_FAIL("grid_noalloc: one problem per call (item_base, scratch_off and tile_items_off must be 0)");
  long long nt = 1;
  for (int d = 0; d < ndim; ++d) nt *= g.ntile[d];
  if (ntiles != nt) MRA_FAIL("grid_noalloc: ntiles does not match the Geom record");
  if (nitems < 0) MRA_FAIL("grid_noalloc: nitems must be non-negative");
  const long long need = (long long)nitems * nbatch * subgrid_points(ndim, P) * (long long)sizeof(float2);
  if (need > mra_grid_noalloc_max_bytes()) {
    char b[256];
    snprintf(b, sizeof(b), "grid_noalloc: needs %lld bytes of scratch (items x batch x subgrid), the static "
             "scratch holds %lld; use mra_grid_partial with a caller buffer", need, mra_grid_noalloc_max_bytes());
    MRA_FAIL(b);
  }
  void* base = nullptr;
  if (noalloc_scratch(&base)) return 1;
  Geom* gslot = (Geom*)base;
  float2* sub = (float2*)((unsigned char*)base + kNoAllocGeomBytes);
  const hipStream_t st = (hipStream_t)stream;
  store_geom_kernel<<<1, 1, 0, st>>>(g, gslot);
  MRA_CHECK_LAUNCH("grid_noalloc: Geom slot");
  LaunchArgs a{};
  a.op = Op::GridPartial;
  a.src = ksp;
  a.aux = dcf;
  a.dst = sub;
  a.wsort = wsort;
  a.lbase = lbase;
  a.perm = perm;
  a.geoms = gslot;
  a.items = (const Item*)items;
  a.nitems = nitems;
  a.nbatch = nbatch;
  a.stream = st;
  if (const int rc = dispatch(ndim, P, a)) return rc;
  LaunchArgs r{};
  r.op = Op::GridReduce;
  r.src = sub;
  r.dst = grid;
  r.geoms = gslot;
  r.prob = 0;
  r.nprob = 1;
  r.tile_items = tile_items;
  r.ntiles = ntiles;
  r.nbatch = nbatch;
  r.accumulate = accumulate;
  r.stream = st;
  return dispatch(ndim, P, r);
  MRA_TRY_END
}

int mra_geom_size() { return (int)sizeof(Geom); }

// Select the phase-1 gridding kernel (1 = matrix cores where compiled, 0 = per-lane LDS
// read-modify-write); returns the previous selection. Both write the same scratch.
int mra_set_grid_kernel(int mode) {
  const int old = g_grid_kernel;
  g_grid_kernel = mode ? 1 : 0;
  return old;
}

int mra_default_tile(int ndim) {
  switch (ndim) {
    case 1: return TileCfg<1>::T;
    case 2: return TileCfg<2>::T;
    case 3: return TileCfg<3>::T;
    default: return 0;
  }
}

int mra_tile(int ndim) { return ndim >= 1 && ndim <= 3 ? g_tile[ndim] : 0; }

int mra_set_tile(int ndim, int tile) {
  if (ndim < 1 || ndim > 3 || !tile_compiled(ndim, tile)) {
    set_last_error("set_tile: no kernels compiled for this tile edge (2D: 8, 16, 32; 1D: 256; 3D: 8)");
    return -1;
  }
  const int old = g_tile[ndim];
  g_tile[ndim] = tile;
  return old;
}

int mra_grad_samples_per_block() { return kThreads * kGradK; }

// Subgrid points per tile (scratch slot size of gridding) for (ndim, P).
long long mra_subgrid_points(int ndim, int P) { return subgrid_points(ndim, P); }

// Largest coil chunk whose subgrids fit the per-block LDS budget.
int mra_max_coil_chunk(int ndim, int P) {
  const int n = subgrid_points(ndim, P);
  int c = kLdsBudget / (n * (int)sizeof(float2));
  // A budget that holds one subgrid still takes two when they fit in 48 KB, so that
  // large kernels (3D, P = 7: 21.4 KB per coil) read coil pairs.
  if (c == 1 && 2 * n * (int)sizeof(float2) <= 48 * 1024) c = 2;
  return c < 1 ? 1 : c;
}

// Fill one Geom record (host memory) for a problem.
int mra_geom_fill(void* dst, long long grid_off, long long grid_stride, long long ksp_off, long long ksp_stride,
                  long long samp_off, int nbatch, int npts, int ndim, const int* G, int tile, const int* wrap,
                  float beta, float width, float out_scale, int item_base, long long scratch_off,
                  int tile_items_off) {
  MRA_TRY_BEGIN
  if (ndim < 1 || ndim > kMaxDim) MRA_FAIL("geom: ndim must be 1..3");
  if (tile != mra_tile(ndim)) MRA_FAIL("geom: tile edge differs from the one selected for this thread (mra_set_tile)");
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
  g.item_base = item_base;
  g.scratch_off = scratch_off;
  g.tile_items_off = tile_items_off;
  *(Geom*)dst = g;
  MRA_TRY_END
}

// Mark a host Geom record's dcf as given in sorted (plan) order: the gridding kernels then
// read it contiguously instead of through perm.
int mra_geom_set_dcf_plan_order(void* geom_host, int on) {
  MRA_TRY_BEGIN
  if (!geom_host) MRA_FAIL("geom: null record");
  ((Geom*)geom_host)->dcf_plan_order = on ? 1 : 0;
  MRA_TRY_END
}

int mra_geom_set_ksp_plan_order(void* geom_host, int on) {
  MRA_TRY_BEGIN
  if (!geom_host) MRA_FAIL("geom: null record");
  ((Geom*)geom_host)->ksp_plan_order = on ? 1 : 0;
  MRA_TRY_END
}

// Per sorted sample KB weights (npts x ndim x P floats) and subgrid offsets.
int mra_plan_weights_dims(int ndim, int P, const float* tsort, int npts, const int* G, const int* wrap, float width,
                          const float* betas, float* wsort, int* lbase, void* stream) {
  MRA_TRY_BEGIN
  if (ndim < 1 || ndim > 3) MRA_FAIL("plan_weights: ndim must be 1..3");
  if (!betas) MRA_FAIL("plan_weights: null betas");
  LaunchArgs a{};
  for (int d = 0; d < kMaxDim; ++d) a.betas.b[d] = betas[d < ndim ? d : 0];
  a.op = Op::Weights;
  a.src = tsort;
  a.dst = wsort;
  a.lbase = lbase;
  a.npts = npts;
  for (int d = 0; d < kMaxDim; ++d) {
    a.geom.G[d] = d < ndim ? G[d] : 1;
    a.geom.wrap[d] = (d < ndim && wrap) ? wrap[d] : 0;
  }
  a.geom.r = 0.5f * width;
  a.geom.beta = betas[0];
  a.stream = (hipStream_t)stream;
  return dispatch(ndim, P, a);
  MRA_TRY_END
}

int mra_plan_weights(int ndim, int P, const float* tsort, int npts, const int* G, const int* wrap, float width,
                     float beta, float* wsort, int* lbase, void* stream) {
  const float b[kMaxDim] = {beta, beta, beta};
  return mra_plan_weights_dims(ndim, P, tsort, npts, G, wrap, width, b, wsort, lbase, stream);
}

int mra_interp(int ndim, int P, const void* grid, void* out, const float* wsort, const int* lbase, const int* perm,
               const void* geoms, const void* items, int nitems, int ccoils, int nchunks, void* stream) {
  MRA_TRY_BEGIN
  if (ccoils < 1 || ccoils > mra_max_coil_chunk(ndim, P)) MRA_FAIL("interp: coil chunk exceeds the LDS budget");
  LaunchArgs a{};
  a.op = Op::Interp;
  a.src = grid;
  a.dst = out;
  a.wsort = wsort;
  a.lbase = lbase;
  a.perm = perm;
  a.geoms = (const Geom*)geoms;
  a.items = (const Item*)items;
  a.nitems = nitems;
  a.ccoils = ccoils;
  a.nchunks = nchunks;
  a.stream = (hipStream_t)stream;
  return dispatch(ndim, P, a);
  MRA_TRY_END
}

// nbatch: the largest batch (coil count) among the problems of the launch.
int mra_grid_partial(int ndim, int P, const void* ksp, const float* dcf, const float* wsort, const int* lbase,
                     const int* perm, const void* geoms, const void* items, int nitems, int nbatch,
                     void* scratch, void* stream) {
  MRA_TRY_BEGIN
  if (nbatch < 1) MRA_FAIL("grid: nbatch must be positive");
  LaunchArgs a{};
  a.op = Op::GridPartial;
  a.src = ksp;
  a.aux = dcf;
  a.dst = scratch;
  a.wsort = wsort;
  a.lbase = lbase;
  a.perm = perm;
  a.geoms = (const Geom*)geoms;
  a.items = (const Item*)items;
  a.nitems = nitems;
  a.nbatch = nbatch;
  a.stream = (hipStream_t)stream;
  return dispatch(ndim, P, a);
  MRA_TRY_END
}

// Problems prob0 .. prob0+nprob-1 (all with ntiles tiles); nbatch = largest batch among them.
int mra_grid_reduce(int ndim, int P, void* grid, const void* scratch, const void* geoms, int prob0, int nprob,
                    const int* tile_items, int ntiles, int nbatch, int accumulate, void* stream) {
  MRA_TRY_BEGIN
  LaunchArgs a{};
  a.op = Op::GridReduce;
  a.src = scratch;
  a.dst = grid;
  a.geoms = (const Geom*)geoms;
  a.prob = prob0;
  a.nprob = nprob;
  a.tile_items = tile_items;
  a.ntiles = ntiles;
  a.nbatch = nbatch;
  a.accumulate = accumulate;
  a.stream = (hipStream_t)stream;
  return dispatch(ndim, P, a);
  MRA_TRY_END
}

int mra_interp_grad(int ndim, int P, const void* grid, const void* a_, float* grad_t, const float* tsort,
                    const int* perm, const void* geoms, const void* items, int nitems, int max_item_samples,
                    void* stream) {
  MRA_TRY_BEGIN
  LaunchArgs a{};
  a.op = Op::Grad;
  a.src = grid;
  a.aux = a_;
  a.dst = grad_t;
  a.tsort = tsort;
  a.perm = perm;
  a.geoms = (const Geom*)geoms;
  a.items = (const Item*)items;
  a.nitems = nitems;
  a.nchunks = (max_item_samples + kThreads * kGradK - 1) / (kThreads * kGradK);
  a.stream = (hipStream_t)stream;
  return dispatch(ndim, P, a);
  MRA_TRY_END
}

}  // extern "C"
