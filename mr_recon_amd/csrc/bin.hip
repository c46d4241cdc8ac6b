// Sample binning: scale/wrap trajectory coordinates exactly as the reference
// does, assign every sample to the spatial tile that holds the lower corner of
// its kernel patch, and produce a stable (sample-index ordered) counting sort.
//
// Coordinate arithmetic follows, bit for bit:
//   triton_nufft._scale_trj  (fourier.py:420-430): t = fl32(fl32(k * fl32(G/N)) + G//2)
//   prep_ungrid_shapes       (ungrid.py:864-865):  t = torch.remainder(t, G)
//                            torch's fp32 remainder is fmod(a, b), plus b when the
//                            result is non-zero and of the opposite sign of b.
//   get_neighborhood         (kernels.py:531-536): lower = int(ceil(t - w/2))
// No FMA contraction is allowed in these steps: built with -ffp-contract=off.
#include "common.h"

namespace mra {

struct BinParams {
  float s[kMaxDim];       // per-dim scale (fp32), used when apply_scale != 0
  float h[kMaxDim];       // per-dim shift added after scaling
  int G[kMaxDim];
  int wrap[kMaxDim];
  int ntile[kMaxDim];
  int T;                  // tile edge (grid points)
  float r;                // half kernel width
  int apply_scale;
  int ntiles;             // product of ntile; key ntiles = non-finite sample
};

template <int D>
__device__ __forceinline__ int sample_key(const float* __restrict__ coords, long long i, const BinParams& bp,
                                          float (&t)[D]) {
  // The rounding of every step must match separate fp32 operations; this file
  // is compiled with -ffp-contract=off (clang's __fmul_rn is an ordinary
  // multiply the backend would otherwise fuse with the add into an FMA).
  bool ok = true;
  int key = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float v = coords[i * D + d];
    if (bp.apply_scale) {
      v = __fmul_rn(v, bp.s[d]);
      v = __fadd_rn(v, bp.h[d]);
    }
    const float Gf = (float)bp.G[d];
    float m = fmodf(v, Gf);
    if (m != 0.0f && m < 0.0f) m = __fadd_rn(m, Gf);
    t[d] = m;
    if (!isfinite(m)) ok = false;
    const int lower = ok ? (int)ceilf(__fsub_rn(m, bp.r)) : 0;
    const int lw = pmod(lower + bp.wrap[d], bp.G[d]);
    key = key * bp.ntile[d] + lw / bp.T;
  }
  return ok ? key : bp.ntiles;
}

// Per-chunk histogram of tile keys, written tile-major: hist[key * nchunks + chunk].
template <int D>
__global__ __launch_bounds__(256) void bin_hist_kernel(const float* __restrict__ coords, long long npts,
                                                       BinParams bp, int* __restrict__ hist, int nchunks,
                                                       int chunk_len) {
  extern __shared__ int sh[];
  const int nkeys = bp.ntiles + 1;
  for (int k = threadIdx.x; k < nkeys; k += blockDim.x) sh[k] = 0;
  __syncthreads();
  const long long c0 = (long long)blockIdx.x * chunk_len;
  const long long c1 = min(c0 + chunk_len, npts);
  for (long long i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
    float t[D];
    const int key = sample_key<D>(coords, i, bp, t);
    atomicAdd(&sh[key], 1);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nkeys; k += blockDim.x) hist[(long long)k * nchunks + blockIdx.x] = sh[k];
}

// Stable scatter: one wave per chunk walks its samples in index order, 64 at a
// time; lanes sharing a key are ranked by lane order, so equal keys keep their
// original order and the result equals a stable sort by key.
template <int D>
__global__ __launch_bounds__(64) void bin_scatter_kernel(const float* __restrict__ coords, long long npts,
                                                         BinParams bp, const int* __restrict__ scanned,
                                                         int nchunks, int chunk_len, float* __restrict__ tsort,
                                                         int* __restrict__ perm) {
  extern __shared__ int cursor[];
  const int lane = threadIdx.x;
  const int nkeys = bp.ntiles + 1;
  for (int k = lane; k < nkeys; k += 64) cursor[k] = -1;
  __syncthreads();
  const long long c0 = (long long)blockIdx.x * chunk_len;
  const long long c1 = min(c0 + chunk_len, npts);
  const unsigned long long lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (long long b = c0; b < c1; b += 64) {
    const long long i = b + lane;
    const bool valid = i < c1;
    float t[D];
    const int key = valid ? sample_key<D>(coords, i, bp, t) : -1;
    unsigned long long active = __ballot(valid);
    while (active) {
      const int leader = __ffsll((long long)active) - 1;
      const int lkey = __shfl(key, leader);
      const unsigned long long mask = __ballot(key == lkey) & active;
      int c = 0;
      if (lane == leader) {
        c = cursor[lkey];
        if (c < 0) c = scanned[(long long)lkey * nchunks + blockIdx.x];
        cursor[lkey] = c + __popcll(mask);
      }
      c = __shfl(c, leader);
      if ((mask >> lane) & 1ull) {
        const long long pos = c + __popcll(mask & lt_mask);
#pragma unroll
        for (int d = 0; d < D; ++d) tsort[pos * D + d] = t[d];
        perm[pos] = (int)i;
      }
      active &= ~mask;
    }
  }
}

// ---- exclusive scan of int32 (3 phases), out has n+1 entries (out[n] = total)
constexpr int kScanBlock = 256;
constexpr int kScanPer = 8;
constexpr int kScanTile = kScanBlock * kScanPer;

__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// Exclusive block scan of one value per thread; returns the block total via *total.
__device__ __forceinline__ int block_excl_scan(int v, int* sh_wave, int* total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int inc = wave_incl_scan(v);
  if (lane == 63) sh_wave[wid] = inc;
  __syncthreads();
  int wofs = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
    const int s = sh_wave[w];
    if (w < wid) wofs += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return wofs + inc - v;
}

__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const int* __restrict__ in, long long n,
                                                                int* __restrict__ bsum) {
  __shared__ int sh[kScanBlock / 64];
  const long long base = (long long)blockIdx.x * kScanTile;
  int s = 0;
  for (int q = 0; q < kScanPer; ++q) {
    const long long i = base + (long long)q * kScanBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  int tot;
  block_excl_scan(s, sh, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kScanBlock) void scan_bsum_kernel(int* __restrict__ bsum, int nb) {
  __shared__ int sh[kScanBlock / 64];
  int carry = 0;
  for (int b0 = 0; b0 < nb; b0 += kScanBlock) {
    const int i = b0 + threadIdx.x;
    const int v = i < nb ? bsum[i] : 0;
    int tot;
    const int ex = block_excl_scan(v, sh, &tot);
    if (i < nb) bsum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) bsum[nb] = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(const int* __restrict__ in, long long n,
                                                               const int* __restrict__ bsum, int nb,
                                                               int* __restrict__ out) {
  __shared__ int sh[kScanBlock / 64];
  const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanPer;
  int v[kScanPer];
  int s = 0;
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    const long long i = base + q;
    v[q] = i < n ? in[i] : 0;
    s += v[q];
  }
  int tot;
  int run = bsum[blockIdx.x] + block_excl_scan(s, sh, &tot);
#pragma unroll
  for (int q = 0; q < kScanPer; ++q) {
    const long long i = base + q;
    if (i < n) out[i] = run;
    run += v[q];
  }
  if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = bsum[nb];
}

// ---- work items
__global__ void items_count_kernel(const int* __restrict__ scanned, int nchunks, int nkeys, int S,
                                   int* __restrict__ nitems) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  const int start = scanned[(long long)k * nchunks];
  const int end = scanned[(long long)(k + 1) * nchunks];
  nitems[k] = (end - start + S - 1) / S;
}

__global__ void items_fill_kernel(Item* __restrict__ items, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) items[i] = make_int4(0, 0, 0, 0);
}

__global__ void items_write_kernel(const int* __restrict__ scanned, int nchunks, int nkeys, int S,
                                   const int* __restrict__ item_off, int prob, Item* __restrict__ items,
                                   int* __restrict__ tile_start) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  const int start = scanned[(long long)k * nchunks];
  const int end = scanned[(long long)(k + 1) * nchunks];
  if (tile_start) {
    tile_start[k] = start;
    if (k == nkeys - 1) tile_start[nkeys] = end;
  }
  const int tile = (k == nkeys - 1) ? -1 : k;
  int o = item_off[k];
  for (int s = start; s < end; s += S) items[o++] = make_int4(prob, tile, s, min(S, end - s));
}

}  // namespace mra

// ------------------------------------------------------------------------------------------
// C-ABI
#include "capi_util.h"

using namespace mra;

static int scan_blocks(long long n) { return (int)((n + kScanTile - 1) / kScanTile); }

extern "C" {

long long mra_scan_workspace(long long n) { return scan_blocks(n) + 1; }

int mra_scan_excl_i32(const int* in, long long n, int* out, int* ws, void* stream) {
  MRA_TRY_BEGIN
  if (n <= 0) MRA_FAIL("scan: n must be positive");
  hipStream_t st = (hipStream_t)stream;
  const int nb = scan_blocks(n);
  scan_reduce_kernel<<<nb, kScanBlock, 0, st>>>(in, n, ws);
  scan_bsum_kernel<<<1, kScanBlock, 0, st>>>(ws, nb);
  scan_apply_kernel<<<nb, kScanBlock, 0, st>>>(in, n, ws, nb, out);
  MRA_CHECK_LAUNCH("scan");
  MRA_TRY_END
}

// Binning workspace sizes for npts samples and ntiles tiles.
int mra_bin_sizes(long long npts, int ntiles, int chunk_len, int* nchunks, long long* hist_elems,
                  long long* scan_ws_elems) {
  MRA_TRY_BEGIN
  if (chunk_len <= 0) MRA_FAIL("bin: chunk_len must be positive");
  const long long nc = npts > 0 ? (npts + chunk_len - 1) / chunk_len : 1;
  *nchunks = (int)nc;
  *hist_elems = (long long)(ntiles + 1) * nc;
  *scan_ws_elems = mra_scan_workspace(*hist_elems + 1);
  MRA_TRY_END
}

int mra_bin_build(int ndim, const float* coords, long long npts, const int* G, const float* scale,
                  const float* shift, int apply_scale, float width, int tile, const int* wrap,
                  int chunk_len, int* hist, int* scanned, int* scan_ws, float* tsort, int* perm,
                  void* stream) {
  MRA_TRY_BEGIN
  if (ndim < 1 || ndim > kMaxDim) MRA_FAIL("bin: ndim must be 1..3");
  if (npts <= 0) MRA_FAIL("bin: npts must be positive");
  if (npts > 0x7fffffffLL) MRA_FAIL("bin: more than 2^31-1 samples");
  if (!(width > 0.0f)) MRA_FAIL("bin: width must be positive");
  if (tile <= 0) MRA_FAIL("bin: tile must be positive");
  BinParams bp{};
  long long nt = 1;
  for (int d = 0; d < ndim; ++d) {
    if (G[d] <= 0) MRA_FAIL("bin: grid size must be positive");
    bp.s[d] = scale ? scale[d] : 1.0f;
    bp.h[d] = shift ? shift[d] : 0.0f;
    bp.G[d] = G[d];
    bp.wrap[d] = wrap ? wrap[d] : 0;
    bp.ntile[d] = (G[d] + tile - 1) / tile;
    nt *= bp.ntile[d];
  }
  if (nt + 1 > 16384) MRA_FAIL("bin: too many tiles (limit 16383); use a larger tile");
  bp.T = tile;
  bp.r = width * 0.5f;
  bp.apply_scale = apply_scale;
  bp.ntiles = (int)nt;
  const int nchunks = (int)((npts + chunk_len - 1) / chunk_len);
  const size_t shbytes = (size_t)(nt + 1) * sizeof(int);
  const long long hist_n = (nt + 1) * (long long)nchunks;
  hipStream_t st = (hipStream_t)stream;
#define MRA_BIN_LAUNCH(DD)                                                                               \
  bin_hist_kernel<DD><<<nchunks, 256, shbytes, st>>>(coords, npts, bp, hist, nchunks, chunk_len);         \
  if (mra_scan_excl_i32(hist, hist_n, scanned, scan_ws, stream)) return 1;                                \
  bin_scatter_kernel<DD><<<nchunks, 64, shbytes, st>>>(coords, npts, bp, scanned, nchunks, chunk_len,     \
                                                       tsort, perm);
  switch (ndim) {
    case 1: MRA_BIN_LAUNCH(1) break;
    case 2: MRA_BIN_LAUNCH(2) break;
    default: MRA_BIN_LAUNCH(3) break;
  }
#undef MRA_BIN_LAUNCH
  MRA_CHECK_LAUNCH("bin_build");
  MRA_TRY_END
}

// Build work items of at most S samples for one problem. items must hold
// max_items = ntiles + 1 + ceil(npts / S) entries; unused entries get count 0.
// nitems_ws needs ntiles + 2 ints, item_off ntiles + 2, scan_ws mra_scan_workspace(ntiles + 1).
int mra_bin_items(const int* scanned, int nchunks, int ntiles, int S, int prob, int* nitems_ws,
                  int* item_off, int* scan_ws, Item* items, int max_items, int* tile_start, void* stream) {
  MRA_TRY_BEGIN
  if (S <= 0) MRA_FAIL("items: S must be positive");
  hipStream_t st = (hipStream_t)stream;
  const int nkeys = ntiles + 1;
  items_fill_kernel<<<(max_items + 255) / 256, 256, 0, st>>>(items, max_items);
  items_count_kernel<<<(nkeys + 255) / 256, 256, 0, st>>>(scanned, nchunks, nkeys, S, nitems_ws);
  if (mra_scan_excl_i32(nitems_ws, nkeys, item_off, scan_ws, stream)) return 1;
  items_write_kernel<<<(nkeys + 255) / 256, 256, 0, st>>>(scanned, nchunks, nkeys, S, item_off, prob, items,
                                                         tile_start);
  MRA_CHECK_LAUNCH("bin_items");
  MRA_TRY_END
}

}  // extern "C"
