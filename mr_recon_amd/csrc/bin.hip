// Sample binning: scale/wrap trajectory coordinates exactly as the reference
// does, assign every sample to the spatial tile that holds the lower corner of
// its kernel patch, and sort the samples by (tile, corner position in the tile),
// stably, so equal keys keep their sample order.
//
// Coordinate arithmetic follows, bit for bit:
//   triton_nufft._scale_trj  (fourier.py:420-430): t = fl32(fl32(k * fl32(G/N)) + G//2)
//   prep_ungrid_shapes       (ungrid.py:864-865):  t = torch.remainder(t, G)
//                            torch's fp32 remainder is fmod(a, b), plus b when the
//                            result is non-zero and of the opposite sign of b.
//   get_neighborhood         (kernels.py:531-536): lower = int(ceil(t - w/2))
// No FMA contraction is allowed in these steps: built with -ffp-contract=off.
//
// The sort is an LSD radix sort of stable counting-sort passes, each over a
// digit of at most kMaxDigit values: first the corner position in the tile
// (T^D values), then the tile id in base kMaxDigit (one pass up to kMaxDigit-1
// tiles, two beyond). Ordering samples of a tile by corner position makes
// neighbouring samples share subgrid addresses: interpolation lanes then hit
// the same or adjacent LDS banks, and gridding can sum runs of samples with
// one corner in registers.
#include "common.h"

namespace mra {

constexpr int kMaxDigit = 8192;  // LDS histogram entries per pass

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

// Position of a corner (a0, a1, a2) inside a 3D tile of edge T (a multiple of 4), ordered
// by cell (a0, a1 / 2, a2 / 4) first, then (a1 % 2, a2 % 4): the samples of one 1x2x4 cell
// of corners are contiguous and a0 never decreases along a tile.
__device__ __forceinline__ int cell_major_local(int a0, int a1, int a2, int T) {
  return (((a0 * (T / 2) + a1 / 2) * (T / 4) + a2 / 4) * 2 + a1 % 2) * 4 + a2 % 4;
}

// Scaled, wrapped coordinates of sample i, its tile id and corner position in the tile.
template <int D>
__device__ __forceinline__ int sample_key(const float* __restrict__ coords, long long i, const BinParams& bp,
                                          float (&t)[D], int& local) {
  // The rounding of every step must match separate fp32 operations; this file
  // is compiled with -ffp-contract=off (clang's __fmul_rn is an ordinary
  // multiply the backend would otherwise fuse with the add into an FMA).
  bool ok = true;
  int key = 0, loc = 0;
  int a[D];
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
    a[d] = lw % bp.T;
    loc = loc * bp.T + a[d];
  }
  if constexpr (D == 3) loc = cell_major_local(a[0], a[1], a[2], bp.T);
  local = ok ? loc : 0;
  return ok ? key : bp.ntiles;
}

template <int D>
__global__ __launch_bounds__(256) void bin_keys_kernel(const float* __restrict__ coords, long long npts,
                                                       BinParams bp, int* __restrict__ tkey,
                                                       int* __restrict__ lkey, float* __restrict__ tco) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  float t[D];
  int loc;
  tkey[i] = sample_key<D>(coords, i, bp, t, loc);
  lkey[i] = loc;
#pragma unroll
  for (int d = 0; d < D; ++d) tco[i * D + d] = t[d];
}

__device__ __forceinline__ int digit_of(int key, int div, int mod) { return (key / div) % mod; }

// Per-chunk histogram of one digit, written digit-major: hist[digit * nchunks + chunk].
__global__ __launch_bounds__(256) void pass_hist_kernel(const int* __restrict__ key, long long npts, int div,
                                                        int mod, int* __restrict__ hist, int nchunks,
                                                        int chunk_len) {
  extern __shared__ int sh[];
  for (int k = threadIdx.x; k < mod; k += blockDim.x) sh[k] = 0;
  __syncthreads();
  const long long c0 = (long long)blockIdx.x * chunk_len;
  const long long c1 = min(c0 + chunk_len, npts);
  for (long long i = c0 + threadIdx.x; i < c1; i += blockDim.x) atomicAdd(&sh[digit_of(key[i], div, mod)], 1);
  __syncthreads();
  for (int k = threadIdx.x; k < mod; k += blockDim.x) hist[(long long)k * nchunks + blockIdx.x] = sh[k];
}

// Stable scatter of one pass: one wave per chunk walks its positions in order, 64 at a
// time. Each lane finds the lanes holding its digit with one ballot per digit bit (the
// lanes whose digits agree on every bit), ranks itself among them by lane order, and the
// lowest of them advances the digit's cursor: equal digits keep their input order. Moves
// the sample index (perm) and the carried tile id.
__global__ __launch_bounds__(64) void pass_scatter_kernel(const int* __restrict__ key, long long npts, int div,
                                                          int mod, const int* __restrict__ scanned, int nchunks,
                                                          int chunk_len, const int* __restrict__ perm_in,
                                                          int* __restrict__ perm_out,
                                                          const int* __restrict__ carry_in,
                                                          int* __restrict__ carry_out) {
  extern __shared__ int cursor[];
  const int lane = threadIdx.x;
  for (int k = lane; k < mod; k += 64) cursor[k] = -1;
  __syncthreads();
  const long long c0 = (long long)blockIdx.x * chunk_len;
  const long long c1 = min(c0 + chunk_len, npts);
  const unsigned long long lt_mask = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int nbits = mod > 1 ? 32 - __clz(mod - 1) : 0;  // bits of the largest digit
  for (long long b = c0; b < c1; b += 64) {
    const long long i = b + lane;
    const bool valid = i < c1;
    const int dig = valid ? digit_of(key[i], div, mod) : 0;
    unsigned long long same = __ballot(valid);
    for (int bit = 0; bit < nbits; ++bit) {
      const bool set = ((dig >> bit) & 1) != 0;
      const unsigned long long bal = __ballot(set);
      same &= set ? bal : ~bal;
    }
    if (valid) {
      // one wave's LDS operations complete in order: every lane reads the cursor before the
      // lowest lane of its digit writes the advanced one
      int c = cursor[dig];
      if (c < 0) c = scanned[(long long)dig * nchunks + blockIdx.x];
      const long long pos = c + __popcll(same & lt_mask);
      perm_out[pos] = perm_in ? perm_in[i] : (int)i;
      carry_out[pos] = carry_in[i];
      if ((same & lt_mask) == 0ull) cursor[dig] = c + __popcll(same);
    }
  }
}

// tile_start[k] = first sorted position of tile k (nkeys + 1 entries, the last = npts),
// from the sorted tile ids; empty tiles get the start of the next non-empty one.
__global__ __launch_bounds__(256) void tile_start_kernel(const int* __restrict__ tkey, long long npts, int nkeys,
                                                         int* __restrict__ tile_start) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const int k = tkey[p];
  const int kp = p > 0 ? tkey[p - 1] : -1;
  for (int j = kp + 1; j <= k; ++j) tile_start[j] = (int)p;
  if (p == npts - 1)
    for (int j = k + 1; j <= nkeys; ++j) tile_start[j] = (int)npts;
}

template <int D>
__global__ __launch_bounds__(256) void gather_coords_kernel(const float* __restrict__ tco,
                                                            const int* __restrict__ perm, long long npts,
                                                            float* __restrict__ tsort) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npts) return;
  const long long o = perm[p];
#pragma unroll
  for (int d = 0; d < D; ++d) tsort[p * D + d] = tco[o * D + d];
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
__global__ void items_count_kernel(const int* __restrict__ tile_start, int nkeys, int S, int* __restrict__ nitems) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  nitems[k] = (tile_start[k + 1] - tile_start[k] + S - 1) / S;
}

__global__ void items_fill_kernel(Item* __restrict__ items, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) items[i] = make_int4(0, 0, 0, 0);
}

__global__ void items_write_kernel(const int* __restrict__ tile_start, int nkeys, int S,
                                   const int* __restrict__ item_off, int prob, Item* __restrict__ items) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  const int start = tile_start[k];
  const int end = tile_start[k + 1];
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

namespace {

struct BinLayout {
  int nchunks, npass, mods[3], divs[3];
  long long hist_elems;  // largest digit range x nchunks
  long long off_hist, off_scanned, off_scanws, off_lkey, off_tkey, off_ta, off_tb, off_pa, off_tco, total;
};

long long align64(long long n) { return (n + 63) / 64 * 64; }

int bin_layout(int ndim, long long npts, long long ntiles, int tile, int chunk_len, BinLayout& L) {
  if (ndim < 1 || ndim > kMaxDim) MRA_FAIL("bin: ndim must be 1..3");
  if (npts <= 0) MRA_FAIL("bin: npts must be positive");
  if (npts > 0x7fffffffLL) MRA_FAIL("bin: more than 2^31-1 samples");
  if (tile <= 0) MRA_FAIL("bin: tile must be positive");
  if (chunk_len <= 0) MRA_FAIL("bin: chunk_len must be positive");
  if (ntiles + 1 > 0x7fffffffLL / 2) MRA_FAIL("bin: too many tiles");
  long long tpd = 1;
  for (int d = 0; d < ndim; ++d) tpd *= tile;
  if (tpd > kMaxDigit) MRA_FAIL("bin: tile^ndim larger than 8192 positions");
  const long long nkeys = ntiles + 1;
  L.nchunks = (int)((npts + chunk_len - 1) / chunk_len);
  L.npass = 0;
  L.mods[L.npass] = (int)tpd;  // corner position in the tile
  L.divs[L.npass++] = 1;
  if (nkeys <= kMaxDigit) {
    L.mods[L.npass] = (int)nkeys;
    L.divs[L.npass++] = 1;
  } else {
    if (nkeys > (long long)kMaxDigit * kMaxDigit) MRA_FAIL("bin: too many tiles (limit 8192^2 - 1)");
    L.mods[L.npass] = kMaxDigit;
    L.divs[L.npass++] = 1;
    L.mods[L.npass] = (int)((nkeys + kMaxDigit - 1) / kMaxDigit);
    L.divs[L.npass++] = kMaxDigit;
  }
  int maxmod = 0;
  for (int p = 0; p < L.npass; ++p) maxmod = L.mods[p] > maxmod ? L.mods[p] : maxmod;
  L.hist_elems = (long long)maxmod * L.nchunks;
  long long o = 0;
  L.off_hist = o;    o += align64(L.hist_elems);
  L.off_scanned = o; o += align64(L.hist_elems + 1);
  L.off_scanws = o;  o += align64(mra_scan_workspace(L.hist_elems + 1));
  L.off_lkey = o;    o += align64(npts);
  L.off_tkey = o;    o += align64(npts);
  L.off_ta = o;      o += align64(npts);
  L.off_tb = o;      o += align64(npts);
  L.off_pa = o;      o += align64(npts);
  L.off_tco = o;     o += align64(npts * ndim);
  L.total = o;
  return 0;
}

}  // namespace

// Workspace (in 4-byte words) of mra_bin_build.
long long mra_bin_workspace(int ndim, long long npts, long long ntiles, int tile, int chunk_len) {
  BinLayout L{};
  if (bin_layout(ndim, npts, ntiles, tile, chunk_len, L)) return -1;
  return L.total;
}

int mra_bin_build(int ndim, const float* coords, long long npts, const int* G, const float* scale,
                  const float* shift, int apply_scale, float width, int tile, const int* wrap,
                  int chunk_len, void* work, float* tsort, int* perm, int* tile_start, void* stream) {
  MRA_TRY_BEGIN
  if (!(width > 0.0f)) MRA_FAIL("bin: width must be positive");
  if (ndim < 1 || ndim > kMaxDim) MRA_FAIL("bin: ndim must be 1..3");
  BinParams bp{};
  long long nt = 1;
  for (int d = 0; d < ndim; ++d) {
    if (G[d] <= 0) MRA_FAIL("bin: grid size must be positive");
    if (tile <= 0) MRA_FAIL("bin: tile must be positive");
    bp.s[d] = scale ? scale[d] : 1.0f;
    bp.h[d] = shift ? shift[d] : 0.0f;
    bp.G[d] = G[d];
    bp.wrap[d] = wrap ? wrap[d] : 0;
    bp.ntile[d] = (G[d] + tile - 1) / tile;
    nt *= bp.ntile[d];
  }
  BinLayout L{};
  if (bin_layout(ndim, npts, nt, tile, chunk_len, L)) return 1;
  bp.T = tile;
  bp.r = width * 0.5f;
  bp.apply_scale = apply_scale;
  bp.ntiles = (int)nt;
  int* w = (int*)work;
  int* hist = w + L.off_hist;
  int* scanned = w + L.off_scanned;
  int* scan_ws = w + L.off_scanws;
  int* lkey = w + L.off_lkey;
  int* tkey = w + L.off_tkey;
  int* tka = w + L.off_ta;
  int* tkb = w + L.off_tb;
  int* pa = w + L.off_pa;
  float* tco = (float*)(w + L.off_tco);
  hipStream_t st = (hipStream_t)stream;
  const unsigned nb = (unsigned)((npts + 255) / 256);
  switch (ndim) {
    case 1: bin_keys_kernel<1><<<nb, 256, 0, st>>>(coords, npts, bp, tkey, lkey, tco); break;
    case 2: bin_keys_kernel<2><<<nb, 256, 0, st>>>(coords, npts, bp, tkey, lkey, tco); break;
    default: bin_keys_kernel<3><<<nb, 256, 0, st>>>(coords, npts, bp, tkey, lkey, tco); break;
  }
  // pass p reads (key, perm, carry) and writes (perm', carry'); the last pass writes perm
  const int* kin = lkey;
  const int* pin = nullptr;
  const int* cin = tkey;
  int* couts[2] = {tka, tkb};
  for (int p = 0; p < L.npass; ++p) {
    const int mod = L.mods[p], div = L.divs[p];
    const size_t shb = (size_t)mod * sizeof(int);
    pass_hist_kernel<<<L.nchunks, 256, shb, st>>>(kin, npts, div, mod, hist, L.nchunks, chunk_len);
    if (mra_scan_excl_i32(hist, (long long)mod * L.nchunks, scanned, scan_ws, stream)) return 1;
    const bool last = p == L.npass - 1;
    // perm ping-pong: the output of the second-to-last pass must not alias perm
    int* pout = last ? perm : ((L.npass - 1 - p) % 2 == 1 ? pa : perm);
    int* cout = couts[p % 2];
    pass_scatter_kernel<<<L.nchunks, 64, shb, st>>>(kin, npts, div, mod, scanned, L.nchunks, chunk_len, pin, pout,
                                                    cin, cout);
    kin = cout;
    cin = cout;
    pin = pout;
  }
  tile_start_kernel<<<nb, 256, 0, st>>>(cin, npts, (int)(nt + 1), tile_start);
  switch (ndim) {
    case 1: gather_coords_kernel<1><<<nb, 256, 0, st>>>(tco, perm, npts, tsort); break;
    case 2: gather_coords_kernel<2><<<nb, 256, 0, st>>>(tco, perm, npts, tsort); break;
    default: gather_coords_kernel<3><<<nb, 256, 0, st>>>(tco, perm, npts, tsort); break;
  }
  MRA_CHECK_LAUNCH("bin_build");
  MRA_TRY_END
}

// Build work items of at most S samples for one problem from tile_start (ntiles + 2
// entries). items must hold max_items = ntiles + 1 + ceil(npts / S) entries; unused
// entries get count 0. nitems_ws and item_off need ntiles + 2 ints, scan_ws
// mra_scan_workspace(ntiles + 1).
int mra_bin_items(const int* tile_start, int ntiles, int S, int prob, int* nitems_ws, int* item_off, int* scan_ws,
                  Item* items, int max_items, void* stream) {
  MRA_TRY_BEGIN
  if (S <= 0) MRA_FAIL("items: S must be positive");
  hipStream_t st = (hipStream_t)stream;
  const int nkeys = ntiles + 1;
  items_fill_kernel<<<(max_items + 255) / 256, 256, 0, st>>>(items, max_items);
  items_count_kernel<<<(nkeys + 255) / 256, 256, 0, st>>>(tile_start, nkeys, S, nitems_ws);
  if (mra_scan_excl_i32(nitems_ws, nkeys, item_off, scan_ws, stream)) return 1;
  items_write_kernel<<<(nkeys + 255) / 256, 256, 0, st>>>(tile_start, nkeys, S, item_off, prob, items);
  MRA_CHECK_LAUNCH("bin_items");
  MRA_TRY_END
}

}  // extern "C"
