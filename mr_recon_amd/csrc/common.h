// Shared device helpers for the MI355X (gfx950) NUFFT kernels.
//
// Layout conventions used by every kernel in this directory:
//   * complex64 data is float2 (re, im), the same bytes as torch.complex64.
//   * an oversampled grid for one batch entry (one coil) is a dense
//     row-major array of G[0]*...*G[D-1] float2, batch entries contiguous.
//   * k-space samples of one batch entry are npts float2, batch entries
//     separated by ksp_stride elements.
//   * per-sample coordinates are already scaled to grid units and wrapped to
//     [0, G] (the value G itself can occur, see bin.hip), stored sorted by
//     spatial tile in `tsort` (npts x D fp32) with `perm` mapping sorted
//     position -> original sample index.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mra {

constexpr int kMaxDim = 3;
constexpr int kThreads = 256;   // threads per work-item block (4 wave64)

// Per-problem geometry. A plain call uses a table with one entry; the batched
// path concatenates many problems and every work item names its problem.
struct Geom {
  long long grid_off;     // float2 offset of batch 0's grid in the grid base
  long long grid_stride;  // float2 elements between consecutive batch grids
  long long ksp_off;      // float2 offset of batch 0's samples in the ksp base
  long long ksp_stride;   // float2 elements between consecutive batch sample vectors
  long long samp_off;     // offset of this problem in tsort/perm (samples)
  int nbatch;             // number of grids (coils x image batch)
  int npts;               // samples per batch entry
  int G[kMaxDim];         // grid size per dim
  int ntile[kMaxDim];     // tiles per dim
  int wrap[kMaxDim];      // added to a grid index before it is wrapped mod G
  float beta;             // Kaiser-Bessel shape
  float r;                // half width (w/2)
  float out_scale;        // multiplies every weight (e.g. 1/w^D)
  int item_base;          // index of this problem's first work item
  long long scratch_off;  // float2 offset of this problem's gridding partials
  int tile_items_off;     // offset of this problem's per-tile item table
  int dcf_plan_order;     // 1: the gridding dcf is indexed by sorted (plan) position, 0: by sample
  int ksp_plan_order;     // 1: k-space (interp output, gridding input, the trajectory gradient's
                          //    coefficients) and dcf are in sorted order; the trajectory
                          //    gradient's output stays in sample order
};

// Position of sorted sample gs in the problem's sample-indexed vectors (k-space, dcf):
// perm[gs], or the sorted position itself when the caller keeps them in plan order.
__device__ __forceinline__ int sample_slot(const int* __restrict__ perm, const Geom& g, long long gs) {
  return g.ksp_plan_order ? (int)(gs - g.samp_off) : perm[gs];
}

// Work item: one tile's slice of sorted samples.
//   x = problem index, y = linear tile id (-1: non-finite samples),
//   z = first sorted sample (relative to the problem), w = sample count.
typedef int4 Item;

__host__ __device__ constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }

__device__ __forceinline__ int pmod(int a, int m) {
  int r = a % m;
  return r < 0 ? r + m : r;
}

// MI355X dispatches workgroups round-robin over its 8 XCDs (block b runs on XCD
// b % 8). Work items are ordered by tile, and trajectories concentrate samples in
// a few regions (the k-space centre), so contiguous item ranges per XCD would load
// some XCDs several times more than others. Items are dealt round-robin instead:
// XCD x takes items x, x+8, x+16, ...; the nsub blocks of one item (coil chunks or
// sample slices, which read the same sample data) run back to back on one XCD and
// share its L2. The grid holds 8 * nsub * ceil(nitems / 8) blocks.
__device__ __forceinline__ void item_of_block(int bid, int nsub, int& item, int& sub) {
  const int k = bid >> 3;
  const int q = k / nsub;
  item = (bid & 7) + 8 * q;
  sub = k - q * nsub;
}

inline unsigned item_grid(long long nitems, int nsub) {
  return (unsigned)(8LL * nsub * ((nitems + 7) / 8));
}

// Zeroth-order modified Bessel function I0(x), x >= 0, by the polynomial
// approximations of Abramowitz & Stegun 9.8.1 (x < 3.75) and 9.8.2 (x >= 3.75),
// the method named by the reference (triton_interp/kernels.py:180-246).
__device__ __forceinline__ float i0_as(float x) {
  if (x < 3.75f) {
    const float t = x / 3.75f;
    const float t2 = t * t;
    return 1.0f + t2 * (3.5156229f + t2 * (3.0899424f + t2 * (1.2067492f +
                 t2 * (0.2659732f + t2 * (0.0360768f + t2 * 0.0045813f)))));
  }
  const float u = 3.75f / x;   // 1/t
  const float poly = 0.39894228f + u * (0.01328592f + u * (0.00225319f + u * (-0.00157565f +
                     u * (0.00916281f + u * (-0.02057706f + u * (0.02635537f +
                     u * (-0.01647633f + u * 0.00392377f)))))));
  return expf(x) / sqrtf(x) * poly;
}

// d/du I0(beta*sqrt(1-u^2)) for u in [0,1], written so that it stays finite
// at u = 1: dz/du = -beta^2 u / z, and both branches of dI0/dz carry a factor
// z (small branch) or are evaluated away from z = 0 (large branch).
__device__ __forceinline__ float di0_du(float u, float beta) {
  const float z = beta * sqrtf(fmaxf(1.0f - u * u, 0.0f));
  if (z < 3.75f) {
    // I0 = sum a_i t^(2i), t = z/3.75  ->  dI0/dz = (z/3.75^2) sum 2i a_i t^(2i-2)
    const float t = z / 3.75f;
    const float t2 = t * t;
    const float s = 2.0f * 3.5156229f + t2 * (4.0f * 3.0899424f + t2 * (6.0f * 1.2067492f +
                    t2 * (8.0f * 0.2659732f + t2 * (10.0f * 0.0360768f + t2 * (12.0f * 0.0045813f)))));
    // dI0/dz * dz/du = (z s / 3.75^2) * (-beta^2 u / z)
    return -(beta * beta) * u * s * (1.0f / (3.75f * 3.75f));
  }
  // I0 = e^z z^-1/2 Q(v), v = 3.75/z
  const float v = 3.75f / z;
  const float q = 0.39894228f + v * (0.01328592f + v * (0.00225319f + v * (-0.00157565f +
                  v * (0.00916281f + v * (-0.02057706f + v * (0.02635537f +
                  v * (-0.01647633f + v * 0.00392377f)))))));
  const float dq = 0.01328592f + v * (2.0f * 0.00225319f + v * (3.0f * -0.00157565f +
                   v * (4.0f * 0.00916281f + v * (5.0f * -0.02057706f + v * (6.0f * 0.02635537f +
                   v * (7.0f * -0.01647633f + v * (8.0f * 0.00392377f)))))));
  const float e = expf(z) / sqrtf(z);
  // dI0/dz = e*(1 - 1/(2z))*q + e*dq*dv/dz,  dv/dz = -3.75/z^2
  const float di0dz = e * ((1.0f - 0.5f / z) * q - dq * v / z);
  return di0dz * (-(beta * beta) * u / z);
}

// Kaiser-Bessel weights along one dimension for a sample at grid coordinate t:
// grid points g = lower + j, lower = ceil(t - r) (fp32, as the reference:
// triton_interp/kernels.py:531-536 and ungrid.py:927-930), support |g - t| <= r
// (kernels.py:306-312). `scale` multiplies every weight.
template <int P>
__device__ __forceinline__ int kb_weights(float t, float r, float beta, float scale, float (&w)[P]) {
  const int lower = (int)ceilf(__fsub_rn(t, r));
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const float d = fabsf(__fsub_rn((float)(lower + j), t));
    w[j] = (d <= r) ? i0_as(beta * sqrtf(1.0f - (d / r) * (d / r))) * scale : 0.0f;
  }
  return lower;
}

// Same points, the derivative of each weight with respect to t.
template <int P>
__device__ __forceinline__ void kb_dweights(float t, float r, float beta, float scale, int lower, float (&dw)[P]) {
#pragma unroll
  for (int j = 0; j < P; ++j) {
    const float dd = __fsub_rn((float)(lower + j), t);
    const float d = fabsf(dd);
    // w = I0(beta sqrt(1-u^2)), u = |g - t|/r,  du/dt = -sign(g - t)/r
    const float sgn = dd > 0.0f ? 1.0f : (dd < 0.0f ? -1.0f : 0.0f);
    dw[j] = (d <= r) ? di0_du(d / r, beta) * (-sgn / r) * scale : 0.0f;
  }
}

}  // namespace mra
