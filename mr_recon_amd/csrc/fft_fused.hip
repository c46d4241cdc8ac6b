// FFTs fused with the zero-pad / crop around them, skipping the known zeros.
//
// Forward (image -> oversampled grid spectrum): the padded grid is zero outside the
// image's N points per dim (pixel n sits at grid index (n - c) mod G). The transform
// along the last axis therefore only runs on rows whose leading indices hold pixels,
// reading the image (times SENSE map, apodization and scale) directly instead of a
// padded grid; each further axis transforms columns whose leading indices hold pixels
// and loads only the pixel rows of the transform axis. The final axis writes every
// grid point.
//
// Adjoint (grid spectrum -> image): the mirror image. The first axis transforms every
// column but stores only pixel rows; later axes load all rows of pixel columns and
// store pixel rows; the last axis transforms the pixel rows of all coils of an image
// and reduces them over coils (times conj(map), apodization and scale) straight into
// the image.
//
// Toeplitz normal operator: forward passes, then along axis 0 one kernel transforms,
// multiplies by the PSF spectrum and transforms back (storing pixel rows only), then
// the adjoint passes.
//
// Replaces, fused: the reference's resize + ifftshift + fftn + fftshift + crop chains of
// triton_nufft.forward_FT_only / adjoint_iFT_only (fourier.py:432-468, 505-525) and
// NUFFT.normal_toeplitz (fourier.py:302-352) with SENSE expand / reduce
// (linops.py:1161, 1220, 1265-1306).
#include "capi_util.h"
#include "fft_lds.h"

namespace mra {
namespace fused {

using namespace fft;

struct Geo3 {
  int N[3], G[3], c[3];   // image size, grid size, pixel at grid index 0 (leading unused = 1/1/0)
  const float* apod[3];   // per-dim apodization by pixel (nullptr = 1)
  float scale;
};

__device__ __forceinline__ int pix_of(int g, int N, int G, int c) {
  if (g < N - c) return g + c;
  if (g >= G - c) return g - G + c;
  return -1;
}
__device__ __forceinline__ int grid_of(int n, int G, int c) { return n - c < 0 ? n - c + G : n - c; }
__device__ __forceinline__ float apod_at(const Geo3& p, int d, int n) { return p.apod[d] ? p.apod[d][n] : 1.0f; }
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Forward, last axis: NT pixel rows per block (rows enumerate (q, n0, n1)); grid batch q
// uses image img_index[q] (or q / ns) and map q (or q % ns).
template <int L>
__global__ __launch_bounds__(kThreadsFFT) void pad_row_kernel(const float2* __restrict__ x, long long x_stride,
                                                               const float2* __restrict__ smaps, long long s_stride,
                                                               long long ns, const int* __restrict__ img_index,
                                                               Geo3 p, long long nrows, float2* __restrict__ grid,
                                                               const float2* __restrict__ twg) {
  using S = Shape<L, true>;
  __shared__ float2 buf[L * S::NTP];
  __shared__ float2 tw[L];
  // per row of the block: image / map offsets of its pixels, apodization, grid row offset
  __shared__ long long rsrc[S::NT], rmap[S::NT], rdst[S::NT];
  __shared__ float rapo[S::NT];
  const int tid = threadIdx.x;
  for (int m = tid; m < L; m += kThreadsFFT) tw[m] = twg[m];
  const long long row0 = (long long)blockIdx.x * S::NT;
  const int N0 = p.N[0], N1 = p.N[1], N2 = p.N[2];
  if (tid < S::NT) {
    const long long row = row0 + tid;
    if (row < nrows) {
      const int n1 = (int)(row % N1);
      const long long r2 = row / N1;
      const int n0 = (int)(r2 % N0);
      const long long q = r2 / N0;
      const long long bs = img_index ? q : q % ns;
      const long long bx = img_index ? img_index[q] : q / ns;
      const long long pix = ((long long)n0 * N1 + n1) * N2;
      rsrc[tid] = bx * x_stride + pix;
      rmap[tid] = bs * s_stride + pix;
      rapo[tid] = p.scale * apod_at(p, 0, n0) * apod_at(p, 1, n1);
      rdst[tid] = ((q * p.G[0] + grid_of(n0, p.G[0], p.c[0])) * p.G[1] + grid_of(n1, p.G[1], p.c[1])) * L;
    }
  }
  __syncthreads();
  const int nr = (int)min((long long)S::NT, nrows - row0);
  for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
    const int c = e / L, j = e - c * L;
    float2 v = make_float2(0.f, 0.f);
    const int n2 = pix_of(j, N2, L, p.c[2]);
    if (c < nr && n2 >= 0) {
      v = x[rsrc[c] + n2];
      if (smaps) v = cmul(v, smaps[rmap[c] + n2]);
      const float a = rapo[c] * apod_at(p, 2, n2);
      v.x *= a;
      v.y *= a;
    }
    buf[j * S::NTP + c] = v;
  }
  __syncthreads();
  fft_lds<L, S>(buf, tw);
  for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
    const int c = e / L, j = e - c * L;
    if (c < nr) grid[rdst[c] + j] = buf[j * S::NTP + c];
  }
}

enum ColMode { kFwd = 0, kAdj = 1, kToep = 2 };

// Column pass along axis a in {0, 1} (of the 3-dim view), NT columns per block. Outer
// index: q (axis 0) or (q, n0) over pixel planes (axis 1). kFwd: load pixel rows only,
// forward, store all rows. kAdj: load all rows, inverse, store pixel rows. kToep: load
// pixel rows, forward, multiply by kern (indexed over one grid), inverse, store pixel rows.
template <int L, int MODE>
__global__ __launch_bounds__(kThreadsFFT) void col_kernel(float2* __restrict__ grid, Geo3 p, int axis,
                                                           long long inner, const float2* __restrict__ twg,
                                                           const float2* __restrict__ kern) {
  using S = Shape<L, false>;
  __shared__ float2 buf[L * S::NTP];
  __shared__ float2 tw[L];
  const int tid = threadIdx.x;
  for (int m = tid; m < L; m += kThreadsFFT) tw[m] = twg[m];
  const long long nblk = (inner + S::NT - 1) / S::NT;
  const long long o = blockIdx.x / nblk;
  const long long i0 = (blockIdx.x - o * nblk) * S::NT;
  const int nvalid = (int)min((long long)S::NT, inner - i0);
  long long base;
  if (axis == 0) {
    base = o * (long long)L * inner;
  } else {  // axis 1: o = q * N0 + n0
    const long long q = o / p.N[0];
    const int n0 = (int)(o - q * p.N[0]);
    base = (q * p.G[0] + grid_of(n0, p.G[0], p.c[0])) * (long long)L * inner;
  }
  const int Na = p.N[axis], ca = p.c[axis];
  float2* col = grid + base + i0;
  for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
    const int j = e / S::NT, c = e - j * S::NT;
    float2 v = make_float2(0.f, 0.f);
    if (c < nvalid) {
      if (MODE == kAdj) {
        v = conjf2(col[(long long)j * inner + c]);
      } else if (pix_of(j, Na, L, ca) >= 0) {
        v = col[(long long)j * inner + c];
      }
    }
    buf[j * S::NTP + c] = v;
  }
  __syncthreads();
  fft_lds<L, S>(buf, tw);
  if (MODE == kToep) {
    const float2* kcol = kern + i0;
    for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
      const int j = e / S::NT, c = e - j * S::NT;
      if (c < nvalid) buf[j * S::NTP + c] = conjf2(cmul(buf[j * S::NTP + c], kcol[(long long)j * inner + c]));
    }
    __syncthreads();
    fft_lds<L, S>(buf, tw);
  }
  for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
    const int j = e / S::NT, c = e - j * S::NT;
    if (c < nvalid) {
      if (MODE == kFwd) {
        col[(long long)j * inner + c] = buf[j * S::NTP + c];
      } else if (pix_of(j, Na, L, ca) >= 0) {
        col[(long long)j * inner + c] = conjf2(buf[j * S::NTP + c]);
      }
    }
  }
}

constexpr int kCropMaxPix = 3;  // pixels per thread along the last axis (N2 <= 768)

// Adjoint, last axis + crop / coil reduce: one block per image row (bx, n0, n1). Grids
// [q0, q1) of image bx are inverse-transformed NT at a time and reduced:
// out[bx][n] (+)= scale * apod(n) * sum_q row_q[g(n)] * (conj) S[q][n].
template <int L>
__global__ __launch_bounds__(kThreadsFFT) void crop_row_kernel(const float2* __restrict__ grid, Geo3 p, long long ns,
                                                                const float2* __restrict__ smaps, long long s_stride,
                                                                const int* __restrict__ coil_off, int conj_maps,
                                                                int accumulate, float2* __restrict__ out,
                                                                long long x_stride, const float2* __restrict__ twg) {
  using S = Shape<L, true>;
  __shared__ float2 buf[L * S::NTP];
  __shared__ float2 tw[L];
  const int tid = threadIdx.x;
  for (int m = tid; m < L; m += kThreadsFFT) tw[m] = twg[m];
  const int N0 = p.N[0], N1 = p.N[1], N2 = p.N[2];
  const long long row = blockIdx.x;
  const int n1 = (int)(row % N1);
  const long long r2 = row / N1;
  const int n0 = (int)(r2 % N0);
  const long long bx = r2 / N0;
  const long long q0 = coil_off ? coil_off[bx] : bx * ns;
  const long long q1 = coil_off ? coil_off[bx + 1] : q0 + ns;
  const long long gsz = (long long)p.G[0] * p.G[1] * L;
  const long long goff = ((long long)grid_of(n0, p.G[0], p.c[0]) * p.G[1] + grid_of(n1, p.G[1], p.c[1])) * L;
  const long long pix0 = ((long long)n0 * N1 + n1) * N2;
  // Short rows (N2 <= 128) use all threads: ng groups of pw threads, group h reduces the
  // coils c = h, h+ng, ... of each chunk for one pixel per thread; the group sums meet in
  // LDS at the end. Longer rows: one group, up to kCropMaxPix pixels per thread.
  const int ng = N2 <= kThreadsFFT / 4 ? 4 : (N2 <= kThreadsFFT / 2 ? 2 : 1);
  const int pw = kThreadsFFT / ng;
  const int grp = tid / pw, tin = tid - grp * pw;
  float2 acc[kCropMaxPix];
#pragma unroll
  for (int k = 0; k < kCropMaxPix; ++k) acc[k] = make_float2(0.f, 0.f);
  for (long long qc = q0; qc < q1; qc += S::NT) {
    const int nr = (int)min((long long)S::NT, q1 - qc);
    for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
      const int c = e / L, j = e - c * L;
      buf[j * S::NTP + c] = c < nr ? conjf2(grid[(qc + c) * gsz + goff + j]) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    fft_lds<L, S>(buf, tw);
#pragma unroll
    for (int k = 0; k < kCropMaxPix; ++k) {
      const int n2 = tin + k * pw;
      if (n2 < N2) {
        const int g2 = grid_of(n2, L, p.c[2]);
        for (int c = grp; c < nr; c += ng) {
          float2 v = conjf2(buf[g2 * S::NTP + c]);
          if (smaps) {
            const float2 s = smaps[(qc + c - (coil_off ? 0 : bx * ns)) * s_stride + pix0 + n2];
            v = conj_maps ? cmul(v, conjf2(s)) : cmul(v, s);
          }
          acc[k].x += v.x;
          acc[k].y += v.y;
        }
      }
    }
    __syncthreads();
  }
  if (ng > 1) {  // one pixel per thread: fold the groups' sums into group 0
    if (grp > 0 && tin < N2) buf[grp * pw + tin] = acc[0];
    __syncthreads();
    if (grp == 0 && tin < N2)
      for (int h = 1; h < ng; ++h) {
        const float2 o = buf[h * pw + tin];
        acc[0].x += o.x;
        acc[0].y += o.y;
      }
  }
  const float a01 = p.scale * apod_at(p, 0, n0) * apod_at(p, 1, n1);
  float2* orow = out + bx * x_stride + pix0;
#pragma unroll
  for (int k = 0; k < kCropMaxPix; ++k) {
    const int n2 = tin + k * pw;
    if (grp == 0 && n2 < N2) {
      const float a = a01 * apod_at(p, 2, n2);
      float2 v = make_float2(acc[k].x * a, acc[k].y * a);
      if (accumulate) {
        const float2 o = orow[n2];
        v.x += o.x;
        v.y += o.y;
      }
      orow[n2] = v;
    }
  }
}

int fill_geo(Geo3& p, int ndim, const int* N, const int* G, const int* center, const float* apod, float scale) {
  if (ndim < 1 || ndim > 3) MRA_FAIL("fused fft: ndim must be 1..3");
  const int lead = 3 - ndim;
  int off = 0;
  for (int d = 0; d < 3; ++d) {
    if (d < lead) {
      p.N[d] = 1;
      p.G[d] = 1;
      p.c[d] = 0;
      p.apod[d] = nullptr;
    } else {
      const int n = N[d - lead], g = G[d - lead];
      if (n <= 0 || g < n) MRA_FAIL("fused fft: need 0 < N <= G in every dim");
      const int c = center ? center[d - lead] : n / 2;
      if (c < 0 || c > n) MRA_FAIL("fused fft: centre index out of range");
      p.N[d] = n;
      p.G[d] = g;
      p.c[d] = c;
      p.apod[d] = apod ? apod + off : nullptr;
      off += n;
    }
  }
  if (p.N[2] > kCropMaxPix * kThreadsFFT) MRA_FAIL("fused fft: last image dim above 768");
  p.scale = scale;
  return 0;
}

// inner extent of the column pass along axis a (of the 3-dim view)
long long inner_of(const Geo3& p, int a) {
  long long in = 1;
  for (int d = a + 1; d < 3; ++d) in *= p.G[d];
  return in;
}

template <int MODE>
int launch_col(int a, float2* grid, const Geo3& p, long long Q, const float2* tw, const float2* kern, hipStream_t st) {
  const long long inner = inner_of(p, a);
  const long long outer = a == 0 ? Q : Q * p.N[0];
  const int rc = with_length(p.G[a], [&](auto LL) {
    constexpr int L = decltype(LL)::value;
    using S = Shape<L, false>;
    if (inner < S::NT) MRA_FAIL("fused fft: inner extent below the column tile");
    const long long nb = outer * ((inner + S::NT - 1) / S::NT);
    if (nb > 0x7fffffffLL) MRA_FAIL("fused fft: grid too large");
    col_kernel<L, MODE><<<(unsigned)nb, kThreadsFFT, 0, st>>>(grid, p, a, inner, tw, kern);
    return 0;
  });
  if (rc < 0) MRA_FAIL("fused fft: length without an LDS plan");
  return rc;
}

int launch_pad_rows(const float2* x, long long x_stride, const float2* smaps, long long s_stride, long long ns,
                    const int* img_index, const Geo3& p, long long Q, float2* grid, const float2* tw, hipStream_t st) {
  const long long nrows = Q * p.N[0] * p.N[1];
  const int rc = with_length(p.G[2], [&](auto LL) {
    constexpr int L = decltype(LL)::value;
    using S = Shape<L, true>;
    const long long nb = (nrows + S::NT - 1) / S::NT;
    if (nb > 0x7fffffffLL) MRA_FAIL("fused fft: grid too large");
    pad_row_kernel<L><<<(unsigned)nb, kThreadsFFT, 0, st>>>(x, x_stride, smaps, s_stride, ns, img_index, p, nrows,
                                                            grid, tw);
    return 0;
  });
  if (rc < 0) MRA_FAIL("fused fft: length without an LDS plan");
  return rc;
}

int launch_crop_rows(const float2* grid, const Geo3& p, long long nx, long long ns, const float2* smaps,
                     long long s_stride, const int* coil_off, int conj_maps, int accumulate, float2* out,
                     long long x_stride, const float2* tw, hipStream_t st) {
  const long long nb = nx * p.N[0] * p.N[1];
  if (nb > 0x7fffffffLL) MRA_FAIL("fused fft: image too large");
  const int rc = with_length(p.G[2], [&](auto LL) {
    constexpr int L = decltype(LL)::value;
    crop_row_kernel<L><<<(unsigned)nb, kThreadsFFT, 0, st>>>(grid, p, ns, smaps, s_stride, coil_off, conj_maps,
                                                             accumulate, out, x_stride, tw);
    return 0;
  });
  if (rc < 0) MRA_FAIL("fused fft: length without an LDS plan");
  return rc;
}

}  // namespace fused
}  // namespace mra

using namespace mra;
using namespace mra::fused;

extern "C" {

// 1 when the fused transforms support grids of shape G (ndim dims): every length has a
// compiled radix plan and every axis but the last has >= 64 points after it.
int mra_fft_fused_supported(int ndim, const int* G) {
  if (ndim < 1 || ndim > 3) return 0;
  long long inner = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    if (!fft::plan_supported(G[d])) return 0;
    if (d < ndim - 1 && inner < 64) return 0;
    inner *= G[d];
  }
  return 1;
}

// Forward: grid[q] = FFT(pad(x[img(q)] * S[map(q)] * apod * scale)), q < nq, with
// img(q) = img_index[q] or q / ns, map(q) = q or q % ns. tw[d] as for mra_fft_lds.
int mra_fft_pad_forward(int ndim, const int* N, const int* G, const int* center, const void* x, long long x_stride,
                        const void* smaps, long long ns, long long s_stride, const int* img_index, long long nq,
                        const float* apod, float scale, void* grid, const void* const* tw, void* stream) {
  MRA_TRY_BEGIN
  if (!mra_fft_fused_supported(ndim, G)) MRA_FAIL("fused fft: grid shape not supported");
  if (nq < 1 || ns < 1) MRA_FAIL("fused fft: batch sizes must be positive");
  Geo3 p{};
  if (fill_geo(p, ndim, N, G, center, apod, scale)) return 1;
  const float2* const* t = (const float2* const*)tw;
  const int lead = 3 - ndim;
  hipStream_t st = (hipStream_t)stream;
  if (launch_pad_rows((const float2*)x, x_stride, (const float2*)smaps, s_stride, ns, img_index, p, nq,
                      (float2*)grid, t[ndim - 1], st))
    return 1;
  for (int a = 1; a >= lead; --a)
    if (launch_col<kFwd>(a, (float2*)grid, p, nq, t[a - lead], nullptr, st)) return 1;
  MRA_CHECK_LAUNCH("fft_pad_forward");
  MRA_TRY_END
}

// Adjoint: out[bx] (+)= scale * apod * sum_{q of bx} crop(iFFT(grid[q])) * (conj) S[q],
// grids of image bx = coil_off[bx] .. coil_off[bx+1]-1 or bx*ns .. bx*ns+ns-1; nq grids
// in all. The grids are overwritten.
int mra_fft_crop_adjoint(int ndim, const int* N, const int* G, const int* center, void* grid, long long nq,
                         long long nx, const void* smaps, long long ns, long long s_stride, const int* coil_off,
                         int conj_maps, const float* apod, float scale, int accumulate, void* out,
                         long long x_stride, const void* const* tw, void* stream) {
  MRA_TRY_BEGIN
  if (!mra_fft_fused_supported(ndim, G)) MRA_FAIL("fused fft: grid shape not supported");
  if (nq < 1 || nx < 1 || ns < 1) MRA_FAIL("fused fft: batch sizes must be positive");
  Geo3 p{};
  if (fill_geo(p, ndim, N, G, center, apod, scale)) return 1;
  const float2* const* t = (const float2* const*)tw;
  const int lead = 3 - ndim;
  hipStream_t st = (hipStream_t)stream;
  for (int a = lead; a <= 1; ++a)
    if (launch_col<kAdj>(a, (float2*)grid, p, nq, t[a - lead], nullptr, st)) return 1;
  if (launch_crop_rows((const float2*)grid, p, nx, ns, (const float2*)smaps, s_stride, coil_off, conj_maps, accumulate,
                       (float2*)out, x_stride, t[ndim - 1], st))
    return 1;
  MRA_CHECK_LAUNCH("fft_crop_adjoint");
  MRA_TRY_END
}

// Toeplitz normal operator of one image x and ns maps (2D / 3D):
// out (+)= sum_q conj(S[q]) crop(iFFT(K * FFT(pad(x * S[q])))), K the PSF spectrum in
// the unshifted layout on grid G (normalisation folded in). work holds ns grids.
int mra_toeplitz_fused(int ndim, const int* N, const int* G, const int* center, const void* x, const void* smaps,
                       long long ns, long long s_stride, const void* kern, void* work, int accumulate, void* out,
                       const void* const* tw, void* stream) {
  MRA_TRY_BEGIN
  if (ndim < 2) MRA_FAIL("fused Toeplitz: 2D or 3D only");
  if (!mra_fft_fused_supported(ndim, G)) MRA_FAIL("fused fft: grid shape not supported");
  if (ns < 1) MRA_FAIL("fused fft: batch sizes must be positive");
  Geo3 p{};
  if (fill_geo(p, ndim, N, G, center, nullptr, 1.0f)) return 1;
  const float2* const* t = (const float2* const*)tw;
  const int lead = 3 - ndim;
  hipStream_t st = (hipStream_t)stream;
  float2* g = (float2*)work;
  if (launch_pad_rows((const float2*)x, 0, (const float2*)smaps, s_stride, ns, nullptr, p, ns, g, t[ndim - 1], st))
    return 1;
  // forward along axis 1 (3D), then axis 0 forward * K inverse, then inverse along axis 1 (3D)
  if (ndim == 3 && launch_col<kFwd>(1, g, p, ns, t[1], nullptr, st)) return 1;
  if (launch_col<kToep>(lead, g, p, ns, t[0], (const float2*)kern, st)) return 1;
  if (ndim == 3 && launch_col<kAdj>(1, g, p, ns, t[1], nullptr, st)) return 1;
  if (launch_crop_rows(g, p, 1, ns, (const float2*)smaps, s_stride, nullptr, 1, accumulate, (float2*)out, 0,
                       t[ndim - 1], st))
    return 1;
  MRA_CHECK_LAUNCH("toeplitz_fused");
  MRA_TRY_END
}

}  // extern "C"
