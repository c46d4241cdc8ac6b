// Streaming kernels around the FFT: apodize + zero-pad (+ SENSE expand) into
// the oversampled grid, crop + deapodize (+ SENSE reduce) out of it, and the
// Toeplitz (PSF) multiply.
//
// The reference centres every FFT with ifftshift / fftshift copies
// (fourier.py:37-61) and zero-pads / crops with `resize` (utils.py:41-84). For
// an image of N points placed centred in G points and then ifftshift-ed, pixel
// n lands at grid index (n - N//2) mod G; the crop after fftshift reads the
// same index back. These kernels apply that placement directly, so the FFT in
// between runs unshifted and no shift or pad copy is ever materialised.
#include "common.h"
#include "capi_util.h"

namespace mra {

struct PadCropParams {
  int N[3];            // image size (leading unused dims = 1)
  int G[3];            // grid size (leading unused dims = 1)
  int c[3];            // pixel c sits at grid index 0 (pixel n at (n - c) mod G)
  long long nx;        // image batch entries
  long long ns;        // maps (coils); grid batch = nx * ns
  long long x_stride;  // float2 elements between image batch entries
  long long s_stride;  // float2 elements between maps
  float scale;
  int conj_maps;       // crop: multiply by conj(map)
  int accumulate;      // crop: out += result
  int roll_out;        // crop: write pixel n at (n - N//2) mod N (ifftshift layout)
  const float* apod[3];  // per-dim apodization (nullptr = 1), indexed by pixel
};

// Grid index -> pixel, inverse of n -> (n - c) mod G (0 <= c <= N <= G); -1 when
// g holds no pixel.
__device__ __forceinline__ int grid_to_pixel(int g, int N, int G, int c) {
  if (g < N - c) return g + c;
  if (g >= G - c) return g - G + c;
  return -1;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// One wave per output row (fixed batch entry and leading indices), lanes along
// the last dimension. out[(bx*ns + bs)][g] = x[bx][n] * S[bs][n] * apod(n) * scale.
__global__ __launch_bounds__(256) void pad_expand_kernel(const float2* __restrict__ x,
                                                         const float2* __restrict__ smaps,
                                                         float2* __restrict__ out, PadCropParams p,
                                                         long long nrows, const int* __restrict__ img_index) {
  const int lane = threadIdx.x & 63;
  const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nrows) return;
  const int G0 = p.G[0], G1 = p.G[1], G2 = p.G[2];
  const int N0 = p.N[0], N1 = p.N[1], N2 = p.N[2];
  const int g1 = (int)(row % G1);
  long long rr = row / G1;
  const int g0 = (int)(rr % G0);
  rr /= G0;  // grid batch index q = bx * ns + bs
  // batched problems: grid q belongs to image img_index[q] and uses map q
  const long long bs = img_index ? rr : rr % p.ns;
  const long long bx = img_index ? img_index[rr] : rr / p.ns;
  float2* orow = out + row * G2;
  const int n0 = grid_to_pixel(g0, N0, G0, p.c[0]), n1 = grid_to_pixel(g1, N1, G1, p.c[1]);
  if (n0 < 0 || n1 < 0) {
    for (int g2 = lane; g2 < G2; g2 += 64) orow[g2] = make_float2(0.f, 0.f);
    return;
  }
  const float a01 = p.scale * (p.apod[0] ? p.apod[0][n0] : 1.0f) * (p.apod[1] ? p.apod[1][n1] : 1.0f);
  const long long pix = ((long long)n0 * N1 + n1) * N2;
  const float2* xrow = x + bx * p.x_stride + pix;
  const float2* srow = smaps ? smaps + bs * p.s_stride + pix : nullptr;
  for (int g2 = lane; g2 < G2; g2 += 64) {
    const int n2 = grid_to_pixel(g2, N2, G2, p.c[2]);
    float2 v = make_float2(0.f, 0.f);
    if (n2 >= 0) {
      v = xrow[n2];
      if (srow) v = cmul(v, srow[n2]);
      const float a = a01 * (p.apod[2] ? p.apod[2][n2] : 1.0f);
      v.x *= a;
      v.y *= a;
    }
    orow[g2] = v;
  }
}

// One wave per output image row; lanes along the last dim; loop over maps.
// out[bx][n] (+)= scale * apod(n) * sum_bs grid[(bx*ns + bs)][g(n)] * conj(S[bs][n])
__global__ __launch_bounds__(256) void crop_reduce_kernel(const float2* __restrict__ grid,
                                                          const float2* __restrict__ smaps,
                                                          float2* __restrict__ out, PadCropParams p,
                                                          long long nrows, const int* __restrict__ coil_off) {
  const int lane = threadIdx.x & 63;
  const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= nrows) return;
  const int G0 = p.G[0], G1 = p.G[1], G2 = p.G[2];
  const int N0 = p.N[0], N1 = p.N[1], N2 = p.N[2];
  const int n1 = (int)(row % N1);
  long long rr = row / N1;
  const int n0 = (int)(rr % N0);
  const long long bx = rr / N0;
  const int g0 = pmod(n0 - p.c[0], G0), g1 = pmod(n1 - p.c[1], G1);
  const long long gtot = (long long)G0 * G1 * G2;
  // grids (and maps) [q0, q1) reduce into image bx; batched problems give the range per image
  const long long q0 = coil_off ? coil_off[bx] : bx * p.ns;
  const long long q1 = coil_off ? coil_off[bx + 1] : q0 + p.ns;
  const float2* grow = grid + q0 * gtot + ((long long)g0 * G1 + g1) * G2;
  const long long pix = ((long long)n0 * N1 + n1) * N2;
  const float a01 = p.scale * (p.apod[0] ? p.apod[0][n0] : 1.0f) * (p.apod[1] ? p.apod[1][n1] : 1.0f);
  long long orow;
  if (p.roll_out) {
    orow = ((long long)pmod(n0 - N0 / 2, N0) * N1 + pmod(n1 - N1 / 2, N1)) * N2;
  } else {
    orow = pix;
  }
  float2* obase = out + bx * p.x_stride + orow;
  for (int n2 = lane; n2 < N2; n2 += 64) {
    const int g2 = pmod(n2 - p.c[2], G2);
    float2 acc = make_float2(0.f, 0.f);
    for (long long bs = 0; bs < q1 - q0; ++bs) {
      float2 v = grow[bs * gtot + g2];
      if (smaps) {
        const float2 s = smaps[((coil_off ? q0 : 0) + bs) * p.s_stride + pix + n2];
        v = p.conj_maps ? cmulc(v, s) : cmul(v, s);
      }
      acc.x += v.x;
      acc.y += v.y;
    }
    const float a = a01 * (p.apod[2] ? p.apod[2][n2] : 1.0f);
    acc.x *= a;
    acc.y *= a;
    const int o2 = p.roll_out ? pmod(n2 - N2 / 2, N2) : n2;
    if (p.accumulate) {
      float2 o = obase[o2];
      o.x += acc.x;
      o.y += acc.y;
      obase[o2] = o;
    } else {
      obase[o2] = acc;
    }
  }
}

// data[b][i] *= k[i]
__global__ __launch_bounds__(256) void cmul_bcast_kernel(float2* __restrict__ data, const float2* __restrict__ k,
                                                         long long n) {
  float2* d = data + (long long)blockIdx.y * n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    d[i] = cmul(d[i], k[i]);
}

}  // namespace mra

using namespace mra;

static int fill_params(PadCropParams& p, int ndim, const int* N, const int* G, const int* center, long long nx,
                       long long ns, long long x_stride, long long s_stride, const float* apod, float scale) {
  if (ndim < 1 || ndim > 3) MRA_FAIL("pad/crop: ndim must be 1..3");
  if (nx <= 0 || ns <= 0) MRA_FAIL("pad/crop: batch sizes must be positive");
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
      if (n <= 0 || g < n) MRA_FAIL("pad/crop: need 0 < N <= G in every dim");
      const int c = center ? center[d - lead] : n / 2;
      if (c < 0 || c > n) MRA_FAIL("pad/crop: centre index out of range");
      p.N[d] = n;
      p.G[d] = g;
      p.c[d] = c;
      p.apod[d] = apod ? apod + off : nullptr;
      off += n;
    }
  }
  p.nx = nx;
  p.ns = ns;
  p.x_stride = x_stride;
  p.s_stride = s_stride;
  p.scale = scale;
  p.conj_maps = 1;
  p.accumulate = 0;
  p.roll_out = 0;
  return 0;
}

extern "C" {

// apod: concatenated per-dim tables (N[0] + ... + N[ndim-1] floats) or null.
// center: per-dim pixel placed at grid index 0 (null: N//2).
int mra_pad_expand(int ndim, const int* N, const int* G, const int* center, const void* x, long long nx,
                   long long x_stride, const void* smaps, long long ns, long long s_stride, const float* apod,
                   float scale, void* out, void* stream) {
  MRA_TRY_BEGIN
  PadCropParams p{};
  if (fill_params(p, ndim, N, G, center, nx, smaps ? ns : 1, x_stride, s_stride, apod, scale)) return 1;
  const long long nrows = p.nx * p.ns * p.G[0] * p.G[1];
  const long long nblk = (nrows + 3) / 4;
  pad_expand_kernel<<<(unsigned)nblk, 256, 0, (hipStream_t)stream>>>((const float2*)x, (const float2*)smaps,
                                                                   (float2*)out, p, nrows, nullptr);
  MRA_CHECK_LAUNCH("pad_expand");
  MRA_TRY_END
}

// Batched problems of one image size: grid q (0 <= q < nq) = pad(apod * x[img_index[q]] * S[q]).
int mra_pad_expand_batched(int ndim, const int* N, const int* G, const int* center, const void* x, long long nimg,
                           const void* smaps, long long nq, const int* img_index, const float* apod, float scale,
                           void* out, void* stream) {
  MRA_TRY_BEGIN
  PadCropParams p{};
  if (fill_params(p, ndim, N, G, center, nimg, 1, 1, 1, apod, scale)) return 1;
  const long long npix = (long long)p.N[0] * p.N[1] * p.N[2];
  p.x_stride = npix;
  p.s_stride = npix;
  const long long nrows = nq * p.G[0] * p.G[1];
  const long long nblk = (nrows + 3) / 4;
  pad_expand_kernel<<<(unsigned)nblk, 256, 0, (hipStream_t)stream>>>((const float2*)x, (const float2*)smaps,
                                                                   (float2*)out, p, nrows, img_index);
  MRA_CHECK_LAUNCH("pad_expand_batched");
  MRA_TRY_END
}

// grid holds nx*ns batch grids; out holds nx images (stride x_stride).
int mra_crop_reduce(int ndim, const int* N, const int* G, const int* center, const void* grid, long long nx,
                    const void* smaps, long long ns, long long s_stride, int conj_maps, const float* apod,
                    float scale, int accumulate, int roll_out, void* out, long long x_stride, void* stream) {
  MRA_TRY_BEGIN
  PadCropParams p{};
  if (fill_params(p, ndim, N, G, center, nx, ns, x_stride, s_stride, apod, scale)) return 1;
  p.conj_maps = conj_maps;
  p.accumulate = accumulate;
  p.roll_out = roll_out;
  const long long nrows = p.nx * p.N[0] * p.N[1];
  const long long nblk = (nrows + 3) / 4;
  crop_reduce_kernel<<<(unsigned)nblk, 256, 0, (hipStream_t)stream>>>((const float2*)grid, (const float2*)smaps,
                                                                    (float2*)out, p, nrows, nullptr);
  MRA_CHECK_LAUNCH("crop_reduce");
  MRA_TRY_END
}

// Batched problems of one image size: image i = scale * apod * sum_{q in [coil_off[i], coil_off[i+1])}
// conj(S[q]) * crop(grid[q]).
int mra_crop_reduce_batched(int ndim, const int* N, const int* G, const int* center, const void* grid,
                            long long nimg, const void* smaps, const int* coil_off, const float* apod, float scale,
                            void* out, void* stream) {
  MRA_TRY_BEGIN
  PadCropParams p{};
  if (fill_params(p, ndim, N, G, center, nimg, 1, 1, 1, apod, scale)) return 1;
  const long long npix = (long long)p.N[0] * p.N[1] * p.N[2];
  p.x_stride = npix;
  p.s_stride = npix;
  const long long nrows = p.nx * p.N[0] * p.N[1];
  const long long nblk = (nrows + 3) / 4;
  crop_reduce_kernel<<<(unsigned)nblk, 256, 0, (hipStream_t)stream>>>((const float2*)grid, (const float2*)smaps,
                                                                    (float2*)out, p, nrows, coil_off);
  MRA_CHECK_LAUNCH("crop_reduce_batched");
  MRA_TRY_END
}

int mra_cmul_bcast(void* data, const void* k, long long nbatch, long long n, void* stream) {
  MRA_TRY_BEGIN
  if (nbatch <= 0 || n <= 0) MRA_FAIL("cmul: sizes must be positive");
  if (nbatch > 65535) MRA_FAIL("cmul: batch larger than 65535");
  long long bx = (n + 255) / 256;
  if (bx > 4096) bx = 4096;
  dim3 grid((unsigned)bx, (unsigned)nbatch);
  cmul_bcast_kernel<<<grid, 256, 0, (hipStream_t)stream>>>((float2*)data, (const float2*)k, n);
  MRA_CHECK_LAUNCH("cmul_bcast");
  MRA_TRY_END
}

}  // extern "C"
