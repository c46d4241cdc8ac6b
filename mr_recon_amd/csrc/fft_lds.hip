// Batched complex FFTs along one axis of a row-major grid, staged through LDS: one
// global read and one global write per element per axis (rocFFT's plan for a 160^3
// batch takes four passes per 3D transform). Unnormalised, in place. The inverse
// transform is conj(F(conj(x))), applied on load and store.
//
// Axis access patterns: along the contiguous last axis a block takes NT consecutive
// rows (each row is L contiguous values); along any other axis (stride `inner`) a block
// takes NT consecutive columns, so every row of the load is NT contiguous values.
// Lengths are compile-time (radix plans in fft_lds.h); others go through rocFFT.
#include "capi_util.h"
#include "fft_lds.h"

namespace mra {
namespace fft {

// ROW: transforms along the contiguous last axis (inner == 1), NT consecutive rows per
// block. COL: transforms along an axis of stride `inner` (inner >= NT), NT consecutive
// columns per block.
template <int L, bool ROW>
__global__ __launch_bounds__(kThreadsFFT) void fft_axis_kernel(float2* __restrict__ data, long long outer,
                                                                long long inner, const float2* __restrict__ twg,
                                                                int inverse) {
  using S = Shape<L, ROW>;
  __shared__ float2 buf[L * S::NTP];
  __shared__ float2 tw[L];
  const int tid = threadIdx.x;
  for (int m = tid; m < L; m += kThreadsFFT) tw[m] = twg[m];
  long long o, i0;
  int nvalid;
  if constexpr (ROW) {
    o = (long long)blockIdx.x * S::NT;
    i0 = 0;
    nvalid = (int)min((long long)S::NT, outer - o);
  } else {
    const long long nblk = (inner + S::NT - 1) / S::NT;
    o = blockIdx.x / nblk;
    i0 = (blockIdx.x - o * nblk) * S::NT;
    nvalid = (int)min((long long)S::NT, inner - i0);
  }
  const float sgn = inverse ? -1.0f : 1.0f;  // conj on load and store for the inverse
  if constexpr (ROW) {
    const float2* src = data + o * L;
    for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
      const int c = e / L, j = e - c * L;
      if (c < nvalid) {
        const float2 x = src[(long long)c * L + j];
        buf[j * S::NTP + c] = make_float2(x.x, sgn * x.y);
      }
    }
  } else {
    const float2* src = data + o * L * inner + i0;
    for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
      const int j = e / S::NT, c = e - j * S::NT;
      if (c < nvalid) {
        const float2 x = src[(long long)j * inner + c];
        buf[j * S::NTP + c] = make_float2(x.x, sgn * x.y);
      }
    }
  }
  __syncthreads();
  fft_lds<L, S>(buf, tw);
  if constexpr (ROW) {
    float2* dst = data + o * L;
    for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
      const int c = e / L, j = e - c * L;
      if (c < nvalid) {
        const float2 x = buf[j * S::NTP + c];
        dst[(long long)c * L + j] = make_float2(x.x, sgn * x.y);
      }
    }
  } else {
    float2* dst = data + o * L * inner + i0;
    for (int e = tid; e < S::NT * L; e += kThreadsFFT) {
      const int j = e / S::NT, c = e - j * S::NT;
      if (c < nvalid) {
        const float2 x = buf[j * S::NTP + c];
        dst[(long long)j * inner + c] = make_float2(x.x, sgn * x.y);
      }
    }
  }
}

template <int L>
int launch_axis(float2* data, long long outer, long long inner, const float2* tw, int inverse, hipStream_t st) {
  if (inner == 1) {
    using S = Shape<L, true>;
    const long long nb = (outer + S::NT - 1) / S::NT;
    if (nb > 0x7fffffffLL) MRA_FAIL("fft: too many transforms");
    fft_axis_kernel<L, true><<<(unsigned)nb, kThreadsFFT, 0, st>>>(data, outer, inner, tw, inverse);
  } else {
    using S = Shape<L, false>;
    if (inner < S::NT) MRA_FAIL("fft: inner extent below the column tile");
    const long long nb = outer * ((inner + S::NT - 1) / S::NT);
    if (nb > 0x7fffffffLL) MRA_FAIL("fft: too many transforms");
    fft_axis_kernel<L, false><<<(unsigned)nb, kThreadsFFT, 0, st>>>(data, outer, inner, tw, inverse);
  }
  MRA_CHECK_LAUNCH("fft axis");
  return 0;
}

int dispatch_axis(int L, float2* data, long long outer, long long inner, const float2* tw, int inverse,
                  hipStream_t st) {
  const int rc = with_length(L, [&](auto LL) {
    return launch_axis<decltype(LL)::value>(data, outer, inner, tw, inverse, st);
  });
  if (rc < 0) MRA_FAIL("fft: length without an LDS plan");
  return rc;
}

bool supported(int L) { return plan_supported(L); }

}  // namespace fft
}  // namespace mra

using namespace mra;

extern "C" {

// 1 when every length of n[0..ndim) has an LDS plan and every non-last axis has at
// least 16 columns after it (the column tile), else 0.
int mra_fft_lds_supported(int ndim, const int* n) {
  if (ndim < 1 || ndim > 3) return 0;
  long long inner = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    if (!fft::supported(n[d])) return 0;
    if (d < ndim - 1 && inner < 64) return 0;
    inner *= n[d];
  }
  return 1;
}

// Twiddle table for length L: tw[m] = exp(-2 pi i m / L), host memory, L complex.
int mra_fft_lds_twiddles(int L, float* tw_host) {
  MRA_TRY_BEGIN
  if (L < 1) MRA_FAIL("fft: bad length");
  for (int m = 0; m < L; ++m) {
    const double a = -2.0 * 3.14159265358979323846 * (double)m / (double)L;
    tw_host[2 * m] = (float)cos(a);
    tw_host[2 * m + 1] = (float)sin(a);
  }
  MRA_TRY_END
}

// In-place unnormalised FFT over the trailing ndim axes of `batch` grids of shape n,
// one LDS pass per axis. tw[d] = device twiddle table of length n[d].
int mra_fft_lds(int ndim, const int* n, long long batch, void* data, const void* const* tw, int inverse,
                void* stream) {
  MRA_TRY_BEGIN
  if (!mra_fft_lds_supported(ndim, n)) MRA_FAIL("fft: shape not supported by the LDS transforms");
  if (batch < 1) MRA_FAIL("fft: batch must be positive");
  long long total = batch;
  for (int d = 0; d < ndim; ++d) total *= n[d];
  long long inner = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    const long long outer = total / ((long long)n[d] * inner);
    if (fft::dispatch_axis(n[d], (float2*)data, outer, inner, (const float2*)tw[d], inverse,
                           (hipStream_t)stream))
      return 1;
    inner *= n[d];
  }
  MRA_TRY_END
}

}  // extern "C"
