// In-place batched complex-to-complex FFTs over the trailing 1..3 dims (rocFFT through
// hipFFT), for the oversampled grids of the NUFFT. The grids are temporaries the
// caller owns, so the transform runs in place: no copy of the input is made (an
// out-of-place transform that must preserve its input costs a full grid copy).
// Unnormalised in both directions, like torch.fft.fftn / ifftn(norm="forward").
//
// Replaces the centred FFTs of the reference NUFFT (fourier.py:37-61 fft/ifft,
// used by triton_nufft.forward_FT_only / adjoint_iFT_only, fourier.py:489-525); the
// centring shifts are index maps in the pad / crop / KB kernels instead.
#include <hipfft/hipfft.h>

#include "capi_util.h"

namespace {

const char* fft_status(hipfftResult r) {
  switch (r) {
    case HIPFFT_SUCCESS: return "success";
    case HIPFFT_INVALID_PLAN: return "invalid plan";
    case HIPFFT_ALLOC_FAILED: return "allocation failed";
    case HIPFFT_INVALID_VALUE: return "invalid value";
    case HIPFFT_INTERNAL_ERROR: return "internal error";
    case HIPFFT_EXEC_FAILED: return "exec failed";
    case HIPFFT_SETUP_FAILED: return "setup failed";
    case HIPFFT_INVALID_SIZE: return "invalid size";
    default: return "hipFFT error";
  }
}

}  // namespace

#define MRA_FFT_CHECK(call, what)                                        \
  do {                                                                   \
    hipfftResult r_ = (call);                                            \
    if (r_ != HIPFFT_SUCCESS) {                                          \
      char b_[160];                                                      \
      snprintf(b_, sizeof(b_), "%s: %s (%d)", what, fft_status(r_), (int)r_); \
      ::mra::set_last_error(b_);                                         \
      return 1;                                                          \
    }                                                                    \
  } while (0)

extern "C" {

// Plan for `batch` contiguous grids of shape n[0..ndim) (row-major), complex64.
int mra_fft_plan(int ndim, const int* n, int batch, void** plan) {
  MRA_TRY_BEGIN
  if (ndim < 1 || ndim > 3) MRA_FAIL("fft: ndim must be 1..3");
  if (batch < 1) MRA_FAIL("fft: batch must be positive");
  int dims[3];
  long long dist = 1;
  for (int d = 0; d < ndim; ++d) {
    if (n[d] < 1) MRA_FAIL("fft: sizes must be positive");
    dims[d] = n[d];
    dist *= n[d];
  }
  if (dist > 0x7fffffffLL) MRA_FAIL("fft: grid larger than 2^31-1 points");
  hipfftHandle h = nullptr;
  MRA_FFT_CHECK(hipfftPlanMany(&h, ndim, dims, nullptr, 1, (int)dist, nullptr, 1, (int)dist, HIPFFT_C2C, batch),
                "fft plan");
  *plan = (void*)h;
  MRA_TRY_END
}

// In-place transform of `data` on `stream`: forward (e^{-i...}) when inverse == 0.
int mra_fft_exec(void* plan, void* data, int inverse, void* stream) {
  MRA_TRY_BEGIN
  if (!plan) MRA_FAIL("fft: null plan");
  hipfftHandle h = (hipfftHandle)plan;
  MRA_FFT_CHECK(hipfftSetStream(h, (hipStream_t)stream), "fft set stream");
  hipfftComplex* x = (hipfftComplex*)data;
  MRA_FFT_CHECK(hipfftExecC2C(h, x, x, inverse ? HIPFFT_BACKWARD : HIPFFT_FORWARD), "fft exec");
  MRA_TRY_END
}

int mra_fft_destroy(void* plan) {
  MRA_TRY_BEGIN
  if (plan) MRA_FFT_CHECK(hipfftDestroy((hipfftHandle)plan), "fft destroy");
  MRA_TRY_END
}

}  // extern "C"
