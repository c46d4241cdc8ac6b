"""GPU parity of the in-place grid FFTs (LDS-staged passes and the rocFFT fallback)
against torch.fft on the same data."""

import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(3, 160, 160, 160), (2, 64, 96, 128), (4, 320, 320), (2, 240, 240),
                                   (1, 512, 512), (5, 640), (3, 128, 80, 120), (2, 192, 384),
                                   (2, 256, 480)])
def test_lds_fft_matches_torch(dev, shape):
    from mr_recon_amd.fourier import _fftn, _ifftn, _lds_fft

    d = len(shape) - 1
    assert _lds_fft.supported(tuple(shape[1:]))
    g = torch.Generator().manual_seed(sum(shape))
    x = torch.randn(shape, dtype=torch.complex64, generator=g).to(dev)
    dims = tuple(range(-d, 0))
    want = torch.fft.fftn(x, dim=dims)
    got = _fftn(x.clone(), d)
    assert rel_l2(got, want) < 5e-6
    want_i = torch.fft.ifftn(x, dim=dims, norm="forward")
    got_i = _ifftn(x.clone(), d)
    assert rel_l2(got_i, want_i) < 5e-6


@pytest.mark.parametrize("shape", [(2, 100, 100), (3, 48, 48, 48), (2, 64, 30)])
def test_fft_fallback_shapes(dev, shape):
    from mr_recon_amd.fourier import _fftn, _lds_fft

    d = len(shape) - 1
    assert not _lds_fft.supported(tuple(shape[1:]))
    x = torch.randn(shape, dtype=torch.complex64, generator=torch.Generator().manual_seed(1)).to(dev)
    want = torch.fft.fftn(x, dim=tuple(range(-d, 0)))
    assert rel_l2(_fftn(x.clone(), d), want) < 5e-6
