"""GPU parity of the autograd seam the NUFFT calls (``interpolate`` / ``interpolate_adjoint``,
reference triton_interp/interp.py:15-177): values against the reference CPU path restated in
fp32, gradients for the values and for the coordinates against autograd through the FP64
oracle (NaN-safe masking), real and complex inputs, 1D to 3D."""

import math

import numpy as np
import pytest
import torch

from helpers import rel_l2
from oracle import interp_ref

pytestmark = pytest.mark.gpu

CASES = [  # (grid size, width, complex values)
    ((64,), 5.0, False),
    ((40, 36), 4.0, True),
    ((40, 36), 6.0, False),
    ((16, 18, 20), 4.0, True),
    ((12, 14, 10), 6.0, False),
]


def _beta(w, os_=1.25):
    return math.pi * math.sqrt(((w / os_) * (os_ - 0.5)) ** 2 - 0.8)


def _inputs(G, cplx, seed, nb=2, npts=700):
    g = torch.Generator().manual_seed(seed)
    D = len(G)
    locs = (torch.rand((npts, D), generator=g) * 3 - 1) * torch.tensor(G, dtype=torch.float32)  # wraps
    shape_g = (nb, *G)
    if cplx:
        vals = torch.complex(torch.randn(shape_g, generator=g), torch.randn(shape_g, generator=g))
        samp = torch.complex(torch.randn((nb, npts), generator=g), torch.randn((nb, npts), generator=g))
    else:
        vals = torch.randn(shape_g, generator=g)
        samp = torch.randn((nb, npts), generator=g)
    return locs, vals, samp


def _loss(out, target):
    return torch.sum(torch.abs(out - target) ** 2)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_interpolate_values_and_gradients(dev, case):
    from mr_recon_amd.hip_interp import interpolate

    G, w, cplx = CASES[case]
    locs, vals, samp = _inputs(G, cplx, seed=case)
    kp = {"beta": _beta(w)}
    vd = vals.to(dev).requires_grad_(True)
    ld = locs.to(dev).requires_grad_(True)
    out = interpolate(vd, ld, w, kernel_params=kp)
    assert out.shape == samp.shape and out.is_complex() == cplx
    ref32 = interp_ref.ungrid(vals, locs, w, beta=kp["beta"])
    assert rel_l2(out, ref32) < 1e-5
    _loss(out, samp.to(dev)).backward()

    v64 = (vals.to(torch.complex128) if cplx else vals.to(torch.float64)).requires_grad_(True)
    l64 = locs.to(torch.float64).requires_grad_(True)
    o64 = interp_ref.ungrid(v64, l64, w, beta=kp["beta"], dtype=torch.float64, coords64=True, safe=True)
    if not cplx:
        o64 = o64.real
    _loss(o64, samp.to(o64.dtype)).backward()
    assert rel_l2(vd.grad, v64.grad) < 1e-5
    assert rel_l2(ld.grad.to(torch.float64), l64.grad) < 1e-3


@pytest.mark.parametrize("case", range(len(CASES)))
def test_interpolate_adjoint_values_and_gradients(dev, case):
    from mr_recon_amd.hip_interp import interpolate_adjoint

    G, w, cplx = CASES[case]
    locs, vals, samp = _inputs(G, cplx, seed=10 + case)
    kp = {"beta": _beta(w)}
    sd = samp.to(dev).requires_grad_(True)
    ld = locs.to(dev).requires_grad_(True)
    out = interpolate_adjoint(sd, ld, G, w, kernel_params=kp)
    assert out.shape == vals.shape and out.is_complex() == cplx
    ref32 = interp_ref.grid(samp, locs, G, w, beta=kp["beta"])
    assert rel_l2(out, ref32) < 1e-5
    _loss(out, vals.to(dev)).backward()

    s64 = (samp.to(torch.complex128) if cplx else samp.to(torch.float64)).requires_grad_(True)
    l64 = locs.to(torch.float64).requires_grad_(True)
    o64 = interp_ref.grid(s64, l64, G, w, beta=kp["beta"], dtype=torch.float64, coords64=True, safe=True)
    if not cplx:
        o64 = o64.real
    _loss(o64, vals.to(o64.dtype)).backward()
    assert rel_l2(sd.grad, s64.grad) < 1e-5
    assert rel_l2(ld.grad.to(torch.float64), l64.grad) < 1e-3


@pytest.mark.parametrize("G,w", [((40, 36), 6.0), ((16, 18, 20), 4.0)])
def test_locs_gradient_on_the_lattice(dev, G, w):
    """Samples with t - w/2 on grid integers reach floor(w) + 1 points per dimension (the
    support edge |u| = 1 included, where the weight is I0(0) and its derivative finite);
    the rest reach one fewer. A third of the samples on the lattice: values and the locs
    gradient against FP64 oracle autograd."""
    from mr_recon_amd.hip_interp import interpolate

    locs, vals, samp = _inputs(G, True, seed=77, npts=900)
    g = torch.Generator().manual_seed(78)
    on = torch.rand(locs.shape[0], generator=g) < 0.33
    lattice = torch.floor(locs) + (w / 2 - math.floor(w / 2))
    locs = torch.where(on[:, None], lattice, locs)
    kp = {"beta": _beta(w)}
    vd = vals.to(dev)
    ld = locs.to(dev).requires_grad_(True)
    out = interpolate(vd, ld, w, kernel_params=kp)
    assert rel_l2(out, interp_ref.ungrid(vals, locs, w, beta=kp["beta"])) < 1e-5
    _loss(out, samp.to(dev)).backward()
    l64 = locs.to(torch.float64).requires_grad_(True)
    o64 = interp_ref.ungrid(vals.to(torch.complex128), l64, w, beta=kp["beta"], dtype=torch.float64,
                            coords64=True, safe=True)
    _loss(o64, samp.to(o64.dtype)).backward()
    assert rel_l2(ld.grad.to(torch.float64), l64.grad) < 1e-3


def test_interpolate_pair_is_adjoint(dev):
    """<interpolate(v), s> = <v, interpolate_adjoint(s)> through the seam (2D complex)."""
    from mr_recon_amd.hip_interp import interpolate, interpolate_adjoint

    G, w = (40, 36), 4.0
    locs, vals, samp = _inputs(G, True, seed=99)
    kp = {"beta": _beta(w)}
    a = interpolate(vals.to(dev), locs.to(dev), w, kernel_params=kp).cpu().to(torch.complex128)
    b = interpolate_adjoint(samp.to(dev), locs.to(dev), G, w, kernel_params=kp).cpu().to(torch.complex128)
    lhs = complex(torch.sum(a.conj() * samp.to(torch.complex128)))
    rhs = complex(torch.sum(vals.to(torch.complex128).conj() * b))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
    assert np.isfinite(lhs.real)


@pytest.mark.parametrize("G,w,cplx", [((6, 7, 5, 8), 4.0, True), ((5, 6, 4, 5, 6), 3.0, False)])
def test_ungrid_grid_4d_5d_vs_oracle(dev, G, w, cplx):
    """4D / 5D (the reference's _ungrid4d / _ungrid5d and grid's torch fallback) against the
    CPU path restated, values and the value gradients through the autograd seam."""
    from mr_recon_amd.hip_interp import grid, interpolate, interpolate_adjoint, ungrid

    locs, vals, samp = _inputs(G, cplx, seed=len(G), npts=300)
    kp = {"beta": _beta(w)}
    y = ungrid(vals.to(dev), locs.to(dev), w, kernel_params=kp)
    assert rel_l2(y, interp_ref.ungrid(vals, locs, w, beta=kp["beta"])) < 1e-5
    z = grid(samp.to(dev), locs.to(dev), G, w, kernel_params=kp)
    assert rel_l2(z, interp_ref.grid(samp, locs, G, w, beta=kp["beta"])) < 1e-5
    vd = vals.to(dev).requires_grad_(True)
    _loss(interpolate(vd, locs.to(dev), w, kernel_params=kp), samp.to(dev)).backward()
    v64 = (vals.to(torch.complex128) if cplx else vals.to(torch.float64)).requires_grad_(True)
    o64 = interp_ref.ungrid(v64, locs, w, beta=kp["beta"], dtype=torch.float64)
    _loss(o64 if cplx else o64.real, samp.to(o64.dtype)).backward()
    assert rel_l2(vd.grad, v64.grad) < 1e-5
    sd = samp.to(dev).requires_grad_(True)
    _loss(interpolate_adjoint(sd, locs.to(dev), G, w, kernel_params=kp), vals.to(dev)).backward()
    s64 = (samp.to(torch.complex128) if cplx else samp.to(torch.float64)).requires_grad_(True)
    o64 = interp_ref.grid(s64, locs, G, w, beta=kp["beta"], dtype=torch.float64)
    _loss(o64 if cplx else o64.real, vals.to(o64.dtype)).backward()
    assert rel_l2(sd.grad, s64.grad) < 1e-5
