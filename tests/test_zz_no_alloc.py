"""The no-allocation gridding entry point (``mra_grid_noalloc``, static scratch) against the
existing one (``mra_grid_partial`` + ``mra_grid_reduce`` with a caller scratch buffer), called
through the C-ABI: the same kernels in the same order, so the grids must be bit-identical.
Two calls in a row, the second with a smaller problem (its subgrids overwrite only part of
what the first left in the static array); sizes past the static scratch are refused."""

import ctypes
import math

import pytest
import torch

from mr_recon_amd import _lib, ops, synth
from mr_recon_amd.fourier import hip_nufft

pytestmark = pytest.mark.gpu


def _noalloc(ksp, plan, out, dcf=None, accumulate=False):
    L = _lib.lib()
    gh = plan.geom_host(out.shape[0])
    st = torch.cuda.current_stream(out.device).cuda_stream
    return L.mra_grid_noalloc(plan.ndim, plan.P, ksp.data_ptr(), None if dcf is None else dcf.data_ptr(),
                              plan.wsort.data_ptr(), plan.lbase.data_ptr(), plan.perm.data_ptr(),
                              ctypes.addressof(gh), plan.items.data_ptr(), plan.nitems, out.shape[0],
                              plan.tile_items.data_ptr(), plan.ntiles, out.data_ptr(), int(accumulate), st)


def _problem(dev, im, C, trj, w, seed, with_dcf):
    nf = hip_nufft(im, oversamp=1.25, width=w)
    plan = nf.plan(trj.reshape(-1, len(im)).to(dev))
    ksp = synth.random_image((C, plan.npts), seed=seed).to(dev)
    dcf = (torch.rand(plan.npts, generator=torch.Generator().manual_seed(seed)) + 0.5).to(dev) if with_dcf else None
    return plan, ksp, dcf


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_gridding_matches_existing_entry_point(dev, dim):
    L = _lib.lib()
    if dim == 2:
        cases = [((96, 96), 12, synth.spiral_2d(16, 1500, 96), 4, True),
                 ((64, 64), 5, synth.radial_2d(40, 128, 64), 6, False)]
    else:
        cases = [((24, 24, 24), 6, synth.kooshball_3d(300, 48, 24), 4, False),
                 ((16, 16, 16), 2, synth.kooshball_3d(120, 32, 16), 4, True)]
    prev = None
    for k, (im, C, trj, w, with_dcf) in enumerate(cases):
        print(f"no-alloc gridding: {dim}D case {k}: image {im}, {C} coils, {trj.numel() // dim} samples", flush=True)
        plan, ksp, dcf = _problem(dev, im, C, trj, w, 10 * dim + k, with_dcf)
        need = plan.nitems * C * plan.subgrid * 8
        assert need <= L.mra_grid_noalloc_max_bytes()
        if prev is not None:
            assert need < prev  # the second call is the smaller one
        prev = need
        ref = torch.empty((C, *plan.grid_size), dtype=torch.complex64, device=dev)
        ops.gridding(ksp, plan, ref, dcf=dcf)
        got = torch.full_like(ref, float("nan"))
        _lib.check(_noalloc(ksp, plan, got, dcf))
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        # accumulate adds to what the output holds, as the existing entry point does
        _lib.check(_noalloc(ksp, plan, got, dcf, accumulate=True))
        ops.gridding(ksp, plan, ref, dcf=dcf, accumulate=True)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
        # the tensor-level wrapper takes the same path
        alt = torch.empty_like(ref)
        ops.gridding_noalloc(ksp, plan, alt, dcf=dcf)
        torch.cuda.synchronize()
        assert torch.equal(alt * 2, got)


def test_noalloc_gridding_refuses_oversized_problems(dev):
    """A plan whose subgrids exceed the static scratch is refused before any launch."""
    L = _lib.lib()
    plan, ksp, _ = _problem(dev, (64, 64, 64), 32, synth.kooshball_3d(2000, 128, 64), 4, 7, False)
    need = plan.nitems * 32 * plan.subgrid * 8
    assert need > L.mra_grid_noalloc_max_bytes()
    out = torch.zeros((32, *plan.grid_size), dtype=torch.complex64, device=dev)
    rc = _noalloc(ksp, plan, out)
    assert rc != 0 and b"static" in L.mra_last_error()
    with pytest.raises(RuntimeError, match="scratch"):
        ops.gridding_noalloc(ksp, plan, out)
    assert int(torch.count_nonzero(out)) == 0
