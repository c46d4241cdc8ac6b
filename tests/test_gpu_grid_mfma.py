"""GPU parity of the 3D matrix-core gridding kernel (cells of samples summed with
v_mfma_f32_16x16x4f32) against the oracle and against the per-lane LDS kernel on the same
plan: kernel widths P = 3..7, coil counts around the 8-coil block edge, hot tiles split
into several items, cells cut by stage and item boundaries, long runs of one corner, DCF."""

import contextlib

import numpy as np
import pytest
import torch

from helpers import edge_coords, rel_l2
from oracle import interp_ref

pytestmark = pytest.mark.gpu

TOL = 1e-5


@contextlib.contextmanager
def grid_kernel(mode):
    from mr_recon_amd import _lib

    old = _lib.lib().mra_set_grid_kernel(mode)
    try:
        yield
    finally:
        _lib.lib().mra_set_grid_kernel(old)


def _coords(G, width, seed):
    rng = np.random.default_rng(seed)
    parts = [edge_coords((G,) * 3, 1500, seed=seed, width=width)]
    # hot tile: far more samples than one item holds
    parts.append((8 + rng.random((3000, 3)) * 8).astype(np.float32))
    # long runs of one corner and near-duplicates (runs cut by the 64-sample stages)
    p = (rng.random((6, 3)) * G).astype(np.float32)
    parts.append(np.repeat(p, [1, 63, 64, 65, 130, 7], axis=0))
    parts.append((p + 1e-3 * rng.random((6, 3))).astype(np.float32))
    k = np.concatenate(parts).astype(np.float32)
    return torch.from_numpy(k[rng.permutation(k.shape[0])])


def _grid_both(dev, k, y, dcf, G, width, beta, S=None):
    from mr_recon_amd import ops
    from mr_recon_amd.plan import SamplePlan

    plan = SamplePlan(k.to(dev), (G,) * 3, width, beta, samples_per_item=S)
    outs = []
    for mode in (1, 0):
        out = torch.empty((y.shape[0], G, G, G), dtype=torch.complex64, device=dev)
        with grid_kernel(mode):
            ops.gridding(y.to(dev), plan, out, dcf=None if dcf is None else dcf.to(dev))
        outs.append(out)
    return outs


@pytest.mark.parametrize("width", [2.0, 3.5, 4.0, 5.0, 6.0])
@pytest.mark.parametrize("C", [1, 8, 9, 20])
def test_mfma_grid_vs_oracle_and_lane_kernel(dev, width, C):
    G = 24
    k = _coords(G, width, seed=int(width * 10) + C)
    g = torch.Generator().manual_seed(C)
    y = torch.randn((C, k.shape[0]), dtype=torch.complex64, generator=g)
    dcf = torch.rand(k.shape[0], generator=g) + 0.5
    beta = 2.3 * width
    mfma, lane = _grid_both(dev, k, y, dcf, G, width, beta)
    ref = interp_ref.grid(y * dcf, k, (G,) * 3, width, beta=beta)
    assert rel_l2(mfma, ref) < TOL
    assert rel_l2(lane, ref) < TOL
    assert rel_l2(mfma, lane) < TOL


@pytest.mark.parametrize("width", [4.0, 5.0, 6.0, 3.5, 7.0])
@pytest.mark.parametrize("frac", [0.0, 0.002, 1.0])
def test_mfma_grid_short_and_full_items(dev, width, frac):
    """An item whose samples never reach the last patch point of a dimension (t - w/2 off the
    grid integers) is gridded with P - 1 points per dimension and 1 x 4 corner cells; one
    with such a sample keeps P points and 2 x 4 cells. No sample, a few (scattered over
    items) and every sample on the lattice, against the oracle and the lane kernel."""
    G = 24
    rng = np.random.default_rng(int(width * 10) + int(frac * 1000))
    n = 6000
    k = rng.random((n, 3)) * G
    on = rng.random(n) < frac
    k[on] = np.floor(k[on]) + (width / 2 - np.floor(width / 2))
    k = torch.from_numpy(k.astype(np.float32))
    g = torch.Generator().manual_seed(3)
    y = torch.randn((6, n), dtype=torch.complex64, generator=g)
    beta = 2.3 * width
    mfma, lane = _grid_both(dev, k, y, None, G, width, beta, S=500)
    ref = interp_ref.grid(y, k, (G,) * 3, width, beta=beta)
    assert rel_l2(mfma, ref) < TOL
    assert rel_l2(mfma, lane) < TOL


@pytest.mark.parametrize("S", [64, 100, 1000])
def test_mfma_grid_item_boundaries(dev, S):
    """Items of S samples cut cells at arbitrary places; each item flushes its own part."""
    G = 32
    k = _coords(G, 4.0, seed=S)
    y = torch.randn((4, k.shape[0]), dtype=torch.complex64, generator=torch.Generator().manual_seed(S))
    mfma, lane = _grid_both(dev, k, y, None, G, 4.0, 9.0, S=S)
    ref = interp_ref.grid(y, k, (G,) * 3, 4.0, beta=9.0)
    assert rel_l2(mfma, ref) < TOL
    assert rel_l2(mfma, lane) < TOL


def test_mfma_grid_kooshball_many_coils(dev):
    """A kooshball trajectory (dense centre, sparse edge) with 32 coils on a 64^3 grid."""
    from mr_recon_amd import synth

    N, G = 48, 60
    k = synth.kooshball_3d(400, 96, N).reshape(-1, 3) * (G / N) + G // 2
    y = torch.randn((32, k.shape[0]), dtype=torch.complex64, generator=torch.Generator().manual_seed(5))
    mfma, lane = _grid_both(dev, k.float(), y, None, G, 4.0, 9.1)
    assert rel_l2(mfma, lane) < TOL


@pytest.mark.parametrize("D", [2, 3])
def test_dcf_in_plan_order(dev, D):
    """dcf handed over in sorted (plan) order (dcf[perm], read contiguously) grids the same
    as dcf in sample order, for both phase-1 kernels."""
    from mr_recon_amd import ops
    from mr_recon_amd.plan import SamplePlan

    G = 32
    k = torch.from_numpy(edge_coords((G,) * D, 4000, seed=D, width=4.0))
    y = torch.randn((5, k.shape[0]), dtype=torch.complex64, generator=torch.Generator().manual_seed(D)).to(dev)
    dcf = (torch.rand(k.shape[0], generator=torch.Generator().manual_seed(9)) + 0.5).to(dev)
    plan = SamplePlan(k.to(dev), (G,) * D, 4.0, 9.0)
    ref = interp_ref.grid(y.cpu() * dcf.cpu(), k, (G,) * D, 4.0, beta=9.0)
    for mode in (1, 0):
        with grid_kernel(mode):
            a = torch.empty((5, *(G,) * D), dtype=torch.complex64, device=dev)
            b = torch.empty_like(a)
            ops.gridding_reduce(ops.gridding_partial(y, plan, 5, dcf=dcf), plan, a)
            ops.gridding_reduce(ops.gridding_partial(y, plan, 5, dcf=dcf[plan.perm.long()].contiguous(),
                                                     dcf_plan_order=True), plan, b)
        assert rel_l2(a, ref) < TOL
        assert torch.equal(a, b)


@pytest.mark.parametrize("D,C", [(2, 1), (2, 4), (3, 1), (3, 5)])
def test_ksp_in_plan_order(dev, D, C):
    """Interpolation output and gridding input (and dcf) kept in sorted (plan) order give
    bit-identical values at the permuted positions, for the single- and coil-pair
    interpolation kernels and both phase-1 gridding kernels."""
    from mr_recon_amd import ops
    from mr_recon_amd.plan import SamplePlan

    G = 32
    k = torch.from_numpy(edge_coords((G,) * D, 4000, seed=D + C, width=4.0))
    gen = torch.Generator().manual_seed(C)
    grid = torch.randn((C, *(G,) * D), dtype=torch.complex64, generator=gen).to(dev)
    y = torch.randn((C, k.shape[0]), dtype=torch.complex64, generator=gen).to(dev)
    dcf = (torch.rand(k.shape[0], generator=gen) + 0.5).to(dev)
    plan = SamplePlan(k.to(dev), (G,) * D, 4.0, 9.0)
    perm = plan.perm.long()
    a = ops.interp(grid, plan)
    b = ops.interp(grid, plan, geom=plan.geom(C, ksp_plan_order=True))
    assert torch.equal(a[:, perm], b)
    for mode in (1, 0):
        with grid_kernel(mode):
            ga = torch.empty((C, *(G,) * D), dtype=torch.complex64, device=dev)
            gb = torch.empty_like(ga)
            ops.gridding_reduce(ops.gridding_partial(y, plan, C, dcf=dcf), plan, ga)
            ops.gridding_reduce(ops.gridding_partial(y[:, perm].contiguous(), plan, C, dcf=dcf[perm].contiguous(),
                                                     ksp_plan_order=True), plan, gb)
        assert torch.equal(ga, gb)
