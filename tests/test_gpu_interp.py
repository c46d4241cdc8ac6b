"""GPU parity: binning plans (bit-exact), KB interpolation and gridding vs the CPU oracle."""

import math

import numpy as np
import pytest
import torch

from helpers import cdot, edge_coords, rel_l2
from oracle import bins as obins
from oracle import interp_ref

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L2, complex64 (fp32 accumulation-order differences only)


def _case(D, G, width, n, seed):
    k = edge_coords((G,) * D, n, seed=seed, width=width)
    return torch.from_numpy(k)


@pytest.mark.parametrize("D,G,width", [(1, 96, 4.0), (2, 64, 4.0), (2, 80, 6.0), (3, 24, 4.0), (3, 20, 6.0),
                                       (2, 40, 3.5), (3, 16, 8.0)])
@pytest.mark.parametrize("wrapoff", [0, -1])
def test_bin_plan_bit_exact(dev, D, G, width, wrapoff):
    from mr_recon_amd.plan import SamplePlan

    k = _case(D, G, width, 3000, seed=D * 10 + int(width))
    grid_size = (G,) * D
    wrap = tuple(wrapoff * (g // 2) for g in grid_size)
    plan = SamplePlan(k.to(dev), grid_size, width, 2.0, wrap=wrap)
    ts, perm, start, _ = obins.bin_plan(k.numpy(), grid_size, width, plan.tile, wrap=wrap)
    got_t = plan.tsort.cpu().numpy()
    assert np.array_equal(plan.perm.cpu().numpy(), perm)
    assert np.array_equal(got_t.view(np.int32), ts.view(np.int32))
    assert np.array_equal(plan.tile_start.cpu().numpy(), start)


@pytest.mark.parametrize("D,N", [(2, 64), (3, 20)])
def test_bin_plan_scaled_bit_exact(dev, D, N):
    """The NUFFT scaling t = fl32(fl32(k * fl32(G/N)) + G//2), including non-dyadic G/N."""
    from mr_recon_amd.plan import SamplePlan

    for G in (math.ceil(1.25 * N), math.ceil(1.3 * N) + 1, 2 * N):
        rng = np.random.default_rng(G)
        k = ((rng.random((4000, D)) - 0.5) * N).astype(np.float32)
        k[:20] = np.float32(-N / 2)
        k[20:40] = np.float32(N / 2)
        sc = [float(np.float32(G / N))] * D
        sh = [float(G // 2)] * D
        grid = (G,) * D
        plan = SamplePlan(torch.from_numpy(k).to(dev), grid, 4.0, 5.0, scale=sc, shift=sh,
                          wrap=tuple(-(g // 2) for g in grid))
        ts, perm, start, _ = obins.bin_plan(k, grid, 4.0, plan.tile, wrap=tuple(-(g // 2) for g in grid), scale=sc,
                                            shift=sh)
        assert np.array_equal(plan.perm.cpu().numpy(), perm)
        assert np.array_equal(plan.tsort.cpu().numpy().view(np.int32), ts.view(np.int32))
        # the oracle's scaling equals what the reference computes with torch ops
        t_ref = torch.from_numpy(k).clone()
        for d in range(D):
            t_ref[:, d] *= G / N
            t_ref[:, d] += G // 2
        t_ref = torch.remainder(t_ref, torch.tensor(grid))
        assert np.array_equal(t_ref.numpy()[perm].reshape(-1).view(np.int32), ts.view(np.int32))


@pytest.mark.parametrize("D,G,width,B", [(1, 128, 4.0, 3), (2, 48, 4.0, 2), (2, 64, 6.0, 4), (3, 20, 4.0, 3),
                                         (3, 16, 6.0, 2), (2, 32, 3.0, 1), (3, 12, 8.0, 1), (2, 40, 2.5, 2)])
def test_ungrid_matches_oracle(dev, D, G, width, B):
    from mr_recon_amd.hip_interp import ungrid

    g = torch.Generator().manual_seed(D * 100 + G)
    vals = torch.randn((B, *(G,) * D), dtype=torch.complex64, generator=g)
    k = _case(D, G, width, 2500, seed=G)
    beta = 7.3
    out = ungrid(vals.to(dev), k.to(dev), width, kernel_params={"beta": beta})
    ref = interp_ref.ungrid(vals, k, width, beta=beta)
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("D,G,width,B", [(2, 48, 4.0, 2), (2, 40, 5.0, 3), (3, 20, 4.0, 2), (3, 20, 4.0, 3),
                                         (3, 16, 6.0, 2), (1, 64, 4.0, 1)])
@pytest.mark.parametrize("frac", [0.0, 0.01, 1.0])
def test_ungrid_full_and_short_patches(dev, D, G, width, B, frac):
    """A sample reaches floor(w) + 1 points along a dimension only where t - w/2 is a grid
    integer; elsewhere the last patch point has weight 0 and the gather sums one point
    fewer per dimension. Waves with no such sample, some, and only such samples (a
    fraction ``frac`` of the samples on the lattice, in one or all dimensions), coil
    chunks of odd and even size, against the oracle."""
    from mr_recon_amd.hip_interp import ungrid

    g = torch.Generator().manual_seed(D * 1000 + G + int(10 * width) + int(100 * frac))
    n = 3000
    k = torch.rand((n, D), generator=g) * G
    on = torch.rand(n, generator=g) < frac
    lattice = torch.floor(k) + (width / 2 - math.floor(width / 2))
    alldims = torch.rand(n, generator=g) < 0.5
    dim0 = torch.zeros_like(k, dtype=torch.bool)
    dim0[:, 0] = True
    sel = on[:, None] & (alldims[:, None] | dim0)
    k = torch.where(sel, lattice, k)
    vals = torch.randn((B, *(G,) * D), dtype=torch.complex64, generator=g)
    beta = 6.7
    out = ungrid(vals.to(dev), k.to(dev), width, kernel_params={"beta": beta})
    ref = interp_ref.ungrid(vals, k, width, beta=beta)
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("D,G,width,B", [(2, 48, 4.0, 3), (2, 40, 6.0, 9), (2, 40, 5.0, 2), (1, 96, 4.0, 2),
                                         (3, 20, 8.0, 2)])
@pytest.mark.parametrize("frac", [0.0, 0.01, 1.0])
def test_grid_full_and_short_patches(dev, D, G, width, B, frac):
    """Lane gridding: an item whose samples all miss the last patch plane (dim-0 weight 0)
    accumulates P-1 planes per sample; one with a sample on the lattice keeps P."""
    from mr_recon_amd.hip_interp import grid

    g = torch.Generator().manual_seed(D * 1000 + G + int(10 * width) + int(100 * frac))
    n = 3000
    k = torch.rand((n, D), generator=g) * G
    on = torch.rand(n, generator=g) < frac
    k[on, 0] = torch.floor(k[on, 0]) + (width / 2 - math.floor(width / 2))
    vals = torch.randn((B, n), dtype=torch.complex64, generator=g)
    out = grid(vals.to(dev), k.to(dev), (G,) * D, width, kernel_params={"beta": 6.3})
    assert rel_l2(out, interp_ref.grid(vals, k, (G,) * D, width, beta=6.3)) < TOL


@pytest.mark.parametrize("D,G,width,B", [(1, 128, 4.0, 2), (2, 48, 4.0, 2), (2, 64, 6.0, 3), (3, 20, 4.0, 2),
                                         (3, 16, 6.0, 1), (2, 32, 3.0, 1)])
def test_grid_matches_oracle(dev, D, G, width, B):
    from mr_recon_amd.hip_interp import grid

    k = _case(D, G, width, 2500, seed=G + 1)
    g = torch.Generator().manual_seed(7)
    vals = torch.randn((B, k.shape[0]), dtype=torch.complex64, generator=g)
    beta = 6.1
    out = grid(vals.to(dev), k.to(dev), (G,) * D, width, kernel_params={"beta": beta})
    ref = interp_ref.grid(vals, k, (G,) * D, width, beta=beta)
    assert out.shape == ref.shape
    assert rel_l2(out, ref) < TOL


@pytest.mark.parametrize("D,G,width", [(1, 3, 4.0), (2, 5, 4.0), (3, 5, 4.0), (3, 11, 6.0), (2, 7, 8.0)])
def test_grid_tiny_grids(dev, D, G, width):
    """Grids smaller than one subgrid: a partial subgrid wraps onto the same grid point
    several times and every wrap must be summed."""
    from mr_recon_amd.hip_interp import grid

    g = torch.Generator().manual_seed(G)
    k = (torch.rand((300, D), generator=g) - 0.3) * 2 * G
    vals = torch.randn((2, 300), dtype=torch.complex64, generator=g)
    out = grid(vals.to(dev), k.to(dev), (G,) * D, width, kernel_params={"beta": 5.0})
    assert rel_l2(out, interp_ref.grid(vals, k, (G,) * D, width, beta=5.0)) < TOL


@pytest.mark.parametrize("D,G", [(3, 50), (3, 37), (2, 100), (2, 70), (1, 700)])
def test_grid_partial_last_tile(dev, D, G):
    """G not a multiple of the tile edge: the last tile's subgrid wraps into tiles 0 and 1."""
    from mr_recon_amd.hip_interp import grid, ungrid

    g = torch.Generator().manual_seed(G)
    k = torch.rand((6000, D), generator=g) * G
    k[:200] = G - torch.rand((200, D), generator=g) * 3  # crowd the wrap seam
    vals = torch.randn((3, 6000), dtype=torch.complex64, generator=g)
    out = grid(vals.to(dev), k.to(dev), (G,) * D, 4.0, kernel_params={"beta": 7.0})
    assert rel_l2(out, interp_ref.grid(vals, k, (G,) * D, 4.0, beta=7.0)) < TOL
    x = torch.randn((3, *(G,) * D), dtype=torch.complex64, generator=g)
    y = ungrid(x.to(dev), k.to(dev), 4.0, kernel_params={"beta": 7.0})
    assert rel_l2(y, interp_ref.ungrid(x, k, 4.0, beta=7.0)) < TOL


@pytest.mark.parametrize("D,G", [(3, 176), (2, 1500)])
def test_many_tiles_three_radix_passes(dev, D, G):
    """More than 8191 tiles: the tile id is sorted in two digits (three passes in all)."""
    from mr_recon_amd.hip_interp import grid, ungrid
    from mr_recon_amd.plan import SamplePlan

    k = _case(D, G, 4.0, 3000, seed=G)
    grid_size = (G,) * D
    plan = SamplePlan(k.to(dev), grid_size, 4.0, 6.0, wrap=tuple(-(g // 2) for g in grid_size))
    assert plan.ntiles + 1 > 8192
    ts, perm, start, _ = obins.bin_plan(k.numpy(), grid_size, 4.0, plan.tile, wrap=tuple(-(g // 2) for g in grid_size))
    assert np.array_equal(plan.perm.cpu().numpy(), perm)
    assert np.array_equal(plan.tsort.cpu().numpy().view(np.int32), ts.view(np.int32))
    assert np.array_equal(plan.tile_start.cpu().numpy(), start)
    g = torch.Generator().manual_seed(G)
    vals = torch.randn((1, k.shape[0]), dtype=torch.complex64, generator=g)
    out = grid(vals.to(dev), k.to(dev), grid_size, 4.0, kernel_params={"beta": 6.0})
    assert rel_l2(out, interp_ref.grid(vals, k, grid_size, 4.0, beta=6.0)) < TOL
    x = torch.randn((1, *grid_size), dtype=torch.complex64, generator=g)
    y = ungrid(x.to(dev), k.to(dev), 4.0, kernel_params={"beta": 6.0})
    assert rel_l2(y, interp_ref.ungrid(x, k, 4.0, beta=6.0)) < TOL


@pytest.mark.parametrize("D", [2, 3])
def test_grid_runs_of_equal_corners(dev, D):
    """Long runs of samples with one patch corner (summed in registers, across the
    32-sample stages) mixed with isolated samples."""
    from mr_recon_amd.hip_interp import grid

    G = 40
    g = torch.Generator().manual_seed(D)
    pts = torch.rand((12, D), generator=g) * G
    k = torch.cat([pts.repeat_interleave(torch.tensor([1, 2, 31, 32, 33, 64, 65, 100, 3, 1, 7, 200]), dim=0),
                   pts[:4] + 1e-3 * torch.rand((4, D), generator=g), torch.rand((500, D), generator=g) * G])
    k = k[torch.randperm(k.shape[0], generator=g)]
    vals = torch.randn((3, k.shape[0]), dtype=torch.complex64, generator=g)
    from mr_recon_amd.plan import SamplePlan

    ref = interp_ref.grid(vals, k, (G,) * D, 4.0, beta=6.0)
    for S in (64, 256, None):  # items that cut runs at different places
        plan = SamplePlan(k.to(dev), (G,) * D, 4.0, 6.0, samples_per_item=S)
        out = grid(vals.to(dev), k.to(dev), (G,) * D, 4.0, kernel_params={"beta": 6.0}, plan=plan)
        assert rel_l2(out, ref) < TOL


def test_plan_items_cover_every_sample(dev):
    """Work items: per tile contiguous, each at most S samples, together exactly the sorted range."""
    from mr_recon_amd.plan import SamplePlan

    k = torch.from_numpy(edge_coords((60, 60), 20000, seed=3))
    k[7, 0] = float("nan")
    plan = SamplePlan(k.to(dev), (60, 60), 4.0, 6.0, samples_per_item=300)
    items = plan.items[: plan.nitems].cpu().numpy()
    ti = plan.tile_items.cpu().numpy()
    start = plan.tile_start.cpu().numpy()
    assert ti[-1] == plan.nitems
    cover = np.zeros(plan.npts, dtype=np.int64)
    for t in range(plan.ntiles + 1):
        its = items[ti[t]:ti[t + 1]]
        assert np.all(its[:, 1] == (t if t < plan.ntiles else -1))
        assert np.all((its[:, 3] > 0) & (its[:, 3] <= 300))
        if len(its):
            assert its[0, 2] == start[t] and its[-1, 2] + its[-1, 3] == start[t + 1]
        for it in its:
            cover[it[2]:it[2] + it[3]] += 1
    assert np.all(cover == 1)


def test_ungrid_bruteforce_tiny(dev):
    """Tiny grids (G smaller than the patch) against the pure-Python definition."""
    from mr_recon_amd.hip_interp import ungrid

    for D, G, w in [(1, 5, 4.0), (2, 6, 4.0), (3, 5, 4.0), (2, 3, 6.0)]:
        g = torch.Generator().manual_seed(G)
        vals = torch.randn((2, *(G,) * D), dtype=torch.complex64, generator=g)
        k = (torch.rand((40, D), generator=g) - 0.3) * 2 * G
        out = ungrid(vals.to(dev), k.to(dev), w, kernel_params={"beta": 4.0}).cpu()
        ref = interp_ref.ungrid_bruteforce(vals, k, w, 4.0)
        assert rel_l2(out, torch.from_numpy(ref)) < TOL


def test_adjointness(dev):
    from mr_recon_amd.hip_interp import grid, ungrid

    for D, G in [(2, 50), (3, 18)]:
        g = torch.Generator().manual_seed(3)
        x = torch.randn((2, *(G,) * D), dtype=torch.complex64, generator=g)
        k = (torch.rand((3000, D), generator=g) * 2 - 0.5) * G
        y = torch.randn((2, 3000), dtype=torch.complex64, generator=g)
        ax = ungrid(x.to(dev), k.to(dev), 4.0, kernel_params={"beta": 6.0})
        aty = grid(y.to(dev), k.to(dev), (G,) * D, 4.0, kernel_params={"beta": 6.0})
        lhs, rhs = cdot(y, ax), cdot(aty, x)
        assert abs(lhs - rhs) / abs(lhs) < 1e-5


def test_nonfinite_coords_give_zero(dev):
    from mr_recon_amd.hip_interp import grid, ungrid

    vals = torch.randn((2, 32, 32), dtype=torch.complex64)
    k = torch.rand((100, 2)) * 32
    k[5, 0] = float("nan")
    k[17, 1] = float("inf")
    out = ungrid(vals.to(dev), k.to(dev), 4.0).cpu()
    assert torch.all(out[:, 5] == 0) and torch.all(out[:, 17] == 0)
    ok = torch.ones(100, dtype=torch.bool)
    ok[[5, 17]] = False
    ref = interp_ref.ungrid(vals, k[ok], 4.0)
    assert rel_l2(out[:, ok], ref) < TOL
    y = torch.randn((2, 100), dtype=torch.complex64)
    gr = grid(y.to(dev), k.to(dev), (32, 32), 4.0).cpu()
    ref_g = interp_ref.grid(y[:, ok], k[ok], (32, 32), 4.0)
    assert rel_l2(gr, ref_g) < TOL


def test_real_and_batched_shapes(dev):
    from mr_recon_amd.hip_interp import grid, ungrid

    vals = torch.randn((2, 3, 40, 40))
    k = (torch.rand((7, 11, 2)) - 0.5) * 40
    out = ungrid(vals.to(dev), k.to(dev), 4.0)
    assert out.shape == (2, 3, 7, 11) and not torch.is_complex(out)
    ref = interp_ref.ungrid(vals.to(torch.complex64), k, 4.0).real
    assert rel_l2(out, ref) < TOL
    y = torch.randn((2, 3, 7, 11))
    gr = grid(y.to(dev), k.to(dev), (40, 40), 4.0)
    assert gr.shape == (2, 3, 40, 40)
    ref = interp_ref.grid(y.to(torch.complex64), k, (40, 40), 4.0).real
    assert rel_l2(gr, ref) < TOL


def test_dense_hot_tile_split(dev):
    """Many samples in one tile (radial centre): items split the tile; results unchanged."""
    from mr_recon_amd.hip_interp import grid, ungrid

    g = torch.Generator().manual_seed(11)
    k = 30 + torch.rand((20000, 2), generator=g) * 3  # all inside one tile
    vals = torch.randn((3, 64, 64), dtype=torch.complex64, generator=g)
    out = ungrid(vals.to(dev), k.to(dev), 6.0, kernel_params={"beta": 9.0})
    assert rel_l2(out, interp_ref.ungrid(vals, k, 6.0, beta=9.0)) < TOL
    y = torch.randn((3, 20000), dtype=torch.complex64, generator=g)
    gr = grid(y.to(dev), k.to(dev), (64, 64), 6.0, kernel_params={"beta": 9.0})
    assert rel_l2(gr, interp_ref.grid(y, k, (64, 64), 6.0, beta=9.0)) < TOL


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("G,width", [(72, 4.0), (50, 6.0), (37, 3.0)])
def test_2d_tile_edges(dev, tile, G, width):
    """The 2D kernels compiled for tile edges 8, 16 and 32: plans bit-exact against the
    oracle's binning with that edge (grid sizes that are and are not multiples of it), and
    gather / gridding against the oracle, a hot tile included."""
    from mr_recon_amd.hip_interp import grid, ungrid
    from mr_recon_amd.plan import SamplePlan

    k = _case(2, G, width, 2500, seed=G + tile)
    rng = np.random.default_rng(tile)
    hot = torch.from_numpy((G / 2 + rng.random((1500, 2)) * 3).astype(np.float32))
    k = torch.cat([k, hot])
    plan = SamplePlan(k.to(dev), (G, G), width, 6.0, tile=tile, samples_per_item=700)
    ts, perm, start, _ = obins.bin_plan(k.numpy(), (G, G), width, tile)
    assert plan.tile == tile and plan.subgrid == (tile + plan.P - 1) ** 2
    assert np.array_equal(plan.perm.cpu().numpy(), perm)
    assert np.array_equal(plan.tsort.cpu().numpy().view(np.int32), ts.view(np.int32))
    assert np.array_equal(plan.tile_start.cpu().numpy(), start)
    from mr_recon_amd import _lib, ops

    g = torch.Generator().manual_seed(tile)
    vals = torch.randn((3, G, G), dtype=torch.complex64, generator=g)
    samp = torch.randn((3, k.shape[0]), dtype=torch.complex64, generator=g)
    out = ops.interp(vals.to(dev), plan)
    assert rel_l2(out, interp_ref.ungrid(vals, k, width, beta=6.0)) < TOL
    gout = torch.empty((3, G, G), dtype=torch.complex64, device=dev)
    ops.gridding(samp.to(dev), plan, gout)
    assert rel_l2(gout, interp_ref.grid(samp, k, (G, G), width, beta=6.0)) < TOL
    with _lib.tile_scope(2, tile):  # the plain surface plans with the thread's edge
        assert rel_l2(ungrid(vals.to(dev), k.to(dev), width, kernel_params={"beta": 6.0}), out) < 1e-6
        assert rel_l2(grid(samp.to(dev), k.to(dev), (G, G), width, kernel_params={"beta": 6.0}), gout) < 1e-6
