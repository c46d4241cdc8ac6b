"""GPU parity of the batched multi-problem path against the plain single-problem path
(which is itself checked against the oracle), including HIP-graph capture and replay."""

import pytest
import torch

from helpers import rel_l2
from mr_recon_amd import synth

pytestmark = pytest.mark.gpu


def _problems(dev, seed=0):
    from mr_recon_amd.batched import Problem

    spec = [(32, 2, "spiral", 900), (48, 3, "radial", 1500), (32, 4, "radial", 700), (48, 2, "spiral", 2200),
            (32, 3, "spiral", 1300), (64, 2, "radial", 1800)]
    probs, ksp = [], []
    for i, (N, C, kind, M) in enumerate(spec):
        if kind == "spiral":
            trj = synth.spiral_2d(4, M // 4, N).reshape(-1, 2)
        else:
            trj = synth.radial_2d(M // (2 * N), 2 * N, N).reshape(-1, 2)
        mps = synth.coil_maps(C, (N, N), seed=seed + i)
        dcf = torch.rand(trj.shape[0], generator=torch.Generator().manual_seed(i)) + 0.5
        probs.append(Problem(trj.to(dev), mps.to(dev), dcf.to(dev)))
        ksp.append(synth.random_image((C, trj.shape[0]), seed=seed + 100 + i).to(dev))
    return probs, ksp


def _plain(pr, dev):
    from mr_recon_amd.fourier import hip_nufft
    from mr_recon_amd.linops import sense_linop

    N = pr.mps.shape[1]
    return sense_linop(pr.trj, pr.mps, dcf=pr.dcf, nufft=hip_nufft((N, N), oversamp=1.25, width=4))


def test_batched_operators_match_plain(dev):
    from mr_recon_amd.batched import BatchedSense

    probs, ksp = _problems(dev)
    B = BatchedSense(probs, oversamp=1.25, width=4)
    assert len(B.groups) == 3
    xs = [synth.random_image(tuple(pr.mps.shape[1:]), seed=i).to(dev) for i, pr in enumerate(probs)]
    Y = B.forward(B.stack_images(xs))  # one interp launch over all three image sizes
    Ys = B.stack_ksp([synth.random_image(k.shape, seed=i + 7).to(dev) for i, k in enumerate(ksp)])
    Z = B.unstack_images(B.adjoint(Ys))  # one gridding launch over all three image sizes
    for i, pr in enumerate(probs):
        A = _plain(pr, dev)
        sl = B.ksp_slice(i)
        assert rel_l2(Y[sl], A.forward(xs[i]).reshape(-1)) < 1e-5
        assert rel_l2(Z[i], A.adjoint(Ys[sl].reshape(ksp[i].shape))) < 1e-5


@pytest.mark.parametrize("tile", [8, 32])
def test_batched_tile_edges_match_plain(dev, tile):
    """A batch planned with 2D tile edge 8 or 32 (one edge for every problem, one launch
    serves them all) against plain operators with the same edge."""
    from mr_recon_amd.batched import BatchedSense
    from mr_recon_amd.fourier import hip_nufft
    from mr_recon_amd.linops import sense_linop

    probs, ksp = _problems(dev, seed=7)
    B = BatchedSense(probs, oversamp=1.25, width=4, tile=tile)
    assert B.tile == tile and all(p.tile == tile for g in B.groups for p in g.plans)
    xs = [synth.random_image(tuple(pr.mps.shape[1:]), seed=40 + i).to(dev) for i, pr in enumerate(probs)]
    Y = B.forward(B.stack_images(xs))
    Z = B.unstack_images(B.adjoint(B.stack_ksp(ksp)))
    for i, pr in enumerate(probs):
        N = pr.mps.shape[1]
        A = sense_linop(pr.trj, pr.mps, dcf=pr.dcf, nufft=hip_nufft((N, N), oversamp=1.25, width=4, tile=tile))
        assert rel_l2(Y[B.ksp_slice(i)], A.forward(xs[i]).reshape(-1)) < 1e-5
        assert rel_l2(Z[i], A.adjoint(ksp[i])) < 1e-5


def test_batched_tile_autotune(dev):
    from mr_recon_amd import autotune
    from mr_recon_amd.batched import BatchedSense

    probs, _ = _problems(dev, seed=8)
    B = BatchedSense(probs, oversamp=1.25, width=4, tile="auto")
    assert B.tile in autotune.TILES[2]


def test_batched_normal_keeps_ksp_in_plan_order(dev):
    """The batched normal operator keeps k-space in plan order: equal to adjoint(forward)."""
    from mr_recon_amd.batched import BatchedSense

    probs, _ = _problems(dev, seed=5)
    B = BatchedSense(probs, oversamp=1.25, width=4)
    xs = [synth.random_image(tuple(pr.mps.shape[1:]), seed=20 + i).to(dev) for i, pr in enumerate(probs)]
    Xs = B.stack_images(xs)
    got = B.normal(Xs, 0.5)
    ref = B.adjoint(B.forward(Xs), 0.5)
    for a, b in zip(got, ref):
        assert torch.equal(a, b)


def test_batched_cg_matches_loop_of_plain_calls(dev):
    from mr_recon_amd.batched import BatchedSense
    from mr_recon_amd.recons import CG_SENSE_recon

    probs, ksp = _problems(dev, seed=3)
    B = BatchedSense(probs, oversamp=1.25, width=4)
    kw = dict(num_iters=6, lamda_l2=1e-3, max_eigen=3.0)
    out = B.cg_sense(ksp, **kw)
    for i, pr in enumerate(probs):
        ref = CG_SENSE_recon(_plain(pr, dev), ksp[i], max_iter=6, lamda_l2=1e-3, max_eigen=3.0, tolerance=0.0,
                             verbose=False)
        assert rel_l2(out[i], ref) < 1e-4


def test_batched_cg_graph_replays(dev):
    """The captured graph replays correctly several times, with eager work and new inputs in between."""
    from mr_recon_amd.batched import BatchedSense

    probs, ksp = _problems(dev, seed=5)
    B = BatchedSense(probs, oversamp=1.25, width=4)
    kw = dict(num_iters=4, lamda_l2=1e-3, max_eigen=3.0)
    B.capture(**kw)
    for rep in range(3):
        ks = [k * (1.0 + 0.5 * rep) + 0.01 * rep for k in ksp]
        got = [t.clone() for t in B.replay(ks)]
        torch.cuda.synchronize()
        eager = B.cg_sense(ks, **kw)
        _ = sum(float(e.abs().sum()) for e in eager)  # eager work + host sync between replays
        for a, b in zip(got, eager):
            assert rel_l2(a, b) < 1e-5
