"""GPU parity of density_compensation (Pipe-Menon on the HIP gridding and
interpolation kernels) against the oracle in the reference's fp32 arithmetic.

fp32, not fp64: the Kaiser-Bessel kernel is 1, not 0, at the edge of its support
(|u| = 1), so a sample within ~1e-14 of a grid integer (a spiral's points on the
axes, k = (-16, 1.8e-14)) keeps or loses one edge point depending on whether
ceil(k - 2) is taken in fp32 (the reference, sigpy and the HIP plan) or fp64;
30 iterations carry that to ~6e-4 on the spiral below."""

import numpy as np
import pytest
import torch

from helpers import rel_l2
from oracle import dcf_ref

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _radial_2d(nspokes, nread, N, seed=0):
    rng = np.random.default_rng(seed)
    ang = np.pi * np.arange(nspokes) / nspokes + 0.01 * rng.random(nspokes)
    r = (np.arange(nread) - nread / 2) / nread * N
    k = np.stack([np.outer(r, np.cos(ang)), np.outer(r, np.sin(ang))], -1)
    return torch.from_numpy(k.astype(np.float32))


def _spiral_2d(ninter, nread, N):
    t = np.linspace(0, 1, nread)[:, None]
    ph = 2 * np.pi * np.arange(ninter)[None] / ninter
    r = t * N / 2
    th = 12 * np.pi * t + ph
    return torch.from_numpy(np.stack([r * np.cos(th), r * np.sin(th)], -1).astype(np.float32))


@pytest.mark.parametrize("kind", ["radial", "spiral"])
def test_dcf_2d_vs_oracle(dev, kind):
    from mr_recon_amd.algs import density_compensation

    N = 48
    k = _radial_2d(40, 64, N) if kind == "radial" else _spiral_2d(6, 400, N)
    ref = dcf_ref.pipe_menon(k, (N, N), max_iter=30, dtype=torch.float32)
    out = density_compensation(k.to(dev), (N, N))
    assert out.shape == k.shape[:-1] and out.dtype == torch.float32
    assert rel_l2(out, ref) < TOL
    assert float(out.max()) == 1.0


def test_dcf_3d_and_1d_vs_oracle(dev):
    from mr_recon_amd import synth
    from mr_recon_amd.algs import density_compensation

    N = 20
    k = synth.kooshball_3d(120, 24, N).float()
    ref = dcf_ref.pipe_menon(k, (N,) * 3, max_iter=12, dtype=torch.float32)
    assert rel_l2(density_compensation(k.to(dev), (N,) * 3, num_iters=12), ref) < TOL
    k1 = torch.from_numpy(np.sort(np.random.default_rng(1).normal(0, 6, 300)).astype(np.float32))[:, None]
    ref1 = dcf_ref.pipe_menon(k1, (32,), max_iter=30, dtype=torch.float32)
    assert rel_l2(density_compensation(k1.to(dev), (32,)), ref1) < TOL


def test_dcf_cartesian_and_zero_iters(dev):
    from mr_recon_amd.algs import density_compensation

    N = 16
    g = np.meshgrid(np.arange(N) - N // 2, np.arange(N) - N // 2, indexing="ij")
    k = torch.from_numpy(np.stack(g, -1).astype(np.float32)).to(dev)
    d = density_compensation(k, (N, N), num_iters=5)
    assert torch.allclose(d, torch.ones_like(d), rtol=1e-6, atol=0)
    k2 = _radial_2d(8, 16, N).to(dev)
    assert torch.equal(density_compensation(k2, (N, N), num_iters=0), torch.ones(16, 8, device=dev))


def test_dcf_bad_arguments(dev):
    from mr_recon_amd.algs import density_compensation

    k = _radial_2d(8, 16, 16).to(dev)
    with pytest.raises(ValueError):
        density_compensation(k, (16, 16), method="cg_ksp")
    with pytest.raises(ValueError):
        density_compensation(k, (16, 16, 16))
