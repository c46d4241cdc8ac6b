"""Pins the Pipe-Menon density-compensation oracle (no GPU): a dense G^H G built from
the kernel definition alone, the Cartesian closed form, and invariances the
iteration must have (duplicated samples, integer shifts, sample order)."""

import numpy as np
import pytest
import torch
from scipy.special import i0

from helpers import rel_l2
from mr_recon_amd import synth
from oracle import dcf_ref


def _radial_2d(nspokes, nread, N, seed=0):
    """Radial spokes in grid units -N/2..N/2 (dense centre, sparse edge)."""
    rng = np.random.default_rng(seed)
    ang = np.pi * np.arange(nspokes) / nspokes + 0.01 * rng.random(nspokes)
    r = (np.arange(nread) - nread / 2) / nread * N
    k = np.stack([np.outer(np.cos(ang), r), np.outer(np.sin(ang), r)], -1)
    return torch.from_numpy(k.astype(np.float32))


@pytest.mark.parametrize("D,G,M", [(1, 12, 9), (2, 10, 40), (3, 6, 25)])
def test_pipe_menon_vs_dense_gram(D, G, M):
    """The oracle's iteration (G and G^H from the gridding / interpolation restatement)
    equals the same iteration on A A^T, A built entry by entry from the kernel."""
    rng = np.random.default_rng(D)
    k = torch.from_numpy(((rng.random((M, D)) - 0.5) * 1.4 * G).astype(np.float32))
    # coordinates on a 2^-10 lattice wrap mod G without rounding: identical iterates
    kq = torch.round(k * 1024) / 1024
    for it in (1, 2, 7):
        a = dcf_ref.pipe_menon(kq, (G,) * D, max_iter=it, normalize=False)
        b = dcf_ref.pipe_menon_dense(kq, (G,) * D, max_iter=it, normalize=False)
        assert rel_l2(a, torch.from_numpy(b)) < 1e-12
    # elsewhere the oracle wraps the fp32 coordinate mod G (as the NUFFT gridding does)
    # while the dense form wraps the integer grid index: one fp32 rounding apart
    a = dcf_ref.pipe_menon(k, (G,) * D, max_iter=7, normalize=False)
    b = dcf_ref.pipe_menon_dense(k, (G,) * D, max_iter=7, normalize=False)
    assert rel_l2(a, torch.from_numpy(b)) < 1e-6


@pytest.mark.parametrize("D", [1, 2, 3])
def test_cartesian_closed_form(D):
    """Every grid point sampled once: G^H G 1 = S^(2D) with S = sum_u I0(8 sqrt(1-(u/2)^2))
    over u = -2..2 (scipy's I0), so w_1 = S^(-2D), w is a fixed point after one
    step, and the normalised dcf is all ones."""
    N = 8
    grids = np.meshgrid(*([np.arange(N) - N // 2] * D), indexing="ij")
    k = torch.from_numpy(np.stack(grids, -1).reshape(-1, D).astype(np.float32))
    S = sum(i0(8.0 * np.sqrt(1.0 - (u / 2.0) ** 2)) for u in range(-2, 3))
    w1 = dcf_ref.pipe_menon(k, (N,) * D, max_iter=1, normalize=False)
    assert torch.allclose(w1, torch.full_like(w1, S ** (-2 * D)), rtol=1e-5, atol=0)
    w5 = dcf_ref.pipe_menon(k, (N,) * D, max_iter=5, normalize=False)
    assert torch.allclose(w5, w1, rtol=1e-12, atol=0)
    assert torch.allclose(dcf_ref.pipe_menon(k, (N,) * D, max_iter=5), torch.ones(N ** D, dtype=torch.float64),
                          rtol=1e-12, atol=0)


def test_invariances():
    """Duplicating every sample halves w but not the normalised dcf; an integer shift
    of all samples (grid indices wrap mod N) and a permutation of the samples change
    nothing but the order."""
    N = 16
    k = _radial_2d(12, 16, N)
    base = dcf_ref.pipe_menon(k, (N, N), max_iter=10)
    dup = dcf_ref.pipe_menon(torch.cat([k, k], 0), (N, N), max_iter=10)
    assert rel_l2(dup[:12], base) < 1e-12 and rel_l2(dup[12:], base) < 1e-12
    raw = dcf_ref.pipe_menon(k, (N, N), max_iter=3, normalize=False)
    raw2 = dcf_ref.pipe_menon(torch.cat([k, k], 0), (N, N), max_iter=3, normalize=False)
    assert rel_l2(raw2[:12], raw / 2) < 1e-12
    shifted = dcf_ref.pipe_menon(k + torch.tensor([3.0, -5.0]), (N, N), max_iter=10)
    assert rel_l2(shifted, base) < 1e-6
    perm = torch.randperm(k.reshape(-1, 2).shape[0], generator=torch.Generator().manual_seed(0))
    p = dcf_ref.pipe_menon(k.reshape(-1, 2)[perm], (N, N), max_iter=10)
    assert rel_l2(p, base.reshape(-1)[perm]) < 1e-12


def test_radial_density_and_fp32():
    """A radial trajectory's dcf grows with |k| (the centre is oversampled), and the
    reference's fp32 arithmetic stays within 1e-5 of the fp64 twin after 30
    iterations."""
    N = 32
    k = _radial_2d(24, 32, N)
    d = dcf_ref.pipe_menon(k, (N, N), max_iter=30)
    r = k.norm(dim=-1)
    inner = d[(r > 1) & (r < 4)].mean()
    outer = d[(r > 8) & (r < 14)].mean()
    assert outer > 2.5 * inner
    d32 = dcf_ref.pipe_menon(k, (N, N), max_iter=30, dtype=torch.float32)
    assert rel_l2(d32, d) < 1e-5


def test_kooshball_fp32_close_to_fp64():
    N = 16
    k = synth.kooshball_3d(40, 16, N).float()
    d = dcf_ref.pipe_menon(k, (N,) * 3, max_iter=30)
    d32 = dcf_ref.pipe_menon(k, (N,) * 3, max_iter=30, dtype=torch.float32)
    assert rel_l2(d32, d) < 1e-5
