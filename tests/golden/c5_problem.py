"""The trajectory-gradient workload (2D 256^2 image, 16 coils, 64 spiral interleaves x 2048
samples, Kaiser-Bessel width 6, oversampling 1.25) and its loss restated in FP64 on the
oracle:

    L(x, k) = ||A(k) x - y||^2 + ||CG_10(A(k), y) - x0||^2

A(k) = SENSE forward (coil maps times the reference NUFFT, fourier.py:354-582);
CG_10 = 10 CG iterations on (A^H A / max_eigen + lamda I) x = A^H y / sqrt(max_eigen)
started at A^H y / sqrt(max_eigen) (recons.py:77-156, algs.py:779-872), lamda 1e-3,
max_eigen 1. Shared by the fixture script (make_c5_fd.py) and the GPU test, so both
build bit-identical inputs.
"""

from __future__ import annotations

import math

import numpy as np
import torch

N, C, NINT, NSAMP = 256, 16, 64, 2048
OS, W = 1.25, 6
LAMDA, MAX_EIGEN, ITERS = 1e-3, 1.0, 10
H = 1e-4  # central-difference step (trajectory units: cycles per FOV; image: value units)


def nufft64():
    from oracle.nufft_ref import TritonNufftRef

    return TritonNufftRef((N, N), oversamp=OS, width=W, dtype=torch.float64, coords64=True)


def sense_fwd64(nf, x, k, mps):
    """(N, N) complex128, (M, 2) float64, (C, N, N) -> (C, M)."""
    return nf.forward((mps * x)[None], k[None])[0]


def sense_adj64(nf, y, k, mps):
    return torch.sum(nf.adjoint(y[None], k[None])[0] * mps.conj(), dim=0)


def build_inputs():
    """Seeded inputs; y = A(k0) x0 (FP64 oracle) + 1 % complex noise, rounded to complex64."""
    from mr_recon_amd import synth

    k0 = synth.spiral_2d(NINT, NSAMP, N).reshape(-1, 2)
    mps = synth.coil_maps(C, (N, N), seed=5)
    x0 = synth.phantom_2d(N, seed=5)
    nf = nufft64()
    y = sense_fwd64(nf, x0.to(torch.complex128), k0.to(torch.float64), mps.to(torch.complex128))
    noise = synth.random_image(tuple(y.shape), seed=9).to(torch.complex128)
    y = (y + 0.01 * noise * y.abs().mean()).to(torch.complex64)
    return {"k0": k0, "mps": mps, "x0": x0, "y": y, "xin": (x0 * 0.9).contiguous(),
            "kin": (k0 + 0.05).contiguous()}


def loss64(inp, x=None, k=None):
    """L at (x, k) (defaults: the perturbed starting point xin, kin) in FP64."""
    nf = nufft64()
    x = (inp["xin"] if x is None else x).to(torch.complex128)
    k = (inp["kin"] if k is None else k).to(torch.float64)
    mps = inp["mps"].to(torch.complex128)
    y = inp["y"].to(torch.complex128)
    x0 = inp["x0"].to(torch.complex128)
    dc = torch.sum(torch.abs(sense_fwd64(nf, x, k, mps) - y) ** 2)

    def normal(v):
        return sense_adj64(nf, sense_fwd64(nf, v, k, mps), k, mps) / MAX_EIGEN + LAMDA * v

    ahb = sense_adj64(nf, y, k, mps) / math.sqrt(MAX_EIGEN)
    xx = ahb
    r = ahb - normal(xx)
    p = r
    rz = torch.real(torch.vdot(r.reshape(-1), r.reshape(-1)))
    for _ in range(ITERS):
        ap = normal(p)
        a = rz / torch.real(torch.vdot(p.reshape(-1), ap.reshape(-1)))
        xx = xx + a * p
        r = r - a * ap
        rzn = torch.real(torch.vdot(r.reshape(-1), r.reshape(-1)))
        p = r + (rzn / rz) * p
        rz = rzn
    return float(dc + torch.sum(torch.abs(xx - x0) ** 2))


def smooth_direction(k, v):
    """Zero the components of v that would move a sample across a Kaiser-Bessel support
    edge within +-H (the reference kernel is discontinuous there: KB(1) = I0(0) = 1)."""
    G = math.ceil(OS * N)
    s = G / N
    t = k.to(torch.float64) * s + G // 2
    dl = H * s * v.abs()
    r = W / 2
    same = (torch.ceil(t - r + dl) == torch.ceil(t - r - dl)) & (torch.floor(t + r + dl) == torch.floor(t + r - dl))
    return torch.where(same, v, torch.zeros_like(v))


def k_directions(kin):
    """Named trajectory directions: three dense random ones and two confined to one interleave."""
    out = []
    for s in range(3):
        v = torch.from_numpy(np.random.default_rng(100 + s).standard_normal(tuple(kin.shape)))
        out.append((f"dense{s}", smooth_direction(kin, v)))
    for j in (0, 37):
        v = torch.zeros(tuple(kin.shape), dtype=torch.float64)
        sl = slice(j * NSAMP, (j + 1) * NSAMP)
        v[sl] = torch.from_numpy(np.random.default_rng(200 + j).standard_normal((NSAMP, 2)))
        out.append((f"interleave{j}", smooth_direction(kin, v)))
    return out


def x_direction():
    g = torch.Generator().manual_seed(300)
    return torch.complex(torch.randn((N, N), generator=g, dtype=torch.float64),
                         torch.randn((N, N), generator=g, dtype=torch.float64))
