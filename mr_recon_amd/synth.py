"""Seeded synthetic MRI inputs: trajectories, coil maps, phantoms.

All coordinates are in the reference's units (cycles per FOV, k in
[-N/2, N/2)); all generators are deterministic given ``seed`` and run on the
CPU (move the results to the device yourself).
"""

from __future__ import annotations

import math

import numpy as np
import torch

GOLDEN_ANGLE_DEG = 111.246


def radial_2d(nspokes: int, nread: int, N: int) -> torch.Tensor:
    """Golden-angle radial spokes (nspokes, nread, 2); readout k_r = (i - nread//2) * N / nread."""
    th = np.deg2rad(GOLDEN_ANGLE_DEG) * np.arange(nspokes)
    kr = (np.arange(nread) - nread // 2) * (N / nread)
    k = np.stack([np.cos(th)[:, None] * kr[None], np.sin(th)[:, None] * kr[None]], axis=-1)
    return torch.from_numpy(k.astype(np.float32))


def spiral_2d(ninterleaves: int, nsamp: int, N: int, alpha: float = 1.5, turns: float | None = None) -> torch.Tensor:
    """Variable-density Archimedean-type spiral (ninterleaves, nsamp, 2), |k| <= N/2 * (1 - 1/nsamp).

    k(tau) = N/2 * tau^alpha * exp(i (2 pi n_turns tau + 2 pi j / ninterleaves)).
    """
    if turns is None:
        turns = N / (2.0 * ninterleaves)
    tau = np.arange(nsamp) / nsamp
    r = (N / 2) * tau ** alpha
    out = []
    for j in range(ninterleaves):
        ph = 2 * np.pi * turns * tau + 2 * np.pi * j / ninterleaves
        out.append(np.stack([r * np.cos(ph), r * np.sin(ph)], axis=-1))
    k = np.stack(out, 0)
    k = np.clip(k, -N / 2, N / 2 - 1e-3)
    return torch.from_numpy(k.astype(np.float32))


def fibonacci_sphere(n: int) -> np.ndarray:
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = np.pi * (3 - math.sqrt(5)) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=-1)


def kooshball_3d(nspokes: int, nread: int, N: int) -> torch.Tensor:
    """3D radial spokes along Fibonacci-sphere directions (nspokes, nread, 3)."""
    dirs = fibonacci_sphere(nspokes)
    kr = (np.arange(nread) - nread // 2) * (N / nread)
    k = dirs[:, None, :] * kr[None, :, None]
    return torch.from_numpy(k.astype(np.float32))


def random_image(shape, seed: int = 0) -> torch.Tensor:
    """i.i.d. complex Gaussian, re/im ~ N(0, 1/2)."""
    g = torch.Generator().manual_seed(seed)
    re = torch.randn(shape, generator=g) / math.sqrt(2)
    im = torch.randn(shape, generator=g) / math.sqrt(2)
    return torch.complex(re, im)


def coil_maps(ncoils: int, im_size, seed: int = 0, radius: float = 0.35, sigma: float = 0.3) -> torch.Tensor:
    """Smooth maps: complex Gaussian bumps centred on a circle (2D) / Fibonacci sphere (3D)
    of ``radius`` FOV, width ``sigma`` FOV, random constant phase, normalised to
    root-sum-of-squares 1 per voxel."""
    d = len(im_size)
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[(np.arange(n) - n // 2) / n for n in im_size], indexing="ij")
    r = np.stack(grids, -1)
    if d == 2:
        ang = 2 * np.pi * np.arange(ncoils) / ncoils
        cen = radius * np.stack([np.cos(ang), np.sin(ang)], -1)
    elif d == 3:
        cen = radius * fibonacci_sphere(ncoils)
    else:
        cen = radius * np.linspace(-1, 1, ncoils)[:, None]
    maps = []
    for c in range(ncoils):
        dist2 = np.sum((r - cen[c]) ** 2, axis=-1)
        ph = rng.uniform(0, 2 * np.pi)
        maps.append(np.exp(-dist2 / (2 * sigma ** 2)) * np.exp(1j * ph))
    m = np.stack(maps, 0)
    m /= np.sqrt(np.sum(np.abs(m) ** 2, axis=0, keepdims=True))
    return torch.from_numpy(m.astype(np.complex64))


# (intensity, a, b, x0, y0, phi_deg): own ten-ellipse head-like table
_ELLIPSES = [
    (1.0, 0.69, 0.92, 0.0, 0.0, 0),
    (-0.8, 0.6624, 0.874, 0.0, -0.0184, 0),
    (-0.2, 0.11, 0.31, 0.22, 0.0, -18),
    (-0.2, 0.16, 0.41, -0.22, 0.0, 18),
    (0.1, 0.21, 0.25, 0.0, 0.35, 0),
    (0.1, 0.046, 0.046, 0.0, 0.1, 0),
    (0.1, 0.046, 0.046, 0.0, -0.1, 0),
    (0.1, 0.046, 0.023, -0.08, -0.605, 0),
    (0.1, 0.023, 0.023, 0.0, -0.606, 0),
    (0.1, 0.023, 0.046, 0.06, -0.605, 0),
]


def phantom_2d(N: int, seed: int = 0, random_phase: bool = True) -> torch.Tensor:
    """Shepp-Logan-like magnitude times a smooth random phase (low-order polynomial)."""
    y, x = np.meshgrid((np.arange(N) - N // 2) / (N / 2), (np.arange(N) - N // 2) / (N / 2), indexing="ij")
    img = np.zeros((N, N))
    for (a0, a, b, x0, y0, phi) in _ELLIPSES:
        p = np.deg2rad(phi)
        xr = (x - x0) * np.cos(p) + (y - y0) * np.sin(p)
        yr = -(x - x0) * np.sin(p) + (y - y0) * np.cos(p)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1] += a0
    phase = np.zeros_like(img)
    if random_phase:
        rng = np.random.default_rng(seed)
        c = rng.normal(scale=0.5, size=6)
        phase = c[0] + c[1] * x + c[2] * y + c[3] * x * y + c[4] * x ** 2 + c[5] * y ** 2
    return torch.from_numpy((img * np.exp(1j * phase)).astype(np.complex64))


def complex_noise(shape, sigma: float, seed: int = 0) -> torch.Tensor:
    g = torch.Generator().manual_seed(seed)
    return torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g)) * (sigma / math.sqrt(2))
