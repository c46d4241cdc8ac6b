"""Shared helpers for the parity tests."""

import numpy as np
import torch


def rel_l2(a, b) -> float:
    a = torch.as_tensor(a).detach().cpu().to(torch.complex128)
    b = torch.as_tensor(b).detach().cpu().to(torch.complex128)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def cdot(a, b) -> complex:
    """<a, b> = sum conj(a) b in complex128."""
    a = torch.as_tensor(a).detach().cpu().to(torch.complex128).reshape(-1)
    b = torch.as_tensor(b).detach().cpu().to(torch.complex128).reshape(-1)
    return complex(torch.sum(a.conj() * b))


def edge_coords(grid_size, n_random, seed=0, width=4.0):
    """Random grid-unit coordinates plus the edge cases of the wrap / corner arithmetic:
    exact integers, exact half-width offsets, negatives, values beyond G, -0.0 and
    tiny negatives that torch.remainder maps to exactly G."""
    rng = np.random.default_rng(seed)
    D = len(grid_size)
    G = np.array(grid_size, dtype=np.float32)
    pts = [(rng.random((n_random, D)) * 3 - 1) * G]
    special = []
    for g in grid_size:
        special += [0.0, -0.0, -1e-30, -1e-8, float(g), float(g) - 0.5, -float(g), 2.0 * g + 0.25, width / 2,
                    width / 2 + 1e-6, 1.0, 3.0, -1.5, 0.5, float(g) - 1e-4]
    sp = np.array(special, dtype=np.float32)
    for d in range(D):
        block = rng.choice(sp, size=(len(sp), D)).astype(np.float32)
        block[:, d] = sp
        pts.append(block)
    return np.concatenate(pts, 0).astype(np.float32)
