"""Pins the type-3 NUFFT oracle (no GPU) against the reference's own exact-sum class
(type3_nufft_naive, linops.py:49-121, restated as naive_forward / naive_adjoint):
the KB construction must reproduce y_k = sum_n x_n exp(-2 pi i phi_n . alpha_k) to
the kernel's approximation error, be adjoint-consistent, and its Toeplitz normal
must equal adjoint(forward)."""

import pytest
import torch

from helpers import cdot, rel_l2
from oracle import type3_ref


def _case(D, n, T, seed):
    g = torch.Generator().manual_seed(seed)
    axes = torch.meshgrid(*([torch.arange(n) - n / 2] * D), indexing="ij")
    phis = torch.stack(axes, 0) / n + 0.05 * torch.randn((D, *(n,) * D), generator=g)
    alphas = (torch.rand((D, T), generator=g) - 0.5) * n
    x = torch.randn((2, *(n,) * D), dtype=torch.complex64, generator=g)
    return phis, alphas, x


@pytest.mark.parametrize("D,n,T,w,bound", [(2, 12, 300, 4.0, 2e-3), (2, 12, 300, 6.0, 4e-4), (1, 40, 120, 4.0, 2e-3),
                                           (3, 6, 200, 4.0, 2e-3)])
def test_type3_oracle_vs_exact_sum(D, n, T, w, bound):
    phis, alphas, x = _case(D, n, T, seed=D)
    A = type3_ref.Type3Ref(phis, alphas, 2.0, w)
    assert rel_l2(A.forward(x), type3_ref.naive_forward(phis, alphas, x)) < bound
    y = torch.randn((2, T), dtype=torch.complex64, generator=torch.Generator().manual_seed(1))
    assert rel_l2(A.adjoint(y), type3_ref.naive_adjoint(phis, alphas, y)) < bound
    fx, ay = A.forward(x), A.adjoint(y)
    assert abs(cdot(y, fx) - cdot(ay, x)) < 1e-5 * abs(cdot(y, fx))


def test_type3_oracle_toeplitz():
    phis, alphas, x = _case(2, 12, 300, seed=3)
    A = type3_ref.Type3Ref(phis, alphas, 2.0, 4.0, use_toep=True)
    assert rel_l2(A.normal(x), A.adjoint(A.forward(x))) < 5e-4
