"""Kaiser-Bessel kernel of the reference (triton_interp/kernels.py:21-82, 180-246).

KB(u) = I0(beta * sqrt(1 - u^2)) with I0 from Abramowitz & Stegun 9.8.1 / 9.8.2.
"""

from __future__ import annotations

import numpy as np
import torch

# A&S 9.8.1 (|x| < 3.75), coefficients of t^2, t^4, ..., t^12 with t = x / 3.75
A_SMALL = (3.5156229, 3.0899424, 1.2067492, 0.2659732, 0.0360768, 0.0045813)
# A&S 9.8.2 (x >= 3.75), coefficients of t^0, t^-1, ..., t^-8
A_BIG = (0.39894228, 0.01328592, 0.00225319, -0.00157565, 0.00916281, -0.02057706, 0.02635537, -0.01647633,
         0.00392377)


def i0_as_np(x: np.ndarray) -> np.ndarray:
    """A&S polynomial I0 in float64 (x >= 0)."""
    x = np.asarray(x, dtype=np.float64)
    t = x / 3.75
    small = 1.0 + sum(c * t ** (2 * (i + 1)) for i, c in enumerate(A_SMALL))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        big = np.exp(x) / np.sqrt(x) * sum(c * t ** (-i) for i, c in enumerate(A_BIG))
    return np.where(x < 3.75, small, big)


def di0_du_np(u: np.ndarray, beta: float) -> np.ndarray:
    """d/du of I0_AS(beta sqrt(1-u^2)) for u in [0, 1] (float64), finite at u = 1."""
    u = np.asarray(u, dtype=np.float64)
    z = beta * np.sqrt(np.maximum(1.0 - u * u, 0.0))
    t = z / 3.75
    s = sum(2 * (i + 1) * c * t ** (2 * i) for i, c in enumerate(A_SMALL))
    small = -(beta ** 2) * u * s / 3.75 ** 2
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        v = 3.75 / z
        q = sum(c * v ** i for i, c in enumerate(A_BIG))
        dq = sum(i * c * v ** (i - 1) for i, c in enumerate(A_BIG) if i > 0)
        e = np.exp(z) / np.sqrt(z)
        di0 = e * ((1.0 - 0.5 / z) * q - dq * v / z)
        big = di0 * (-(beta ** 2) * u / z)
    return np.where(z < 3.75, small, big)


def kaiser_bessel_torch(x: torch.Tensor, beta: float) -> torch.Tensor:
    """Elementwise KB of the reference's torch path (kernels.py:21-82), same operation order."""
    x = beta * torch.sqrt(1 - x ** 2)
    smallmask = x < 3.75
    t = x / 3.75
    t2 = t * t
    t4 = t2 * t2
    t6 = t4 * t2
    t8 = t4 * t4
    t10 = t6 * t4
    t12 = t6 * t6
    t3 = t2 * t
    t5 = t3 * t2
    t7 = t5 * t2
    small = (1 + A_SMALL[0] * t2 + A_SMALL[1] * t4 + A_SMALL[2] * t6 + A_SMALL[3] * t8 + A_SMALL[4] * t10
             + A_SMALL[5] * t12)
    big = torch.exp(x) / torch.sqrt(x) * (A_BIG[0] + A_BIG[1] / t + A_BIG[2] / t2 + A_BIG[3] / t3 + A_BIG[4] / t4
                                          + A_BIG[5] / t5 + A_BIG[6] / t6 + A_BIG[7] / t7 + A_BIG[8] / t8)
    return torch.where(smallmask, small, big)


def kaiser_bessel_torch_safe(x: torch.Tensor, beta: float) -> torch.Tensor:
    """Same values (the A&S polynomials of kaiser_bessel_torch), written so that autograd is
    finite everywhere, for finite-difference and gradient checks: points outside |x| <= 1
    are masked before any square root (the reference masks NaNs after it, kernels.py:132,
    which makes autograd produce 0 * NaN); the small-argument branch is evaluated from
    t^2 = beta^2 (1 - x^2) / 3.75^2 directly, a polynomial in x with no square root (its
    derivative stays finite at |x| = 1, where z = 0); the large-argument branch only sees
    arguments >= 3.75."""
    inside = x.abs() <= 1
    xs = torch.where(inside, x, torch.zeros_like(x))
    t2 = (beta * beta / (3.75 * 3.75)) * (1 - xs * xs)
    t4 = t2 * t2
    t6 = t4 * t2
    t8 = t4 * t4
    t10 = t6 * t4
    t12 = t6 * t6
    small = (1 + A_SMALL[0] * t2 + A_SMALL[1] * t4 + A_SMALL[2] * t6 + A_SMALL[3] * t8 + A_SMALL[4] * t10
             + A_SMALL[5] * t12)
    smallmask = t2 < 1.0
    z = 3.75 * torch.sqrt(torch.where(smallmask, torch.ones_like(t2), t2))  # >= 3.75 on the branch used
    t = z / 3.75
    big = torch.exp(z) / torch.sqrt(z) * (A_BIG[0] + A_BIG[1] / t + A_BIG[2] / t ** 2 + A_BIG[3] / t ** 3
                                          + A_BIG[4] / t ** 4 + A_BIG[5] / t ** 5 + A_BIG[6] / t ** 6
                                          + A_BIG[7] / t ** 7 + A_BIG[8] / t ** 8)
    return torch.where(inside, torch.where(smallmask, small, big), torch.zeros_like(x))
