"""GPU parity of the differentiable NUFFT / SENSE: gradients with respect to the image and the
trajectory vs (a) FP64 central finite differences of the CPU oracle and (b) autograd through the
oracle's pure-torch path (restated with NaN-safe masking before the square root)."""

import numpy as np
import pytest
import torch

from helpers import rel_l2
from mr_recon_amd import synth
from oracle.nufft_ref import TritonNufftRef

pytestmark = pytest.mark.gpu


def _problem(seed=0, im=(32, 32), C=4, nint=8, nsamp=64):
    trj = synth.spiral_2d(nint, nsamp, im[0]).reshape(-1, 2)
    g = torch.Generator().manual_seed(seed)
    trj = trj + 0.3 * (torch.rand(trj.shape, generator=g) - 0.5)  # break symmetries
    mps = synth.coil_maps(C, im, seed=seed)
    x = synth.random_image(im, seed=seed + 1)
    y = synth.random_image((C, trj.shape[0]), seed=seed + 2) * 0.1
    return trj, mps, x, y


def _oracle_loss(x, trj, mps, y, os_, w):
    """L = sum |A(k) x - y|^2 with the reference algorithm in FP64 (float64 coordinates)."""
    nf = TritonNufftRef(x.shape, oversamp=os_, width=w, dtype=torch.float64, coords64=True, safe=True)
    ax = torch.stack([nf.forward((mps[c] * x)[None], trj[None])[0] for c in range(mps.shape[0])])
    return torch.sum(torch.abs(ax - y) ** 2)


@pytest.mark.parametrize("os_,w", [(1.25, 6), (2.0, 4)])
def test_forward_gradients_vs_oracle_autograd(dev, os_, w):
    from mr_recon_amd.diff import sense_forward
    from mr_recon_amd.fourier import hip_nufft

    trj, mps, x, y = _problem()
    nf = hip_nufft(x.shape, oversamp=os_, width=w)
    xd = x.to(dev).requires_grad_(True)
    kd = trj.to(dev).requires_grad_(True)
    loss = torch.sum(torch.abs(sense_forward(xd, kd, mps.to(dev), nf) - y.to(dev)) ** 2)
    loss.backward()

    x64 = x.to(torch.complex128).requires_grad_(True)
    k64 = trj.to(torch.float64).requires_grad_(True)
    l64 = _oracle_loss(x64, k64, mps.to(torch.complex128), y.to(torch.complex128), os_, w)
    l64.backward()
    assert abs(float(loss.detach()) - float(l64)) / float(l64) < 1e-4
    assert rel_l2(xd.grad, x64.grad) < 1e-4
    assert rel_l2(kd.grad.to(torch.float64), k64.grad) < 1e-3


def test_forward_gradients_3d_vs_oracle_autograd(dev):
    """3D SENSE forward (kooshball, 3 coils): dL/dx and dL/dk vs FP64 autograd through the
    oracle, the normal function and the adjoint function included through one loss."""
    from mr_recon_amd.diff import sense_adjoint, sense_forward, sense_normal
    from mr_recon_amd.fourier import hip_nufft

    im, C = (12, 12, 12), 3
    trj = synth.kooshball_3d(24, 20, 12).reshape(-1, 3)
    g = torch.Generator().manual_seed(21)
    trj = trj + 0.2 * (torch.rand(trj.shape, generator=g) - 0.5)
    mps = synth.coil_maps(C, im, seed=4)
    x = synth.random_image(im, seed=5)
    y = synth.random_image((C, trj.shape[0]), seed=6) * 0.1
    nf = hip_nufft(im, oversamp=1.25, width=4)
    xd = x.to(dev).requires_grad_(True)
    kd = trj.to(dev).requires_grad_(True)
    md = mps.to(dev)
    loss = torch.sum(torch.abs(sense_forward(xd, kd, md, nf) - y.to(dev)) ** 2)
    # the normal and adjoint functions in the same loss: their k-terms add to dL/dk
    loss = loss + 1e-3 * torch.sum(torch.abs(sense_normal(xd, kd, md, nf)) ** 2)
    loss = loss + 1e-2 * torch.sum(torch.abs(sense_adjoint(y.to(dev), kd, md, nf)) ** 2)
    loss.backward()

    ref = TritonNufftRef(im, oversamp=1.25, width=4, dtype=torch.float64, coords64=True, safe=True)
    m64 = mps.to(torch.complex128)
    x64 = x.to(torch.complex128).requires_grad_(True)
    k64 = trj.to(torch.float64).requires_grad_(True)
    y64 = y.to(torch.complex128)

    def fwd(v):
        return torch.stack([ref.forward((m64[c] * v)[None], k64[None])[0] for c in range(C)])

    def adj(w):
        return sum(m64[c].conj() * ref.adjoint(w[c][None], k64[None])[0] for c in range(C))

    l64 = torch.sum(torch.abs(fwd(x64) - y64) ** 2) + 1e-3 * torch.sum(torch.abs(adj(fwd(x64))) ** 2) \
        + 1e-2 * torch.sum(torch.abs(adj(y64)) ** 2)
    l64.backward()
    assert abs(float(loss.detach()) - float(l64)) / float(l64) < 1e-4
    assert rel_l2(xd.grad, x64.grad) < 1e-4
    assert rel_l2(kd.grad.to(torch.float64), k64.grad) < 1e-3


def _smooth_direction(trj, v, N, os_, w, h):
    """Zero the components of v that would move a sample across a KB support edge within
    +-h: the reference kernel is discontinuous there (KB(1) = I0(0) = 1, not 0), so a
    central difference straddling an edge measures the jump, not the derivative."""
    import math

    G = math.ceil(os_ * N)
    s = G / N
    t = trj.to(torch.float64) * s + G // 2
    dl = h * s * v.abs()
    r = w / 2
    same = (torch.ceil(t - r + dl) == torch.ceil(t - r - dl)) & (torch.floor(t + r + dl) == torch.floor(t + r - dl))
    return torch.where(same, v, torch.zeros_like(v))


def test_trajectory_gradient_vs_fp64_finite_differences(dev):
    from mr_recon_amd.diff import sense_forward
    from mr_recon_amd.fourier import hip_nufft

    trj, mps, x, y = _problem(seed=3)
    nf = hip_nufft(x.shape, oversamp=1.25, width=6)
    kd = trj.to(dev).requires_grad_(True)
    loss = torch.sum(torch.abs(sense_forward(x.to(dev), kd, mps.to(dev), nf) - y.to(dev)) ** 2)
    loss.backward()
    gk = kd.grad.cpu().to(torch.float64)
    rng = np.random.default_rng(0)
    h = 1e-4
    x64, m64, y64 = x.to(torch.complex128), mps.to(torch.complex128), y.to(torch.complex128)
    for _ in range(3):
        v = _smooth_direction(trj, torch.from_numpy(rng.standard_normal(trj.shape)), 32, 1.25, 6, h)
        k64 = trj.to(torch.float64)
        lp = float(_oracle_loss(x64, k64 + h * v, m64, y64, 1.25, 6))
        lm = float(_oracle_loss(x64, k64 - h * v, m64, y64, 1.25, 6))
        fd = (lp - lm) / (2 * h)
        an = float(torch.sum(gk * v))
        assert abs(an - fd) / abs(fd) < 2e-3, (an, fd)


def test_adjoint_gradients_vs_oracle_autograd(dev):
    from mr_recon_amd.diff import sense_adjoint
    from mr_recon_amd.fourier import hip_nufft

    trj, mps, x, y = _problem(seed=5)
    dcf = torch.rand(trj.shape[0], generator=torch.Generator().manual_seed(1)) + 0.5
    target = synth.random_image(x.shape, seed=9)
    nf = hip_nufft(x.shape, oversamp=1.25, width=6)
    yd = y.to(dev).requires_grad_(True)
    kd = trj.to(dev).requires_grad_(True)
    out = sense_adjoint(yd, kd, mps.to(dev), nf, dcf=dcf.to(dev))
    loss = torch.sum(torch.abs(out - target.to(dev)) ** 2)
    loss.backward()

    ref = TritonNufftRef(x.shape, oversamp=1.25, width=6, dtype=torch.float64, coords64=True, safe=True)
    y64 = y.to(torch.complex128).requires_grad_(True)
    k64 = trj.to(torch.float64).requires_grad_(True)
    m64 = mps.to(torch.complex128)
    img = sum(ref.adjoint((y64[c] * dcf)[None, None], k64[None])[0, 0] * m64[c].conj() for c in range(4))
    l64 = torch.sum(torch.abs(img - target.to(torch.complex128)) ** 2)
    l64.backward()
    assert rel_l2(yd.grad, y64.grad) < 1e-4
    assert rel_l2(kd.grad.to(torch.float64), k64.grad) < 1e-3


def test_unrolled_cg_gradients_checkpointed(dev):
    """Unrolled 5-iteration CG-SENSE: checkpointed and plain backward agree, and the trajectory
    gradient matches FP64 central differences of the same unrolled recon on the oracle."""
    from mr_recon_amd.diff import unrolled_cg_sense
    from mr_recon_amd.fourier import hip_nufft

    trj, mps, x, _ = _problem(seed=7, im=(24, 24), C=3, nint=6, nsamp=60)
    nf = hip_nufft(x.shape, oversamp=1.25, width=6)
    md = mps.to(dev)
    with torch.no_grad():
        from mr_recon_amd.diff import sense_forward

        y = sense_forward(x.to(dev), trj.to(dev), md, nf)
    target = x.to(dev)

    def run(ck):
        kd = trj.to(dev).requires_grad_(True)
        rec = unrolled_cg_sense(y, kd, md, nf, num_iters=5, lamda_l2=1e-2, max_eigen=2.0, checkpoint=ck)
        loss = torch.sum(torch.abs(rec - target) ** 2)
        loss.backward()
        return float(loss), kd.grad.clone()

    l1, g1 = run(True)
    l2, g2 = run(False)
    assert abs(l1 - l2) <= 1e-5 * abs(l2) and rel_l2(g1, g2) < 1e-4

    def oracle_unrolled(k64):
        ref = TritonNufftRef(x.shape, oversamp=1.25, width=6, dtype=torch.float64, coords64=True)
        m64 = mps.to(torch.complex128)
        y64 = y.cpu().to(torch.complex128)

        def fwd(v):
            return torch.stack([ref.forward((m64[c] * v)[None], k64[None])[0] for c in range(3)])

        def adj(w):
            return sum(ref.adjoint(w[c][None, None], k64[None])[0, 0] * m64[c].conj() for c in range(3))

        inv = 0.5
        ahb = adj(y64) * 2.0 ** -0.5
        xx = ahb
        r = ahb - (adj(fwd(xx)) * inv + 1e-2 * xx)
        p = r
        rz = torch.real(torch.sum(r.conj() * r))
        for _ in range(5):
            ap = adj(fwd(p)) * inv + 1e-2 * p
            a = rz / torch.real(torch.sum(p.conj() * ap))
            xx = xx + a * p
            r = r - a * ap
            rzn = torch.real(torch.sum(r.conj() * r))
            p = r + (rzn / rz) * p
            rz = rzn
        return float(torch.sum(torch.abs(xx - x.to(torch.complex128)) ** 2))

    assert abs(oracle_unrolled(trj.to(torch.float64)) - l1) / l1 < 1e-3
    h = 1e-4
    v = _smooth_direction(trj, torch.from_numpy(np.random.default_rng(2).standard_normal(trj.shape)), 24, 1.25, 6, h)
    k64 = trj.to(torch.float64)
    fd = (oracle_unrolled(k64 + h * v) - oracle_unrolled(k64 - h * v)) / (2 * h)
    an = float(torch.sum(g1.cpu().to(torch.float64) * v))
    assert abs(an - fd) / abs(fd) < 5e-3, (an, fd)


@pytest.mark.parametrize("dim,use_dcf", [(2, True), (2, False), (3, True)])
def test_normal_function_matches_composition(dev, dim, use_dcf):
    """SenseNormalFn (one autograd function, k-space kept in plan order) against
    sense_adjoint(sense_forward(x)): output, dL/dx and dL/dk for a loss that uses both."""
    from mr_recon_amd.diff import SenseNormalFn, sense_adjoint, sense_forward
    from mr_recon_amd.fourier import hip_nufft

    if dim == 2:
        trj, mps, x, _ = _problem(seed=11)
    else:
        im = (14, 12, 16)
        trj = synth.kooshball_3d(40, 24, 14).reshape(-1, 3) + 0.2 * (torch.rand(960, 3) - 0.5)
        mps = synth.coil_maps(3, im, seed=3)
        x = synth.random_image(im, seed=4)
    dcf = (torch.rand(trj.shape[0], generator=torch.Generator().manual_seed(2)) + 0.5).to(dev) if use_dcf else None
    nf = hip_nufft(tuple(x.shape), oversamp=1.25, width=4)
    target = synth.random_image(tuple(x.shape), seed=9).to(dev)

    def run(fused):
        xd = x.to(dev).requires_grad_(True)
        kd = trj.to(dev).requires_grad_(True)
        if fused:
            out = SenseNormalFn.apply(xd, kd, nf, mps.to(dev), dcf)
        else:
            out = sense_adjoint(sense_forward(xd, kd, mps.to(dev), nf), kd, mps.to(dev), nf, dcf)
        loss = torch.sum(torch.abs(out - target) ** 2)
        loss.backward()
        return out.detach(), xd.grad, kd.grad

    o1, gx1, gk1 = run(True)
    o2, gx2, gk2 = run(False)
    assert rel_l2(o1, o2) < 1e-6
    assert rel_l2(gx1, gx2) < 1e-6
    assert rel_l2(gk1, gk2) < 1e-6


def _cg_step_torch(x, r, p, rz, nrm, inv, lam):
    """The CG update as the composition of torch ops (algs.py:779-872, P = identity): the
    closed form the fused step is differentiated against, run in complex128 on the CPU."""
    ap = nrm * inv + lam * p
    pap = torch.real(torch.sum(p.conj() * ap))
    alpha = rz / pap
    x1 = x + alpha * p
    r1 = r - alpha * ap
    rz1 = torch.real(torch.sum(r1.conj() * r1))
    p1 = r1 + (rz1 / rz) * p
    return x1, r1, p1, rz1


def _cg_step_torch32(x, r, p, rz, nrm, inv, lam):
    """The same composition with complex64 vectors and FP64 inner products and step sizes
    (the torch ops the fused step replaces): the fp32 yardstick of the tolerance."""
    ap = nrm * inv + lam * p
    pap = torch.real(torch.sum(p.conj() * ap, dtype=torch.complex128))
    alpha = (rz / pap).to(torch.float32)
    x1 = x + alpha * p
    r1 = r - alpha * ap
    rz1 = torch.real(torch.sum(r1.conj() * r1, dtype=torch.complex128))
    p1 = r1 + (rz1 / rz).to(torch.float32) * p
    return x1, r1, p1, rz1


@pytest.mark.parametrize("which", ["all", "x_only", "rz_only"])
def test_cg_step_fused_vs_fp64_composition(dev, which):
    """cg_step (fused kernels, fp32 vectors, FP64 scalars) against the complex128 autograd of
    the torch composition: outputs and the gradients of a random linear functional of them
    with respect to x, r, p, rz and nrm, with some cotangents absent. Bound per gradient:
    max(2e-5, 5x the error of the fp32 torch composition itself; the rz gradient goes through
    Re<r1, ap>, a near-cancelling sum of fp32-rounded r1)."""
    from mr_recon_amd.diff import cg_step

    g = torch.Generator().manual_seed(11)

    def rnd(*shape):
        return torch.complex(torch.randn(shape, generator=g), torch.randn(shape, generator=g)).to(torch.complex64)

    n = (48, 40)
    x, r, p, nrm = rnd(*n), rnd(*n), rnd(*n), rnd(*n)
    p = r + 0.3 * p  # a realistic search direction: pAp > 0 below
    nrm = nrm * 0.1 + 2.0 * p  # a positive-definite-like response
    inv, lam = 0.7, 0.05
    rz = torch.real(torch.sum(r.conj().to(torch.complex128) * r.to(torch.complex128)))
    cot = [rnd(*n), rnd(*n), rnd(*n), torch.tensor(0.37, dtype=torch.float64)]
    if which == "x_only":
        cot = [cot[0], None, None, None]
    elif which == "rz_only":
        cot = [None, None, None, cot[3]]

    def functional(outs):
        tot = 0.0
        for o, c in zip(outs, cot):
            if c is None:
                continue
            if o.is_complex():
                tot = tot + torch.real(torch.sum(c.to(o.device, o.dtype).conj() * o))
            else:
                tot = tot + (c.to(o.device) * o)
        return tot

    ins64 = [t.to(torch.complex128).requires_grad_(True) for t in (x, r, p)]
    rz64 = rz.clone().requires_grad_(True)
    nrm64 = nrm.to(torch.complex128).requires_grad_(True)
    out64 = _cg_step_torch(*ins64, rz64, nrm64, inv, lam)
    grads64 = torch.autograd.grad(functional(out64), [*ins64, rz64, nrm64], allow_unused=True)

    insd = [t.to(dev).requires_grad_(True) for t in (x, r, p)]
    rzd = rz.to(dev).requires_grad_(True)
    nrmd = nrm.to(dev).requires_grad_(True)
    outd = cg_step(*insd, rzd, nrmd, inv, lam)
    for a, b in zip(outd, out64):
        assert rel_l2(a.detach().cpu().to(b.dtype), b.detach()) < 1e-6
    gradsd = torch.autograd.grad(functional(outd), [*insd, rzd, nrmd], allow_unused=True)
    ins32 = [t.clone().requires_grad_(True) for t in (x, r, p)]
    rz32 = rz.clone().requires_grad_(True)
    nrm32 = nrm.clone().requires_grad_(True)
    grads32 = torch.autograd.grad(functional(_cg_step_torch32(*ins32, rz32, nrm32, inv, lam)),
                                  [*ins32, rz32, nrm32], allow_unused=True)

    def err(a, b):
        return float(torch.linalg.vector_norm(a.detach().cpu().to(b.dtype) - b) / torch.linalg.vector_norm(b))

    for name, a, b, c in zip(["x", "r", "p", "rz", "nrm"], gradsd, grads64, grads32):
        if b is None or not torch.any(b != 0):
            assert a is None or float(torch.linalg.vector_norm(a.detach().cpu().to(torch.complex128))) == 0.0, name
            continue
        bound = max(2e-5, 5 * err(c, b))
        assert err(a, b) < bound, (name, err(a, b), bound)


@pytest.mark.parametrize("fn", ["forward", "normal"])
def test_kept_and_recomputed_forward_grids_same_bits(dev, monkeypatch, fn):
    """The SENSE functions keep their forward's oversampled grids for the trajectory gradient
    up to KEEP_GRIDS_BYTES and recompute them beyond: the same bits for the forward function;
    the normal function's kept branch sums its two k-terms in one gather over 2C grids (another
    fp32 summation order), its dL/dx is bit-identical."""
    from mr_recon_amd import diff
    from mr_recon_amd.fourier import hip_nufft

    trj, mps, x, y = _problem(seed=13)
    nf = hip_nufft(x.shape, oversamp=1.25, width=6)
    md, yd = mps.to(dev), y.to(dev)

    def run():
        xd = x.to(dev).requires_grad_(True)
        kd = trj.to(dev).requires_grad_(True)
        if fn == "forward":
            loss = torch.sum(torch.abs(diff.sense_forward(xd, kd, md, nf) - yd) ** 2)
        else:
            loss = torch.sum(torch.abs(diff.sense_normal(xd, kd, md, nf)) ** 2)
        loss.backward()
        return xd.grad.clone(), kd.grad.clone()

    gx_keep, gk_keep = run()
    monkeypatch.setattr(diff, "KEEP_GRIDS_BYTES", 0)
    gx_re, gk_re = run()
    assert torch.equal(gx_keep, gx_re)
    if fn == "forward":
        assert torch.equal(gk_keep, gk_re)
    else:
        assert rel_l2(gk_keep, gk_re) < 1e-6


def test_cg_step_zero_direction_stays_finite(dev):
    """A search direction of zeros (pAp = 0, a converged CG) gives alpha = 0 instead of 0/0:
    the iterate and residual stay put, p becomes beta * 0 + r, and the gradients are finite."""
    from mr_recon_amd.diff import cg_step

    g = torch.Generator().manual_seed(3)
    n = (16, 24)
    x = torch.randn(n, generator=g).to(torch.complex64).to(dev).requires_grad_(True)
    r = torch.randn(n, generator=g).to(torch.complex64).to(dev).requires_grad_(True)
    p = torch.zeros(n, dtype=torch.complex64, device=dev, requires_grad=True)
    nrm = torch.zeros(n, dtype=torch.complex64, device=dev, requires_grad=True)
    rz = torch.real(torch.sum(r.detach().conj().to(torch.complex128) * r.detach())).requires_grad_(True)
    x1, r1, p1, rz1 = cg_step(x, r, p, rz, nrm, 0.5, 0.01)
    assert torch.equal(x1, x) and torch.equal(r1, r) and torch.equal(p1, r)
    assert abs(float(rz1) - float(rz)) <= 1e-12 * float(rz)
    loss = torch.sum(torch.abs(x1) ** 2) + torch.sum(torch.abs(p1) ** 2) + rz1
    loss.backward()
    for t in (x, r, p, nrm, rz):
        assert torch.all(torch.isfinite(torch.view_as_real(t.grad) if t.grad.is_complex() else t.grad))
