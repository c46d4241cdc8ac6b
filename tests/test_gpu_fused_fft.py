"""GPU parity of the FFTs fused with zero-pad / crop (ops.pad_fft, ops.ifft_crop,
ops.toeplitz_normal) against a plain torch fp32 restatement: place pixel n at grid index
(n - c) mod G, torch.fft, and the mirror crop / coil reduction. Covers odd image sizes,
centres at both ends, 1D / 2D / 3D, the 128^3 -> 160^3 and 256^3 production grids, SENSE
maps, accumulation, batched image / coil indexing and the unfused fallback shapes."""

import math

import pytest
import torch

from helpers import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L2, complex64 against torch's complex64 FFTs


def _rand(shape, seed, dev):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(shape, dtype=torch.complex64, generator=g).to(dev)


def _apod(N, seed, dev):
    g = torch.Generator().manual_seed(seed)
    return (torch.rand(sum(N), generator=g) + 0.5).to(dev)


def _apod_grid(apod, N):
    if apod is None:
        return 1.0
    out = torch.ones(N, device=apod.device)
    off = 0
    for d, n in enumerate(N):
        shape = [1] * len(N)
        shape[d] = n
        out = out * apod[off:off + n].reshape(shape)
        off += n
    return out


def _centre(N, center):
    return tuple(n // 2 for n in N) if center is None else tuple(center)


def ref_pad_fft(x, N, G, smaps=None, apod=None, scale=1.0, center=None, img_index=None):
    D = len(N)
    if img_index is not None:
        v = x[img_index.long()]
        if smaps is not None:
            v = v * smaps[: v.shape[0]]
    elif smaps is not None:
        v = (x[:, None] * smaps[None]).reshape(-1, *N)
    else:
        v = x.reshape(-1, *N)
    v = v * _apod_grid(apod, N) * scale
    g = torch.zeros((v.shape[0], *G), dtype=torch.complex64, device=x.device)
    g[(slice(None),) + tuple(slice(0, n) for n in N)] = v
    c = _centre(N, center)
    g = torch.roll(g, shifts=tuple(-ci for ci in c), dims=tuple(range(-D, 0)))
    return torch.fft.fftn(g, dim=tuple(range(-D, 0)))


def ref_ifft_crop(grid, N, G, nx, smaps=None, apod=None, scale=1.0, center=None, conj_maps=True, coil_off=None):
    D = len(N)
    g = grid.reshape(-1, *G)
    img = torch.fft.ifftn(g, dim=tuple(range(-D, 0)), norm="forward")
    c = _centre(N, center)
    img = torch.roll(img, shifts=c, dims=tuple(range(-D, 0)))
    img = img[(slice(None),) + tuple(slice(0, n) for n in N)]
    if coil_off is None:
        q_of = torch.arange(img.shape[0], device=grid.device) // (img.shape[0] // nx)
        map_of = torch.arange(img.shape[0], device=grid.device) % (img.shape[0] // nx)
    else:
        off = coil_off.cpu().tolist()
        q_of = torch.tensor([b for b in range(nx) for _ in range(off[b], off[b + 1])], device=grid.device)
        map_of = torch.arange(img.shape[0], device=grid.device)
    if smaps is not None:
        s = smaps[map_of]
        img = img * (s.conj() if conj_maps else s)
    out = torch.zeros((nx, *N), dtype=torch.complex64, device=grid.device)
    out.index_add_(0, q_of, img)
    return out * _apod_grid(apod, N) * scale


# (N, G, nx, ns, center)
CASES = [
    ((37, 50, 61), (64, 80, 96), 2, 3, None),
    ((128, 128, 128), (160, 160, 160), 1, 4, None),
    ((100, 120), (128, 160), 3, 2, (0, 120)),
    ((200, 256), (256, 320), 1, 0, None),
    ((64, 80), (64, 80), 2, 1, None),
    ((500,), (640,), 4, 0, (250,)),
    ((20, 30, 600), (64, 64, 640), 1, 2, None),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_pad_fft_matches_torch(dev, case):
    from mr_recon_amd import ops

    N, G, nx, ns, center = CASES[case]
    assert ops.fused_supported(G)
    x = _rand((nx, *N), case, dev)
    smaps = _rand((ns, *N), 50 + case, dev) if ns else None
    apod = _apod(N, case, dev)
    scale = 1.0 / math.sqrt(math.prod(N))
    got = ops.pad_fft(x, N, G, smaps=smaps, apod=apod, scale=scale, center=center)
    want = ref_pad_fft(x, N, G, smaps=smaps, apod=apod, scale=scale, center=center)
    assert tuple(got.shape) == ((nx, ns, *G) if ns else (nx, *G))
    assert rel_l2(got.reshape(want.shape), want) < TOL


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("accumulate,conj_maps", [(False, True), (True, False)])
def test_ifft_crop_matches_torch(dev, case, accumulate, conj_maps):
    from mr_recon_amd import ops

    N, G, nx, ns, center = CASES[case]
    nq = nx * max(ns, 1)
    grid = _rand((nq, *G), 7 + case, dev)
    smaps = _rand((ns, *N), 60 + case, dev) if ns else None
    apod = _apod(N, 3 + case, dev)
    scale = 0.37
    prev = _rand((nx, *N), 90 + case, dev)
    out = prev.clone() if accumulate else None
    want = ref_ifft_crop(grid, N, G, nx, smaps=smaps, apod=apod, scale=scale, center=center, conj_maps=conj_maps)
    if accumulate:
        want = want + prev
    got = ops.ifft_crop(grid.clone(), N, G, nx, smaps=smaps, apod=apod, scale=scale, out=out,
                        accumulate=accumulate, conj_maps=conj_maps, center=center)
    assert rel_l2(got, want) < TOL


def test_batched_index_tables(dev):
    """Grid q pads image img_index[q] times map q; image b reduces grids coil_off[b]..coil_off[b+1]-1."""
    from mr_recon_amd import ops

    N, G = (96, 100), (120, 128)
    counts = [2, 1, 3, 2]
    nx, nq = len(counts), sum(counts)
    img_index = torch.repeat_interleave(torch.arange(nx, dtype=torch.int32), torch.tensor(counts)).to(dev)
    coil_off = torch.tensor([0] + list(torch.cumsum(torch.tensor(counts), 0).tolist()), dtype=torch.int32,
                            device=dev)
    x = _rand((nx, *N), 1, dev)
    smaps = _rand((nq, *N), 2, dev)
    apod = _apod(N, 4, dev)
    got = ops.pad_fft(x, N, G, smaps=smaps, apod=apod, scale=0.5, img_index=img_index, nq=nq)
    want = ref_pad_fft(x, N, G, smaps=smaps, apod=apod, scale=0.5, img_index=img_index)
    assert rel_l2(got, want) < TOL
    grid = _rand((nq, *G), 3, dev)
    want = ref_ifft_crop(grid, N, G, nx, smaps=smaps, apod=apod, scale=0.25, coil_off=coil_off)
    got = ops.ifft_crop(grid.clone(), N, G, nx, smaps=smaps, apod=apod, scale=0.25, coil_off=coil_off)
    assert rel_l2(got, want) < TOL


@pytest.mark.parametrize("N,G,ns", [((100, 120), (200, 240), 3), ((40, 48, 64), (80, 96, 128), 2),
                                    ((128, 128, 128), (256, 256, 256), 2)])
@pytest.mark.parametrize("accumulate", [False, True])
def test_toeplitz_fused_matches_torch(dev, N, G, ns, accumulate):
    from mr_recon_amd import ops
    from mr_recon_amd.pad import PadLast

    seed = sum(N) + ns
    x = _rand(N, seed, dev)
    smaps = _rand((ns, *N), seed + 1, dev)
    kern = _rand(G, seed + 2, dev) / math.prod(G)
    cen = PadLast(G, N).center
    prev = _rand((1, *N), seed + 3, dev)
    out = prev.clone() if accumulate else None
    spec = ref_pad_fft(x[None], N, G, smaps=smaps, center=cen) * kern
    want = ref_ifft_crop(spec, N, G, 1, smaps=smaps, center=cen)
    if accumulate:
        want = want + prev
    got = ops.toeplitz_normal(x, N, G, smaps, kern, center=cen, out=out, accumulate=accumulate)
    assert rel_l2(got, want) < TOL


@pytest.mark.parametrize("N,G", [((80, 80), (100, 100)), ((30, 40, 50), (40, 48, 60))])
def test_fallback_shapes(dev, N, G):
    """Grids without a compiled radix plan go through pad/crop kernels and rocFFT."""
    from mr_recon_amd import ops

    assert not ops.fused_supported(G)
    x = _rand((2, *N), 5, dev)
    smaps = _rand((2, *N), 6, dev)
    apod = _apod(N, 7, dev)
    got = ops.pad_fft(x, N, G, smaps=smaps, apod=apod, scale=0.3)
    assert rel_l2(got.reshape(4, *G), ref_pad_fft(x, N, G, smaps=smaps, apod=apod, scale=0.3)) < TOL
    grid = _rand((4, *G), 8, dev)
    want = ref_ifft_crop(grid, N, G, 2, smaps=smaps, apod=apod, scale=0.3)
    assert rel_l2(ops.ifft_crop(grid.clone(), N, G, 2, smaps=smaps, apod=apod, scale=0.3), want) < TOL
    kern = _rand(G, 9, dev) / math.prod(G)
    spec = ref_pad_fft(x[:1], N, G, smaps=smaps) * kern
    want = ref_ifft_crop(spec, N, G, 1, smaps=smaps)
    assert rel_l2(ops.toeplitz_normal(x[0], N, G, smaps, kern), want) < TOL
