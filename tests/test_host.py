"""CPU tests of the host side: the C-ABI library loads and exports every declared
symbol, and the Python boundary validates inputs before any launch."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest
import torch

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "mr_recon_amd.h"


def _declared():
    txt = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mra_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    from mr_recon_amd import _lib

    assert set(_declared()) == set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    from mr_recon_amd import _lib

    lib = _lib.lib()
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (mra_[a-z0-9_]+)", out))
    assert set(_declared()) <= exported


def test_query_functions_without_gpu():
    from mr_recon_amd import _lib

    L = _lib.lib()
    assert L.mra_version() >= 100
    assert L.mra_default_tile(3) == 8 and L.mra_default_tile(2) == 16
    assert L.mra_subgrid_points(3, 5) == 12 ** 3
    assert L.mra_max_coil_chunk(3, 5) * 12 ** 3 * 8 <= 64 * 1024
    assert L.mra_geom_size() % 8 == 0


def test_tile_selection_is_per_thread_and_validated():
    import threading

    from mr_recon_amd import _lib

    L = _lib.lib()
    assert L.mra_tile(2) == 16
    assert L.mra_set_tile(2, 12) == -1 and b"tile" in L.mra_last_error()  # no kernels for 12
    assert L.mra_set_tile(3, 16) == -1 and L.mra_set_tile(1, 128) == -1
    seen = {}

    def other():
        seen["before"] = L.mra_tile(2)
        with _lib.tile_scope(2, 8):
            seen["inside"] = L.mra_tile(2)
            seen["subgrid"] = L.mra_subgrid_points(2, 5)

    with _lib.tile_scope(2, 32):
        assert L.mra_tile(2) == 32 and L.mra_subgrid_points(2, 5) == 36 ** 2
        th = threading.Thread(target=other)
        th.start()
        th.join()
        assert L.mra_tile(2) == 32  # the other thread's choice did not leak
    assert L.mra_tile(2) == 16
    assert seen == {"before": 16, "inside": 8, "subgrid": 12 ** 2}
    with pytest.raises(ValueError):
        with _lib.tile_scope(2, 24):
            pass


def test_error_path_reports_message():
    from mr_recon_amd import _lib

    L = _lib.lib()
    buf = (ctypes.c_ubyte * L.mra_geom_size())()
    rc = L.mra_geom_fill(ctypes.addressof(buf), 0, 1, 0, 1, 0, 1, 1, 4, _lib.ints([4, 4, 4, 4]), 8, None,
                         1.0, 4.0, 1.0, 0, 0, 0)
    assert rc != 0
    assert b"ndim" in L.mra_last_error()
    with pytest.raises(RuntimeError, match="ndim"):
        _lib.check(rc)


def test_cpu_tensors_are_rejected_loudly():
    from mr_recon_amd.fourier import hip_nufft
    from mr_recon_amd.hip_interp import grid, ungrid

    x = torch.zeros((1, 8, 8), dtype=torch.complex64)
    k = torch.zeros((1, 10, 2))
    with pytest.raises(RuntimeError, match="GPU only"):
        hip_nufft((8, 8)).forward(x, k)
    with pytest.raises(RuntimeError, match="GPU only"):
        ungrid(torch.zeros((8, 8), dtype=torch.complex64), torch.zeros((5, 2)), 4.0)
    with pytest.raises(RuntimeError, match="GPU only"):
        grid(torch.zeros(5, dtype=torch.complex64), torch.zeros((5, 2)), (8, 8), 4.0)
    from mr_recon_amd.algs import density_compensation
    from mr_recon_amd.fourier import sigpy_nufft

    with pytest.raises(RuntimeError, match="GPU only"):
        density_compensation(torch.zeros((10, 2)), (8, 8))
    with pytest.raises(RuntimeError, match="GPU only"):
        sigpy_nufft((8, 8)).forward(x, k)


def test_option_validation_matches_reference_errors():
    from mr_recon_amd.hip_interp import grid, ungrid

    v = torch.zeros((8, 8), dtype=torch.complex64)
    with pytest.raises(ValueError, match="same length"):
        ungrid(v, torch.zeros((5, 2)), (4.0, 4.0, 4.0))
    with pytest.raises(ValueError, match="Expected last dim of locs"):
        grid(torch.zeros(5, dtype=torch.complex64), torch.zeros((5, 3)), (8, 8), 4.0)
    with pytest.raises(ValueError, match="Unrecognized kernel"):
        ungrid(v, torch.zeros((5, 2)), 4.0, kernel="gauss")
    with pytest.raises(NotImplementedError):
        ungrid(v, torch.zeros((5, 2)), 4.0, kernel="spline")
    with pytest.raises(ValueError, match="exceeds maximum width"):
        ungrid(v, torch.zeros((5, 2)), 40.0)


def test_padlast_matches_reference_layout():
    from mr_recon_amd.pad import PadLast

    p = PadLast((8, 10), (4, 6))
    x = torch.arange(24.0).reshape(4, 6)
    y = p.forward(x)
    assert y.shape == (8, 10) and torch.equal(y[2:6, 2:8], x) and y.sum() == x.sum()
    assert torch.equal(p.adjoint(y), x)
    assert p.center == (2, 3)
    with pytest.raises(AssertionError):
        PadLast((7, 10), (4, 6))


def test_resize_and_gen_grd():
    from mr_recon_amd.utils import batch_iterator, gen_grd, resize

    assert resize(torch.arange(5.0), (8,)).tolist() == [0, 0, 0, 1, 2, 3, 4, 0]
    assert resize(torch.arange(8.0), (5,)).tolist() == [2, 3, 4, 5, 6]
    g = gen_grd((4, 3))
    assert g.shape == (4, 3, 2) and torch.allclose(g[0, 0], torch.tensor([-0.5, -1 / 3]))
    assert list(batch_iterator(5, 2)) == [(0, 2), (2, 4), (4, 5)]


def test_beta_and_shapes_of_hip_nufft():
    import numpy as np

    from mr_recon_amd.fourier import hip_nufft, kb_beta, triton_nufft

    n = hip_nufft((128, 128, 128), oversamp=1.25, width=4)
    assert n.os_shape() == (160, 160, 160)
    assert abs(kb_beta(4, 1.25) - np.pi * ((3.2 * 0.75) ** 2 - 0.8) ** 0.5) < 1e-12
    assert triton_nufft is hip_nufft
    sc, sh = n._scale_shift((150, 100))
    assert sc[0] == float(np.float32(188 / 150)) and sh == [94.0, 62.0]


def test_sigpy_nufft_beta_default(capsys):
    """sigpy_nufft's beta: the KB formula, or 1.0 (with the reference's warning unless
    width == 1) where its radicand is negative (fourier.py:740-747)."""
    from mr_recon_amd.fourier import kb_beta, sigpy_nufft

    assert sigpy_nufft((16, 16), oversamp=2.0, width=6).beta == kb_beta(6, 2.0)
    assert sigpy_nufft((16, 16), oversamp=1.25, width=1).beta == 1.0
    assert "WARNING" not in capsys.readouterr().out
    assert sigpy_nufft((16, 16), oversamp=1.0, width=1.5).beta == 1.0
    assert "WARNING: beta is set to 1.0" in capsys.readouterr().out
    assert sigpy_nufft((16, 16), beta=3.0).beta == 3.0


def test_synthetic_generators():
    from mr_recon_amd import synth

    k = synth.kooshball_3d(100, 64, 32)
    assert k.shape == (100, 64, 3) and float(k.abs().max()) <= 16
    m = synth.coil_maps(8, (16, 16), seed=0)
    rss = (m.abs() ** 2).sum(0)
    assert torch.allclose(rss, torch.ones_like(rss), atol=1e-5)
    s = synth.spiral_2d(4, 100, 64)
    assert float(s.abs().max()) < 32


def test_lazy_conj_views_are_refused():
    """x.conj() shares x's memory with a conjugate bit; its raw address would hand a
    kernel the unconjugated values, so the binding refuses it (inputs of the ops
    materialise it first)."""
    from mr_recon_amd import _lib
    from mr_recon_amd.ops import _c64

    x = torch.randn(4, dtype=torch.complex64)
    with pytest.raises(ValueError, match="resolve_conj"):
        _lib.ptr(x.conj())
    y = _c64(x.conj(), "x")
    assert not y.is_conj() and torch.equal(y, torch.conj_physical(x))
    assert _lib.ptr(x) is not None


def test_gridded_scatter_chunks_vs_groupby():
    """gridded_nufft._chunks (torch ops, run here on the CPU) against a brute-force group-by of
    the sorted indices: chunks tile every run in order, none longer than the chunk size, each
    names its run's index, and the permutation is the stable sort."""
    import itertools

    from mr_recon_amd.fourier import gridded_nufft

    g = torch.Generator().manual_seed(0)
    idx = torch.cat([torch.randint(0, 50, (500,), generator=g), torch.full((300,), 7), torch.full((129,), 3),
                     torch.full((64,), 11), torch.full((65,), 12)]).to(torch.int32)
    idx = idx[torch.randperm(idx.numel(), generator=g)]
    for chunk in (1, 4, 64):
        perm, cs, ce, ci = gridded_nufft._chunks(idx, chunk)
        order = sorted(range(idx.numel()), key=lambda i: (int(idx[i]), i))
        assert perm.tolist() == order
        want = []
        pos = 0
        for key, grp in itertools.groupby([int(idx[i]) for i in order]):
            n = len(list(grp))
            for s in range(0, n, chunk):
                want.append((pos + s, pos + min(s + chunk, n), key))
            pos += n
        assert list(zip(cs.tolist(), ce.tolist(), ci.tolist())) == want


def test_cg_step_entry_points_validate_before_any_launch():
    """The differentiable CG step's entry points refuse empty vectors and null buffers with a
    message, before touching a device (so this runs without a GPU)."""
    from mr_recon_amd import _lib

    L = _lib.lib()
    rc = L.mra_cg_step_fwd(None, None, None, None, None, 0, 1.0, 0.0, None, None, None, None, None, None, None)
    assert rc != 0 and b"n must be positive" in L.mra_last_error()
    rc = L.mra_cg_step_fwd(None, None, None, None, None, 16, 1.0, 0.0, None, None, None, None, None, None, None)
    assert rc != 0 and b"null pointer" in L.mra_last_error()
    rc = L.mra_cg_step_bwd(None, None, None, None, None, None, None, None, 16, 1.0, 0.0, None, None, None, None,
                           None, None, None)
    assert rc != 0 and b"null pointer" in L.mra_last_error()


def test_binning_chunk_bounds():
    """Radix chunks stay within 128..4096 samples, and the histogram (digits x chunks) within
    2^24 counts wherever 4096-sample chunks allow it."""
    from mr_recon_amd.plan import MAX_DIGIT, binning_chunk

    for npts in (1, 100, 3000, 131072, 5_120_000, 50_000_000):
        for ntiles, tile, d in ((27, 8, 3), (400, 16, 2), (8000, 8, 3), (64000, 8, 3)):
            c = binning_chunk(npts, ntiles, tile, d)
            assert 128 <= c <= 4096
            digits = max(tile ** d, min(ntiles + 1, MAX_DIGIT))
            chunks = -(-npts // c)
            assert c == 4096 or digits * chunks <= 1 << 24
