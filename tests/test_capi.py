"""The C-ABI library on the CPU (no GPU): it loads, exports every symbol
include/daw.h declares, its host-side geometry is bit-identical to the
reference's (via the oracle), and argument errors come back as codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "daw.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(daw_\w+)\s*\(", txt, re.M)))


def test_exports_every_declared_symbol(daw):
    L = daw.lib()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(daw.EXPORTS)
    assert L.daw_version() == 2


def test_host_geometry_bitexact(daw):
    assert np.array_equal(daw.hicks1d().view(np.uint32), O.hicks1d().view(np.uint32))
    rng = np.random.default_rng(2)
    for _ in range(300):
        a, b, p, ofs = int(rng.integers(1, 101)), int(rng.integers(1, 101)), int(rng.integers(0, 11)), \
            int(rng.integers(-1, 2))
        assert np.array_equal(daw.hicks2d_tile(a, b, p, ofs).view(np.uint32), O.hicks2d_tile(a, b, p, ofs).view(np.uint32))
    z = rng.uniform(0, 300, 3000).astype(np.float32)
    x = rng.uniform(0, 6000, 3000).astype(np.float32)
    for h in (10.0, 8.0, 6.25, 3.125):
        for ofs in (0, 1):
            assert np.array_equal(daw.subacqui(z, x, np.float32(h), 20, ofs), O.acqui_table(z, x, np.float32(h), 20, ofs))
    gold = np.load(os.path.join(ROOT, "tests", "golden", "ref_small.npz"), allow_pickle=False)
    assert np.array_equal(daw.subacqui(gold["acqui_z"], gold["acqui_x"], gold["acqui_h"], int(gold["acqui_npml"]), 1),
                          gold["acqui_tab_ofs1"])


def test_filter_length_checked_before_gpu_work(daw):
    # the band-pass runs on the GPU; a trace longer than the FFT length list
    # (65536) is refused before any device work
    with pytest.raises(daw.DawError, match="65536"):
        daw.filmartin2d(np.zeros((1, 70000), np.float32), 1e-3, 1, 2, 1, 1)


def test_error_codes(daw):
    L = daw.lib()
    # propagators refuse to run before kwsinc2d, as the reference needs its table first
    L.daw_deallocatekws2d()
    daw._hicks_ofs = None
    vp = np.full((10, 10), 1500.0, np.float32)
    s = np.zeros(5, np.float32)
    is5 = np.array([12, 12, 50, 50, 0], np.int32)
    ir = np.array([[12], [12], [50], [50], [0]], np.int32)
    sis = np.zeros((1, 5), np.float32)
    rc = L.daw_submodeling(vp, 10, 10, 10, C.c_float(10.0), 1, 5, C.c_float(1e-3), s, 0, 0, is5, ir, 1, sis, None)
    assert rc == -2 and b"kwsinc2d" in L.daw_last_error()
    L.daw_kwsinc2d(1)
    rc = L.daw_submodeling(vp, 10, 10, 2, C.c_float(10.0), 1, 5, C.c_float(1e-3), s, 0, 0, is5, ir, 1, sis, None)
    assert rc == -1          # npml < 3
    rc = L.daw_hicks2d_tile(0, 1, 0, 1, np.zeros(441, np.float32))
    assert rc == -1
    L.daw_deallocatekws2d()


def _window_from_weights(w1, w2, ip, ofs):
    """A 21x21 Hicks window [n2][n1] restated from two 1-D weight vectors
    (kwsinc.f90:268-269, 336-355): w2*w1; with a free surface the node row ip
    is zero and the rows ip+k (k = 1..ip) are w2*w1(ip+k) - w2*w1(ip-k); an
    all-zero window for slices ip >= 1 without one."""
    c = np.zeros((21, 21), np.float32)
    if ip > 0 and abs(ofs) != 1:
        return c
    for n1 in range(21):
        if ip > 0 and n1 == ip:
            continue
        if ip > 0 and ip < n1 <= 2 * ip:
            c[:, n1] = (w2 * w1[n1]) - (w2 * w1[2 * ip - n1])
        else:
            c[:, n1] = w2 * w1[n1]
    return c


def test_window_is_tensor_product_of_1d_table(daw):
    """Every 2-D window the library builds is the tensor product of two rows of
    its 1-D table, with the free-surface slices' zero row and odd mirror, bit
    for bit (an independent statement of kwsinc2d's construction)."""
    t1 = daw.hicks1d()
    rng = np.random.default_rng(11)
    for ofs in (1, 0):
        daw.kwsinc2d(ofs)
        try:
            for _ in range(60):
                a1, a2 = (int(v) for v in rng.integers(1, 101, 2))
                for ip in (0, int(rng.integers(1, 11)), 10):
                    want = daw.hicks2d_tile(a1, a2, ip, ofs)
                    got = _window_from_weights(t1[a1 - 1], t1[a2 - 1], ip, ofs)
                    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (ofs, a1, a2, ip)
        finally:
            daw.deallocatekws2d()
