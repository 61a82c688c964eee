"""Pin the CPU oracle (CPU only, no GPU).

1. Bit-exact against the golden fixtures written by the reference's own code
   (tests/golden/gen_golden.py).
2. Against what the mathematics fixes, independent of the reference: the
   Hicks tables' interpolation properties, I0 against scipy, the NR real FFT
   against numpy, the filter's pass band, the propagator's linearity and
   lateral symmetry, MINRES against scipy's MINRES on the same operator, and
   the Hessian-type operator's relation between its two passes.
3. Against the reference build itself on fresh random inputs when
   oracle/_ref/libdawref.so exists (this container).
"""
import os

import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_small.npz"), allow_pickle=False)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32 if np.asarray(a).dtype == np.float32 else np.uint64)


# ---------------------------------------------------------------- golden fixtures

def test_hicks_tables_golden():
    assert np.array_equal(bits(O.hicks1d()), bits(GOLD["hicks1d"]))
    for ofs in (1, 0):
        for (a, b, p), t in zip(GOLD[f"hicks2d_idx_ofs{ofs}"], GOLD[f"hicks2d_ofs{ofs}"]):
            assert np.array_equal(bits(O.hicks2d_tile(a, b, p, ofs)), bits(t)), (ofs, a, b, p)
    for x, y in zip(GOLD["bessi0_x"], GOLD["bessi0_y"]):
        assert O.bessi0(x) == y


def test_acquisition_golden():
    z, x, h, npml = GOLD["acqui_z"], GOLD["acqui_x"], GOLD["acqui_h"], int(GOLD["acqui_npml"])
    for ofs in (1, 0):
        assert np.array_equal(O.acqui_table(z, x, h, npml, ofs), GOLD[f"acqui_tab_ofs{ofs}"])


def _g():
    g = GOLD
    return g["vp"], g["vp0"], int(g["npml"]), g["h"], int(g["nt"]), g["dt"], g["s"], g["src"], g["rec"], g["dat"]


def test_propagators_golden():
    vp, vp0, npml, h, nt, dt, s, src, rec, dat = _g()
    sis, wf = O.forward(vp, npml, h, 1, nt, dt, s, src, rec, want_wavefield=True)
    assert np.array_equal(sis, GOLD["fwd_sis"]) and np.array_equal(wf, GOLD["fwd_wf"])
    dcal = O.forward(vp0, npml, h, 1, nt, dt, s, src, rec)
    res = O.residuals(sis, dcal)
    assert np.array_equal(res, GOLD["res"])
    assert np.array_equal(O.esource_data(vp0, npml, h, 1, nt, dt, res, rec), GOLD["esd_sis"])
    esc, wfe = O.esource(vp0, npml, h, 1, nt, dt, res, rec, want_wavefield=True)
    assert np.array_equal(esc, GOLD["es_sis"]) and np.array_equal(wfe[::10], GOLD["es_wf_slice"])
    assert np.array_equal(O.resample(wf, npml, dat, 1), GOLD["resample_sis"])


def test_minres_golden():
    vp, vp0, npml, h, nt, dt, s, src, rec, dat = _g()
    x, hist, info = O.linearminres(vp0, npml, h, 1, nt, dt, GOLD["res"], rec, 1e-2, 5)
    assert np.array_equal(x, GOLD["lmr_deltade"])
    sc = GOLD["minres_scalars"]
    assert np.float32(info["fac"]) == sc[0, 1]
    # hist[k] = qrnorm/beta1 after k+1 iterations, a REAL division: the
    # reference's rnorm after k+1 iterations divided by its beta1 in FP32
    # gives the same bits
    assert hist.dtype == np.float32
    assert np.array_equal(sc[:, 0] / np.float32(info["beta1"]), hist)
    # the final scalars of the 5-iteration solve, bit for bit
    for j, k in ((0, "rnorm"), (2, "Anorm"), (3, "Acond"), (4, "ynorm")):
        assert np.float32(info[k]) == sc[-1, j], k


def test_filter_golden():
    d = GOLD["filt_in"]
    assert np.array_equal(O.filmartin2d(d, 1e-3, 3.0, 30.0, 2.0, 10.0), GOLD["filt_out"])
    assert np.array_equal(O.filmartin2d(d, 8e-4, 2.0, 700.0, 1.0, 5.0), GOLD["filt_out2"])


# ---------------------------------------------------------------- mathematics

def test_hicks_interpolation_properties():
    c = O.hicks1d().astype(np.float64)            # (100 alphas, 21 taps)
    # alpha = 0 (ia = 51): the window collapses onto the node
    # (sin(pi*n) with the REAL pi leaves O(1e-8) side taps, kept by the 1e-15 cut)
    assert c[50, 10] == 1.0 and np.abs(np.delete(c[50], 10)).max() < 1e-7
    # windowed sinc reproduces constants to the Kaiser window's accuracy
    assert np.all(np.abs(c.sum(1) - 1.0) < 2e-3)
    # mirror symmetry c(n, alpha) = c(-n, -alpha): ia <-> 102 - ia
    for ia in range(2, 101):
        assert np.allclose(c[ia - 1], c[101 - ia][::-1], atol=1e-6)
    # first moment: sum c(n) * n = -alpha (the sinc centre sits at x = n + alpha = 0)
    alpha = 0.5 - np.arange(100) / 100.0
    assert np.allclose((c * np.arange(-10, 11)).sum(1), -alpha, atol=2e-2)


def test_free_surface_mirror():
    # ip nodes above the surface: surface node zero, odd mirror below it
    base = O.hicks2d_tile(37, 64, 0, 1).astype(np.float64)
    for ip in (1, 4, 10):
        t = O.hicks2d_tile(37, 64, ip, 1).astype(np.float64)
        surf = ip
        assert np.all(t[:, surf] == 0.0)
        for k in range(1, ip + 1):
            assert np.allclose(t[:, surf + k], base[:, surf + k] - base[:, surf - k], atol=1e-7)
        assert np.array_equal(t[:, :surf], base[:, :surf].astype(np.float32))
    assert not O.hicks2d_tile(37, 64, 3, 0).any()      # no free surface: slice stays zero


def test_bessi0_vs_scipy():
    from scipy.special import i0
    x = np.linspace(0.0, 14.18, 200).astype(np.float32)
    y = np.array([O.bessi0(v) for v in x], np.float64)
    assert np.all(np.abs(y / i0(x.astype(np.float64)) - 1.0) < 2e-7 * 10)


def test_realft_vs_numpy():
    rng = np.random.default_rng(0)
    n = 256
    x = rng.standard_normal(n)
    re, im = O.realft2s(x, np.zeros(n), 2)
    X = np.fft.fft(x)   # NR sign convention: exp(+i): conj of numpy's
    assert np.allclose(re[0], X[0].real, rtol=1e-6) and np.allclose(im[0], X[n // 2].real, rtol=1e-6)
    assert np.allclose(re[1:n // 2], X[1:n // 2].real, rtol=1e-5, atol=1e-5)
    assert np.allclose(im[1:n // 2], -X[1:n // 2].imag, rtol=1e-5, atol=1e-5)
    back, _ = O.realft2s(re, im, 1)
    assert np.allclose(back, x, atol=1e-5)


def test_filter_passband():
    nt, dt = 1000, 1e-3
    t = np.arange(nt) * dt
    x = np.sin(2 * np.pi * 15.0 * t).astype(np.float32)[None, :]
    y = O.filmartin2d(x, dt, 3.0, 30.0, 2.0, 10.0)
    mid = slice(200, 800)
    assert rel_l2(y[0, mid], x[0, mid]) < 0.02           # pass band
    z = O.filmartin2d(np.sin(2 * np.pi * 200.0 * t).astype(np.float32)[None, :], dt, 3.0, 30.0, 2.0, 10.0)
    assert np.abs(z[0, mid]).max() < 1e-3                 # stop band (0.0001 floor)


def test_propagator_linear_and_symmetric():
    c = Case(nz=30, nx=61, nt=200)
    vp = np.full((c.n2, c.n1), 2000.0, np.float32)
    xm = (c.n2 - 1) * float(c.h) / 2
    src = O.acqui_table([100.0], [xm], c.h, c.npml, 1)[:, 0].copy()
    rx = np.array([xm - 150.0, xm + 150.0])
    rec = O.acqui_table([100.0, 100.0], rx, c.h, c.npml, 1)
    a = O.forward(vp, c.npml, c.h, 1, c.nt, c.dt, c.s, src, rec, prec="f64")
    b = O.forward(vp, c.npml, c.h, 1, c.nt, c.dt, 2.0 * c.s, src, rec, prec="f64")
    assert np.allclose(b, 2.0 * a, rtol=1e-12, atol=1e-12 * np.abs(a).max())
    assert np.allclose(a[0], a[1], rtol=1e-9, atol=1e-9 * np.abs(a).max())   # lateral symmetry
    assert np.abs(a).max() > 0


def test_minres_vs_scipy():
    """The oracle's MINRES recurrences (minresModule.f90:165-303) against
    scipy's MINRES on the same operator (time-reversed esource_data), FP64."""
    from scipy.sparse.linalg import LinearOperator, minres
    c = Case(nz=24, nx=40, nt=80, nrec=5)
    dobs = O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec, prec="f64")
    dcal = O.forward(c.vp0, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec, prec="f64")
    b = np.ascontiguousarray((dobs - dcal)[:, ::-1])
    nr, nt = b.shape
    k = 6
    x, hist, info = O.linearminres(c.vp0, c.npml, c.h, 1, c.nt, c.dt, b, c.rec, 1e-2, k, prec="f64")

    def mv(v):
        y = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, np.ascontiguousarray(v.reshape(nr, nt)), c.rec,
                           prec="f64")
        return np.ascontiguousarray(y[:, ::-1]).ravel()

    A = LinearOperator((nr * nt, nr * nt), matvec=mv, dtype=np.float64)
    xs, _ = minres(A, b.ravel(), shift=-info["fac"], maxiter=k, rtol=0.0)
    assert rel_l2(x.ravel(), xs) < 1e-8


def test_hessian_passes_consistent():
    """esource and esource_data differ only by the pressure extended-source
    range (submodeling2.f90:777-778 vs :1121-1122): with a zero snapshot on
    the interior rim they must agree, and each is linear in res."""
    c = Case(nz=26, nx=50, nt=120)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt))
    a = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, prec="f64")
    b = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, 3.0 * res, c.rec, prec="f64")
    assert np.allclose(b, 3.0 * a, rtol=1e-10, atol=1e-10 * np.abs(a).max())
    e = O.esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, prec="f64")
    assert 0 < rel_l2(e, a) < 0.5


# ---------------------------------------------------------------- live reference

ref = pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")


@ref
def test_oracle_equals_reference_random():
    c = Case(nz=28, nx=52, nt=140, ofs=0, seed=9)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    O.ref_init(0)
    sis, wf = O.ref_forward(c.vp, c.npml, c.h, 0, c.nt, c.dt, c.s, c.src, c.rec)
    a, wa = O.forward(c.vp, c.npml, c.h, 0, c.nt, c.dt, c.s, c.src, c.rec, want_wavefield=True)
    assert np.array_equal(a, sis) and np.array_equal(wa, wf)
    assert np.array_equal(O.esource_data(c.vp0, c.npml, c.h, 0, c.nt, c.dt, res, c.rec),
                          O.ref_esource_data(c.vp0, c.npml, c.h, 0, c.nt, c.dt, res, c.rec))
    O.ref_init(1)
    rng = np.random.default_rng(4)
    for _ in range(2000):
        z, x = np.float32(rng.uniform(0, 250)), np.float32(rng.uniform(0, 4000))
        h = np.float32(rng.choice([10.0, 8.0, 6.25, 12.5, 3.125]))
        a = O.subsourcekws(z, x, h, 1)
        b = O.ref_subsourcekws(z, x, h, 1)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_exact_adjoint_dot_product():
    """FP64 dot test of the exact transpose of the forward scheme, and symmetry
    of the exact Hessian (time reversal o esource_data_exact)."""
    for ofs in (1, 0):
        c = Case(nz=26, nx=50, nt=120, ofs=ofs, nrec=9)
        rng = np.random.default_rng(0)
        s = rng.standard_normal(c.nt).astype(np.float32)
        d = rng.standard_normal((c.rec.shape[1], c.nt))
        Fs = O.forward(c.vp, c.npml, c.h, ofs, c.nt, c.dt, s, c.src, c.rec, prec="f64")
        g = O.forward_adjoint(c.vp, c.npml, c.h, ofs, c.nt, c.dt, d, c.src, c.rec, prec="f64")
        lhs, rhs = (Fs * d).sum(), (s.astype(np.float64) * g).sum()
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
        x = rng.standard_normal(d.shape)
        A = lambda v: O.esource_data_exact(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, v, c.rec)[:, ::-1]
        Ax, Ad = A(x), A(d)
        assert abs((x * Ad).sum() - (Ax * d).sum()) <= 1e-12 * abs((x * Ad).sum())
        assert (x * Ax).sum() > 0
