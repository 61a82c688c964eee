"""GPU parity of the HIP path against the CPU oracle (restatement of the
reference, itself pinned bit-for-bit to the reference build).

Single-precision propagators, Hicks gathers and resampling are compared
bit-exactly: the kernels perform the reference's IEEE operations in the
reference's order with FP contraction off.  MINRES uses FP64 dot products and
Lanczos scalars where the reference uses FP32, so its solution is compared
within 1e-4 relative L2 (the bar for MINRES solutions) against both the FP32
oracle and the FP64 oracle.
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("ofs,type_src", [(1, 0), (0, 0), (1, 1)])
def test_submodeling_bitexact(gpu_daw, ofs, type_src):
    c = Case(ofs=ofs)
    _say("submodeling", ofs, type_src, c.n1, c.n2, c.nt)
    sis, wf = gpu_daw.submodeling(c.vp, c.npml, c.h, ofs, c.nt, c.dt, c.s, c.src, c.rec, type_src=type_src,
                                  want_wavefield=True)
    osis, owf = O.forward(c.vp, c.npml, c.h, ofs, c.nt, c.dt, c.s, c.src, c.rec, type_src=type_src,
                          want_wavefield=True)
    assert np.abs(osis).max() > 0
    assert np.array_equal(sis, osis), rel_l2(sis, osis)
    assert np.array_equal(wf, owf), rel_l2(wf, owf)


def test_submodeling_no_record(gpu_daw):
    c = Case()
    sis, wf = gpu_daw.submodeling(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec, type_rec=1,
                                  want_wavefield=True)
    assert not sis.any() and not wf.any()


@pytest.mark.parametrize("ofs", [1, 0])
def test_esource_data_bitexact(gpu_daw, ofs):
    c = Case(ofs=ofs)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    _say("esource_data", ofs)
    a = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
    b = O.esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
    assert np.abs(b).max() > 0
    assert np.array_equal(a, b), rel_l2(a, b)


def test_esource_bitexact(gpu_daw):
    c = Case()
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    a, wa = gpu_daw.submodeling_adjoint_esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, want_wavefield=True)
    b, wb = O.esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, want_wavefield=True)
    assert np.array_equal(a, b), rel_l2(a, b)
    assert np.array_equal(wa, wb), rel_l2(wa, wb)


def test_resample_and_residuals_bitexact(gpu_daw):
    c = Case()
    wf = c.rng.standard_normal((c.nt, c.n2, c.n1)).astype(np.float32)
    a = gpu_daw.subwavefieldresample(wf, c.npml, c.dat, 1)
    b = O.resample(wf, c.npml, c.dat, 1)
    assert np.array_equal(a, b)
    # receivers near the edges exercise the interior clip (subwavefieldresample.f90:27-30)
    edge = O.acqui_table([3.0, 15.0, 380.0], [2.0, 885.0, 400.0], c.h, c.npml, 1)
    assert np.array_equal(gpu_daw.subwavefieldresample(wf, c.npml, edge, 1), O.resample(wf, c.npml, edge, 1))
    d1 = c.rng.standard_normal((7, 50)).astype(np.float32)
    d2 = c.rng.standard_normal((7, 50)).astype(np.float32)
    assert np.array_equal(gpu_daw.subresiduals(d1, d2), O.residuals(d1, d2))


def test_sublinearminres(gpu_daw):
    c = Case()
    dobs = O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
    dcal = O.forward(c.vp0, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
    dr = O.residuals(dobs, dcal)
    its = 6
    x, hist, info = gpu_daw.sublinearminres(c.vp0, c.npml, c.h, 1, c.nt, c.dt, dr, c.rec, 1e-2, its)
    x32, h32, i32 = O.linearminres(c.vp0, c.npml, c.h, 1, c.nt, c.dt, dr, c.rec, 1e-2, its)
    x64, h64, i64 = O.linearminres(c.vp0, c.npml, c.h, 1, c.nt, c.dt, dr, c.rec, 1e-2, its, prec="f64")
    _say("minres rel vs f32", rel_l2(x, x32), "vs f64", rel_l2(x, x64), "f32 vs f64", rel_l2(x32, x64))
    assert info["itn"] == its and info["istop"] == 5
    assert rel_l2(x, x64) <= 1e-4
    assert rel_l2(x, x32) <= 1e-4
    assert np.allclose(hist, h64, rtol=1e-4, atol=0)
    assert abs(info["fac"] / i64["fac"] - 1) < 1e-5


@pytest.mark.parametrize("tile", [(8, 128), (16, 64), (16, 128), (20, 64), (24, 64), (32, 64), (32, 128), (64, 64)])
def test_tile_shapes_bitexact(gpu_daw, tile):
    """Every step-kernel tile shape gives the reference's fields bit for bit."""
    c = Case(nz=45, nx=150, nt=160)
    gpu_daw.set_tile(*tile)
    try:
        res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
        a = gpu_daw.submodeling(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
        b = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    finally:
        gpu_daw.set_tile(0)
    assert np.array_equal(a, O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec))
    assert np.array_equal(b, O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec))
