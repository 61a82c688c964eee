"""The exact discrete adjoint on the GPU (no reference counterpart: the
reference back-propagates with a sign-flipped forward scheme).

* submodeling_transpose: the dot test <F s, d> = <s, F^T d> with F the GPU
  forward, normalised by ||F s|| ||d|| (FP32 rounding in both operators), and
  agreement with the FP64 oracle transpose;
* the exact Hessian (daw_set_hessian(1)): agreement with the FP64 oracle's
  exact operator, and symmetry of Rev o H, which the reference's operator
  does not have.
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ofs", [1, 0])
def test_transpose_dot_product(gpu_daw, ofs):
    c = Case(nz=36, nx=80, nt=220, ofs=ofs, nrec=23)
    rng = np.random.default_rng(11)
    s = rng.standard_normal(c.nt).astype(np.float32)
    d = rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    print("transpose dot test ofs", ofs, flush=True)
    Fs = gpu_daw.submodeling(c.vp, c.npml, c.h, ofs, c.nt, c.dt, s, c.src, c.rec).astype(np.float64)
    g = gpu_daw.submodeling_transpose(c.vp, c.npml, c.h, ofs, c.nt, c.dt, d, c.src, c.rec).astype(np.float64)
    lhs, rhs = float((Fs * d).sum()), float((s.astype(np.float64) * g).sum())
    err = abs(lhs - rhs) / (np.linalg.norm(Fs) * np.linalg.norm(d))
    print("dot test", lhs, rhs, err, flush=True)
    assert err <= 1e-6
    g64 = O.forward_adjoint(c.vp, c.npml, c.h, ofs, c.nt, c.dt, d, c.src, c.rec, prec="f64")
    assert rel_l2(g, g64) <= 1e-5


def test_exact_hessian(gpu_daw):
    c = Case(nz=30, nx=70, nt=180, nrec=15)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    y = rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    gpu_daw.set_hessian(1)
    try:
        Hx = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, x, c.rec)
        Hy = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, y, c.rec)
    finally:
        gpu_daw.set_hessian(0)
    ref = O.esource_data_exact(c.vp0, c.npml, c.h, 1, c.nt, c.dt, x, c.rec, prec="f64")
    assert rel_l2(Hx, ref) <= 1e-5
    Ax, Ay = Hx[:, ::-1].astype(np.float64), Hy[:, ::-1].astype(np.float64)
    l, r = float((x * Ay).sum()), float((Ax * y).sum())
    err = abs(l - r) / (np.linalg.norm(Ax) * np.linalg.norm(y))
    print("exact Hessian symmetry", l, r, err, flush=True)
    assert err <= 1e-6
    # the default operator is the reference's one again
    Hd = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, x, c.rec)
    assert np.array_equal(Hd, O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, x, c.rec))


def test_transpose_tile_shapes(gpu_daw):
    """The adjoint step gives the same result for every tile shape."""
    c = Case(nz=45, nx=150, nt=120, nrec=11)
    rng = np.random.default_rng(2)
    d = rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    outs = []
    try:
        for tile in [(8, 128), (16, 64), (16, 128), (20, 64), (24, 64), (32, 64), (32, 128), (64, 64)]:
            gpu_daw.set_tile(*tile)
            outs.append(gpu_daw.submodeling_transpose(c.vp, c.npml, c.h, 1, c.nt, c.dt, d, c.src, c.rec))
    finally:
        gpu_daw.set_tile(0)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
