"""The column-march step kernel (tile "R x 58": one wavefront per 58 owned
columns walking R rows, x-neighbours by lane shifts, no LDS) against the CPU
oracle, bit for bit: forward (both source types, both free-surface settings,
full wavefield), the Hessian-type operator (back-propagation with FP32
snapshots, then re-injection) and the source-extension operator.  The grid
of 45 x 150 (65 x 170 extended with the sponges) gives three strips, the
last one partial, and 3 to 9 row segments; the source and receiver rows and
the free-surface mirror row fall inside segments and on their edges.
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu

ROWS = [8, 12, 16, 24, 32]


def _say(*a):
    print(*a, flush=True)


@pytest.fixture
def march(gpu_daw, request):
    gpu_daw.set_tile(request.param, 58)
    assert gpu_daw.get_tile() == (request.param, 58)
    yield request.param
    gpu_daw.set_tile(0)


@pytest.mark.parametrize("march", ROWS, indirect=True)
@pytest.mark.parametrize("ofs,type_src", [(1, 0), (0, 0), (1, 1)])
def test_march_forward_bitexact(gpu_daw, march, ofs, type_src):
    c = Case(nz=45, nx=150, nt=160, ofs=ofs)
    _say("march forward", march, ofs, type_src)
    sis, wf = gpu_daw.submodeling(c.vp, c.npml, c.h, ofs, c.nt, c.dt, c.s, c.src, c.rec, type_src=type_src,
                                  want_wavefield=True)
    osis, owf = O.forward(c.vp, c.npml, c.h, ofs, c.nt, c.dt, c.s, c.src, c.rec, type_src=type_src,
                          want_wavefield=True)
    assert np.abs(osis).max() > 0
    assert np.array_equal(sis, osis), rel_l2(sis, osis)
    assert np.array_equal(wf, owf), rel_l2(wf, owf)


@pytest.mark.parametrize("march", ROWS, indirect=True)
@pytest.mark.parametrize("ofs", [1, 0])
def test_march_esource_data_bitexact(gpu_daw, march, ofs):
    c = Case(nz=45, nx=150, nt=160, ofs=ofs)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    _say("march esource_data", march, ofs)
    a = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
    b = O.esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
    assert np.abs(b).max() > 0
    assert np.array_equal(a, b), rel_l2(a, b)


@pytest.mark.parametrize("march", [8, 16], indirect=True)
def test_march_esource_bitexact(gpu_daw, march):
    c = Case(nz=45, nx=150, nt=160)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    _say("march esource", march)
    a, wa = gpu_daw.submodeling_adjoint_esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, want_wavefield=True)
    b, wb = O.esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, want_wavefield=True)
    assert np.array_equal(a, b), rel_l2(a, b)
    assert np.array_equal(wa, wb), rel_l2(wa, wb)


@pytest.mark.parametrize("march", [12], indirect=True)
def test_march_narrow_grid(gpu_daw, march):
    """A grid narrower than one strip (one partial strip per row segment)."""
    c = Case(nz=30, nx=24, nt=120, nrec=9)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    _say("march narrow", march)
    a = gpu_daw.submodeling(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
    assert np.array_equal(a, O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec))
    b = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    assert np.array_equal(b, O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec))


# ---- the two-step march of the re-injection (tile "R x 52": one walk
# advances two time steps, 52 owned columns per wavefront)

ROWS2 = [16, 24, 32, 40, 48]


@pytest.fixture
def march2(gpu_daw, request):
    gpu_daw.set_tile(request.param, 52)
    assert gpu_daw.get_tile() == (request.param, 52)
    yield request.param
    gpu_daw.set_tile(0)


@pytest.mark.parametrize("march2", ROWS2, indirect=True)
@pytest.mark.parametrize("ofs,nt", [(1, 160), (0, 160), (1, 161)])
def test_march2_esource_data_bitexact(gpu_daw, march2, ofs, nt):
    """Odd nt: the last step runs on the one-step march."""
    c = Case(nz=45, nx=150, nt=nt, ofs=ofs)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    _say("march2 esource_data", march2, ofs, nt)
    a = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
    assert gpu_daw.get_launch(2) == (2, march2, 52)
    b = O.esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
    assert np.abs(b).max() > 0
    assert np.array_equal(a, b), rel_l2(a, b)


@pytest.mark.parametrize("march2", [16, 32], indirect=True)
@pytest.mark.parametrize("nt", [160, 159])
def test_march2_esource_wavefield_bitexact(gpu_daw, march2, nt):
    """The source-extension operator records the full wavefield of both steps."""
    c = Case(nz=45, nx=150, nt=nt)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    _say("march2 esource", march2, nt)
    a, wa = gpu_daw.submodeling_adjoint_esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, want_wavefield=True)
    b, wb = O.esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, want_wavefield=True)
    assert np.array_equal(a, b), rel_l2(a, b)
    assert np.array_equal(wa, wb), rel_l2(wa, wb)


@pytest.mark.parametrize("march2", [16], indirect=True)
@pytest.mark.parametrize("npml", [19, 9, 21])
def test_march2_free_surface_segments(gpu_daw, march2, npml):
    """Free-surface mirror rows near segment edges: with npml = 19 the first
    16-row segment's step-A velocity rows would end at row npml-1 (the host
    shortens it by one row); 9 and 21 put the mirror rows inside the first and
    the second segment's rings."""
    c = Case(nz=45, nx=150, npml=npml, nt=120)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    _say("march2 free surface", npml)
    a = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    b = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    assert np.array_equal(a, b), rel_l2(a, b)


@pytest.mark.parametrize("march2", [24], indirect=True)
def test_march2_narrow_grid_and_other_sweeps(gpu_daw, march2):
    """A grid narrower than one 52-column strip; the forward sweep, which has
    no two-step march, falls back to 16x64 tiles."""
    c = Case(nz=30, nx=24, nt=121, nrec=9)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    _say("march2 narrow")
    a = gpu_daw.submodeling(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
    assert np.array_equal(a, O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec))
    assert gpu_daw.get_launch(0) == (1, 16, 64)
    b = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    assert np.array_equal(b, O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec))
