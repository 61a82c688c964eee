"""Hicks records two ways, each bit for bit against the CPU oracle: record
blocks inside the step launches, and batched record launches every few steps
(workgroups staging the band of up to 64 neighbouring points in LDS), for the
forward sweep (acquisition and, through the batched pipeline, datum
receivers) and the Hessian-type operator's data record, on the tile kernel and
on the column march.
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 1], ids=["in-launch", "batched"])
def rec_mode(gpu_daw, request):
    gpu_daw.set_record_batch(request.param)
    yield request.param
    gpu_daw.set_record_batch(1)


@pytest.mark.parametrize("tile", [(16, 64), (12, 58)], ids=["tile16x64", "march12"])
def test_records_bitexact(gpu_daw, rec_mode, tile):
    c = Case(nz=45, nx=150, nt=170, nrec=61)
    gpu_daw.set_tile(*tile)
    try:
        res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
        print("records", rec_mode, tile, flush=True)
        a = gpu_daw.submodeling(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
        b = gpu_daw.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    finally:
        gpu_daw.set_tile(0)
    oa = O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
    ob = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    assert np.abs(oa).max() > 0 and np.abs(ob).max() > 0
    assert np.array_equal(a, oa), rel_l2(a, oa)
    assert np.array_equal(b, ob), rel_l2(b, ob)


def test_records_edge_receivers(gpu_daw, rec_mode):
    """Receivers whose windows are clipped by the grid edges (both clip rules
    through the datum record of subwavefieldresample) and sparse receivers
    that split into several record groups."""
    c = Case(nz=40, nx=90, nt=150)
    rx = np.array([0.0, 5.0, 13.0, 300.0, 610.0, 880.0, 890.0])
    rec = O.acqui_table(np.full(rx.size, 12.0), rx, c.h, c.npml, 1)
    sis = gpu_daw.submodeling(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, rec)
    assert np.array_equal(sis, O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, rec))
