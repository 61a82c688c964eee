"""The two-step kernel (time steps it and it+1 in one launch, the wave state
staged in LDS with a 6-cell ring) gives the reference's fields bit for bit,
like the one-step kernel: forward with sources of both types and the
wavefield, the Hessian-type operators with FP32 and 16-bit snapshots, odd nt
(a last launch of one step), free-surface rows at different positions in the
tiles, and the batched pipeline.
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu

TILES = [(16, 64), (32, 64), (16, 128)]


@pytest.fixture
def two(gpu_daw):
    gpu_daw.set_steps(2)
    yield gpu_daw
    gpu_daw.set_steps(0)
    gpu_daw.set_tile(0)
    gpu_daw.set_snapshot_format("fp32")


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("npml,ofs,nt", [(10, 1, 161), (15, 1, 160), (12, 0, 161)])
def test_forward_two_step(two, tile, npml, ofs, nt):
    c = Case(nz=45, nx=150, npml=npml, ofs=ofs, nt=nt)
    two.set_tile(*tile)
    print("forward two-step", tile, npml, ofs, nt, flush=True)
    for type_src in (0, 1):
        sis, wf = two.submodeling(c.vp, c.npml, c.h, ofs, c.nt, c.dt, c.s, c.src, c.rec, type_src=type_src,
                                  want_wavefield=True)
        osis, owf = O.forward(c.vp, c.npml, c.h, ofs, c.nt, c.dt, c.s, c.src, c.rec, type_src=type_src,
                              want_wavefield=True)
        assert np.abs(osis).max() > 0
        assert np.array_equal(sis, osis), (type_src, rel_l2(sis, osis))
        assert np.array_equal(wf, owf), (type_src, rel_l2(wf, owf))


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("npml,ofs,nt", [(10, 1, 161), (15, 1, 160), (12, 0, 160)])
def test_hessian_two_step(two, tile, npml, ofs, nt):
    c = Case(nz=45, nx=150, npml=npml, ofs=ofs, nt=nt)
    two.set_tile(*tile)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    print("hessian two-step", tile, npml, ofs, nt, flush=True)
    a = two.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
    assert np.array_equal(a, O.esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec))
    e, we = two.submodeling_adjoint_esource(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec, want_wavefield=True)
    oe, owe = O.esource(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec, want_wavefield=True)
    assert np.array_equal(e, oe) and np.array_equal(we, owe)
    for snap in ("bf16", "fp16"):
        two.set_snapshot_format(snap)
        q = two.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
        assert np.array_equal(q, O.esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec, snap=snap)), snap
    two.set_snapshot_format("fp32")


def test_redatum_batch_two_step_equals_one_step(two):
    import torch
    c = Case(nt=251)
    ns, its, mu = 3, 3, 1e-2
    xmax = (c.n2 - 1) * float(c.h)
    src = O.acqui_table(np.full(ns, 10.0), np.linspace(0.3, 0.7, ns) * xmax, c.h, c.npml, 1)
    nr, nd = c.rec.shape[1], c.dat.shape[1]
    dobs_np = np.stack([O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, src[:, s].copy(), c.rec) for s in range(ns)])
    s_dev = torch.from_numpy(c.s).cuda()
    dobs = torch.from_numpy(dobs_np).cuda()
    outs = []
    for steps, graph in ((1, False), (2, False), (2, True)):
        two.set_steps(steps)
        two.set_graph(graph)
        two.set_tile(16, 64)
        datum = torch.zeros((ns, nd, c.nt), dtype=torch.float32, device="cuda")
        dde = torch.zeros((ns, nr, c.nt), dtype=torch.float32, device="cuda")
        print("redatum_batch steps", steps, "graph", graph, flush=True)
        hist = two.redatum_batch(c.vp0, c.npml, c.h, 1, c.nt, c.dt, s_dev, src, c.rec, c.dat, dobs, mu, its,
                                 datum_dev=datum, deltade_dev=dde)
        outs.append((datum.cpu().numpy(), dde.cpu().numpy(), hist))
    two.set_graph(False)
    assert np.abs(outs[0][0]).max() > 0
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
