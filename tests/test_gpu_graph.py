"""Captured sweeps (hipGraph) give the same results as step-by-step launches.

One redatum_batch call with 4 MINRES iterations replays the captured
back-propagation and re-injection sweeps 5 times each, with the solver's own
eager kernels, host uploads of the shot descriptors and record items, and
stream synchronisations between the replays.  Every replay feeds the next
Lanczos step, so bit-identical outputs check each replay against its eager
counterpart.
"""
import numpy as np
import pytest

import oracle as O
from cases import Case

pytestmark = pytest.mark.gpu


@pytest.fixture
def graph_daw(gpu_daw):
    yield gpu_daw
    gpu_daw.set_graph(False)
    gpu_daw.set_snapshot_format("fp32")


@pytest.mark.parametrize("snap", ["fp32", "bf16"])
def test_redatum_batch_graph_equals_eager(graph_daw, snap):
    import torch
    c = Case(nt=250)
    ns, its, mu = 3, 4, 1e-2
    xmax = (c.n2 - 1) * float(c.h)
    src = O.acqui_table(np.full(ns, 10.0), np.linspace(0.3, 0.7, ns) * xmax, c.h, c.npml, 1)
    nr, nd = c.rec.shape[1], c.dat.shape[1]
    dobs_np = np.stack([O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, src[:, s].copy(), c.rec) for s in range(ns)])
    s_dev = torch.from_numpy(c.s).cuda()
    dobs = torch.from_numpy(dobs_np).cuda()
    graph_daw.set_snapshot_format(snap)
    outs = []
    for graph in (False, True, True):
        graph_daw.set_graph(graph)
        datum = torch.zeros((ns, nd, c.nt), dtype=torch.float32, device="cuda")
        datae = torch.zeros((ns, nr, c.nt), dtype=torch.float32, device="cuda")
        dde = torch.zeros((ns, nr, c.nt), dtype=torch.float32, device="cuda")
        print("redatum_batch graph", graph, snap, flush=True)
        hist = graph_daw.redatum_batch(c.vp0, c.npml, c.h, 1, c.nt, c.dt, s_dev, src, c.rec, c.dat, dobs, mu, its,
                                       datum_dev=datum, datae_dev=datae, deltade_dev=dde)
        outs.append((datum.cpu().numpy(), datae.cpu().numpy(), dde.cpu().numpy(), hist))
    assert np.abs(outs[0][0]).max() > 0
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)
