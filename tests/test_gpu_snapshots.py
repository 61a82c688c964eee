"""GPU parity of the 16-bit snapshot store (BF16 / FP16 with a power-of-two
scale per 64-column segment of each snapshot row).

The Hessian-type operator with 16-bit snapshots is compared bit for bit with
the oracle's operator using the same codec (pinned in test_snapshot_codec),
and the batched redatuming pipeline with 16-bit snapshots against the FP64
oracle of the plain FP32 method (bound 1e-4 relative L2 for FP16, the bar for mixed precision;
BF16, 8 significant bits, is held to 1e-3 and its error printed).
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture
def snap16(gpu_daw):
    yield gpu_daw
    gpu_daw.set_snapshot_format("fp32")
    gpu_daw.set_krylov_format("fp32")


@pytest.mark.parametrize("snap", ["bf16", "fp16"])
@pytest.mark.parametrize("ofs", [1, 0])
def test_esource_data_16bit_bitexact(snap16, snap, ofs):
    c = Case(ofs=ofs, nx=150)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    print("esource_data", snap, "ofs", ofs, flush=True)
    snap16.set_snapshot_format(snap)
    a = snap16.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec)
    b = O.esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec, snap=snap)
    assert np.array_equal(a, b), rel_l2(a, b)
    f64 = O.esource_data(c.vp0, c.npml, c.h, ofs, c.nt, c.dt, res, c.rec, prec="f64")
    print(snap, "rel vs f64", rel_l2(a, f64), flush=True)


@pytest.mark.parametrize("tile", [(8, 128), (20, 64), (24, 64), (32, 64), (64, 64)])
def test_esource_data_16bit_tile_invariant(snap16, tile):
    c = Case(nz=45, nx=150, nt=160)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    snap16.set_snapshot_format("bf16")
    snap16.set_tile(*tile)
    try:
        a = snap16.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    finally:
        snap16.set_tile(0)
    assert np.array_equal(a, O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, snap="bf16"))


@pytest.mark.parametrize("snap,tol", [("fp16", 1e-4), ("bf16", 1e-3)])
def test_redatum_batch_16bit_vs_f64(snap16, snap, tol):
    import torch
    c = Case(nt=250)
    ns, its, mu = 3, 4, 1e-2
    xmax = (c.n2 - 1) * float(c.h)
    src = O.acqui_table(np.full(ns, 10.0), np.linspace(0.3, 0.7, ns) * xmax, c.h, c.npml, 1)
    nr, nd = c.rec.shape[1], c.dat.shape[1]
    dobs_np = np.stack([O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, src[:, s].copy(), c.rec) for s in range(ns)])
    s_dev = torch.from_numpy(c.s).cuda()
    dobs = torch.from_numpy(dobs_np).cuda()
    datum = torch.zeros((ns, nd, c.nt), dtype=torch.float32, device="cuda")
    dde = torch.zeros((ns, nr, c.nt), dtype=torch.float32, device="cuda")
    snap16.set_snapshot_format(snap)
    print("redatum_batch", snap, flush=True)
    hist = snap16.redatum_batch(c.vp0, c.npml, c.h, 1, c.nt, c.dt, s_dev, src, c.rec, c.dat, dobs, mu, its,
                                datum_dev=datum, deltade_dev=dde)
    datum, dde = datum.cpu().numpy(), dde.cpu().numpy()
    for s in range(ns):
        row = src[:, s].copy()
        dr, wfr = O.forward(c.vp0, c.npml, c.h, 1, c.nt, c.dt, c.s, row, c.rec, want_wavefield=True, prec="f64")
        res = np.ascontiguousarray((dobs_np[s].astype(np.float64) - dr)[:, ::-1])
        x, h, info = O.linearminres(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, mu, its, prec="f64")
        esc, wfe = O.esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, x, c.rec, want_wavefield=True, prec="f64")
        ref_datum = O.resample(wfr + wfe, c.npml, c.dat, 1, prec="f64")
        e = (rel_l2(dde[s], x), rel_l2(datum[s], ref_datum), float(np.max(np.abs(hist[s] / h - 1))))
        print(s, snap, "deltade / datum rel L2, max rel hist", e, flush=True)
        assert max(e) <= tol, (s, snap, e)


@pytest.mark.parametrize("kry,tol", [("fp16", 1e-3), ("bf16", 1e-2)])
def test_redatum_batch_16bit_krylov_vs_f64(snap16, kry, tol):
    """MINRES with r1, r2, w, w1, w2 stored in 16 bits (FP64 dot products and
    scalars): gathers and history against the FP64 oracle; the two shots of a
    batch equal the same shots redatumed one at a time (the chunk scales are
    per shot)."""
    import torch
    c = Case(nt=250)
    ns, its, mu = 2, 5, 1e-2
    xmax = (c.n2 - 1) * float(c.h)
    src = O.acqui_table(np.full(ns, 10.0), np.linspace(0.3, 0.7, ns) * xmax, c.h, c.npml, 1)
    nd = c.dat.shape[1]
    dobs_np = np.stack([O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, src[:, s].copy(), c.rec) for s in range(ns)])
    s_dev = torch.from_numpy(c.s).cuda()
    dobs = torch.from_numpy(dobs_np).cuda()
    snap16.set_krylov_format(kry)
    datum = torch.zeros((ns, nd, c.nt), dtype=torch.float32, device="cuda")
    print("redatum_batch, Krylov vectors", kry, flush=True)
    hist = snap16.redatum_batch(c.vp0, c.npml, c.h, 1, c.nt, c.dt, s_dev, src, c.rec, c.dat, dobs, mu, its,
                                datum_dev=datum)
    one = torch.zeros((ns, nd, c.nt), dtype=torch.float32, device="cuda")
    hist1 = snap16.redatum_batch(c.vp0, c.npml, c.h, 1, c.nt, c.dt, s_dev, src, c.rec, c.dat, dobs, mu, its,
                                 datum_dev=one, max_batch=1)
    assert torch.equal(datum, one) and np.array_equal(hist, hist1)
    datum = datum.cpu().numpy()
    for s in range(ns):
        dr, wfr = O.forward(c.vp0, c.npml, c.h, 1, c.nt, c.dt, c.s, src[:, s].copy(), c.rec, want_wavefield=True,
                            prec="f64")
        res = np.ascontiguousarray((dobs_np[s].astype(np.float64) - dr)[:, ::-1])
        x, h, _ = O.linearminres(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, mu, its, prec="f64")
        _, wfe = O.esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, x, c.rec, want_wavefield=True, prec="f64")
        ref = O.resample(wfr + wfe, c.npml, c.dat, 1, prec="f64")
        e = (rel_l2(datum[s], ref), float(np.max(np.abs(hist[s] / h - 1))))
        print(s, kry, "datum rel L2, max rel history vs FP64", e, flush=True)
        assert max(e) <= tol, (s, kry, e)
