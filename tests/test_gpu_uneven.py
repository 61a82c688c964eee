"""Uneven batches: shots on their own subdomains of the model, with their
own receiver sets and nt, advanced together in one launch per step.  Each
shot is checked against the oracle run on that shot's cropped model,
acquisition and nt (the reference's result for it), and an uneven batch of
identical shots equals the even batch bit for bit.
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu


def _shots(c):
    """Three subdomains of a 45 x 150 model with 31..47 receivers and nt 200..240."""
    x0 = np.array([0, 30, 62], np.int32)
    n2s = np.array([90, 110, 84], np.int32)
    nts = np.array([200, 240, 220], np.int32)
    h = float(c.h)
    src, recs, dats = [], [], []
    for k, (a, w) in enumerate(zip(x0, n2s)):
        xmax = (w - 1) * h
        src.append(O.acqui_table([10.0], [0.37 * xmax + 7.0 * k], c.h, c.npml, 1)[:, 0])
        nr = 31 + 8 * k
        rx = np.linspace(0.06, 0.94, nr) * xmax
        recs.append(O.acqui_table(np.full(nr, 15.0), rx, c.h, c.npml, 1))
        dats.append(O.acqui_table(np.full(nr - 5, 200.0), rx[:nr - 5], c.h, c.npml, 1))
    return x0, n2s, nts, np.stack(src, axis=1), recs, dats


def test_redatum_uneven_vs_oracle(gpu_daw):
    import torch
    c = Case(nz=45, nx=150, nt=240)
    x0, n2s, nts, src, recs, dats = _shots(c)
    mu, its = 1e-2, 3
    dobs_np = [O.forward(c.vp[a:a + w], c.npml, c.h, 1, int(nt), c.dt, c.s[:nt], src[:, k].copy(), recs[k])
               for k, (a, w, nt) in enumerate(zip(x0, n2s, nts))]
    dobs = torch.from_numpy(np.concatenate([d.ravel() for d in dobs_np])).cuda()
    nv = dobs.numel()
    nd = sum(d.shape[1] * int(nt) for d, nt in zip(dats, nts))
    datum = torch.zeros(nd, dtype=torch.float32, device="cuda")
    dde = torch.zeros(nv, dtype=torch.float32, device="cuda")
    s_dev = torch.from_numpy(c.s).cuda()
    print("redatum_uneven 3 shots", flush=True)
    hist = gpu_daw.redatum_uneven(c.vp0, c.npml, c.h, 1, c.dt, s_dev, x0, n2s, nts, src, recs, dobs, mu, its,
                                  dat_tabs=dats, datum_dev=datum, deltade_dev=dde)
    datum, dde = datum.cpu().numpy(), dde.cpu().numpy()
    ro = do = 0
    for k, (a, w, nt) in enumerate(zip(x0, n2s, nts)):
        nt = int(nt)
        vp0 = np.ascontiguousarray(c.vp0[a:a + w])
        nr, ndk = recs[k].shape[1], dats[k].shape[1]
        got_x = dde[ro:ro + nr * nt].reshape(nr, nt)
        got_d = datum[do:do + ndk * nt].reshape(ndk, nt)
        ro += nr * nt
        do += ndk * nt
        for prec in ("f32", "f64"):
            dr, wfr = O.forward(vp0, c.npml, c.h, 1, nt, c.dt, c.s[:nt], src[:, k].copy(), recs[k],
                                want_wavefield=True, prec=prec)
            res = np.ascontiguousarray((dobs_np[k].astype(dr.dtype) - dr)[:, ::-1])
            x, hh, _ = O.linearminres(vp0, c.npml, c.h, 1, nt, c.dt, res, recs[k], mu, its, prec=prec)
            esc, wfe = O.esource(vp0, c.npml, c.h, 1, nt, c.dt, x, recs[k], want_wavefield=True, prec=prec)
            ref = O.resample(wfr + wfe, c.npml, dats[k], 1, prec=prec)
            e = (rel_l2(got_x, x), rel_l2(got_d, ref))
            print(k, prec, "deltade / datum rel L2", e, flush=True)
            assert max(e) <= 1e-4, (k, prec, e)
            assert np.allclose(hist[k], hh, rtol=1e-4)


@pytest.mark.parametrize("method", [1, 3])
def test_uneven_of_even_shots_equals_even_batch(gpu_daw, method):
    """Shots that all span the whole model are an even batch: the uneven entry
    point gives the even one's results bit for bit, for the DAW (mode 1) and
    the Symes (mode 3) variants, MINRES log records included."""
    import torch
    c = Case(nt=200)
    ns, its, mu = 3, 3, 1e-2
    xmax = (c.n2 - 1) * float(c.h)
    src = O.acqui_table(np.full(ns, 10.0), np.linspace(0.3, 0.7, ns) * xmax, c.h, c.npml, 1)
    dobs_np = np.stack([O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, src[:, s].copy(), c.rec) for s in range(ns)])
    dobs = torch.from_numpy(dobs_np).cuda()
    s_dev = torch.from_numpy(c.s).cuda()
    nr, nd = c.rec.shape[1], c.dat.shape[1]
    d1 = torch.zeros((ns, nd, c.nt), dtype=torch.float32, device="cuda")
    x1 = torch.zeros((ns, nr, c.nt), dtype=torch.float32, device="cuda")
    l1 = np.zeros((ns, 9 + 5 * its))
    l2 = np.zeros((ns, 9 + 5 * its))
    print("uneven vs even, method", method, flush=True)
    h1 = gpu_daw.redatum_batch(c.vp0, c.npml, c.h, 1, c.nt, c.dt, s_dev, src, c.rec, c.dat, dobs, mu, its,
                               datum_dev=d1, deltade_dev=x1, method=method, minres_log=l1)
    d2 = torch.zeros(ns * nd * c.nt, dtype=torch.float32, device="cuda")
    x2 = torch.zeros(ns * nr * c.nt, dtype=torch.float32, device="cuda")
    h2 = gpu_daw.redatum_uneven(c.vp0, c.npml, c.h, 1, c.dt, s_dev, [0] * ns, [c.n2] * ns, [c.nt] * ns, src,
                                [c.rec] * ns, dobs.reshape(-1), mu, its, dat_tabs=[c.dat] * ns, datum_dev=d2,
                                deltade_dev=x2, method=method, minres_log=l2)
    assert np.abs(d1.cpu().numpy()).max() > 0
    assert np.array_equal(d1.cpu().numpy().ravel(), d2.cpu().numpy())
    assert np.array_equal(x1.cpu().numpy().ravel(), x2.cpu().numpy())
    assert np.array_equal(h1, h2)
    assert np.array_equal(l1, l2) and np.all(l1[:, 8] == its)
    if method == 3:   # the Symes solve's right-hand side is the observed data itself (:175-179)
        for s in range(ns):
            x, hist, info = O.linearminres(c.vp0, c.npml, c.h, 1, c.nt, c.dt, dobs_np[s][:, ::-1].copy(), c.rec, mu,
                                           its, prec="f64")
            assert rel_l2(x1[s].cpu().numpy(), x) <= 1e-4, s


@pytest.mark.parametrize("tile", [(24, 52), (0, 64)], ids=["two-step-march", "autotuned"])
def test_uneven_odd_nt_shots_alone_equal_batch(gpu_daw, tile):
    """Shots of odd and even nt in one uneven batch give, bit for bit, what
    each gives alone: with the two-step march of the re-injection, a shot
    whose nt is odd takes its last step alone inside a two-step launch."""
    import torch
    c = Case(nz=45, nx=150, nt=241)
    x0, n2s, _, src, recs, dats = _shots(c)
    nts = np.array([201, 240, 219], np.int32)
    mu, its = 1e-2, 2
    dobs_np = [O.forward(c.vp[a:a + w], c.npml, c.h, 1, int(nt), c.dt, c.s[:nt], src[:, k].copy(), recs[k])
               for k, (a, w, nt) in enumerate(zip(x0, n2s, nts))]
    s_dev = torch.from_numpy(c.s).cuda()

    def run(ks):
        dobs = torch.from_numpy(np.concatenate([dobs_np[k].ravel() for k in ks])).cuda()
        nd = sum(dats[k].shape[1] * int(nts[k]) for k in ks)
        datum = torch.zeros(nd, dtype=torch.float32, device="cuda")
        dde = torch.zeros(dobs.numel(), dtype=torch.float32, device="cuda")
        gpu_daw.redatum_uneven(c.vp0, c.npml, c.h, 1, c.dt, s_dev, x0[ks], n2s[ks], nts[ks],
                               np.ascontiguousarray(src[:, ks]), [recs[k] for k in ks], dobs, mu, its,
                               dat_tabs=[dats[k] for k in ks], datum_dev=datum, deltade_dev=dde)
        return datum.cpu().numpy(), dde.cpu().numpy()

    gpu_daw.set_tile(*tile)
    try:
        print("uneven odd nt", tile, flush=True)
        datum, dde = run([0, 1, 2])
        ro = do = 0
        for k in range(3):
            nr, ndk, nt = recs[k].shape[1], dats[k].shape[1], int(nts[k])
            d1, x1 = run([k])
            assert np.array_equal(dde[ro:ro + nr * nt], x1), (tile, k)
            assert np.array_equal(datum[do:do + ndk * nt], d1), (tile, k)
            ro += nr * nt
            do += ndk * nt
    finally:
        gpu_daw.set_tile(0)
