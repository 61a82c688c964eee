"""GPU parity of the batched entry points (many shots advanced per launch)
against the oracle run shot by shot, as sub_redatum_DAW.f90 runs them."""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu


def _shots(c, n):
    xmax = (c.n2 - 1) * float(c.h)
    return O.acqui_table(np.full(n, 10.0), np.linspace(0.3, 0.7, n) * xmax, c.h, c.npml, c.ofs)


def test_modeling_batch_bitexact(gpu_daw):
    import torch
    c = Case()
    src = _shots(c, 3)
    nr, nd = c.rec.shape[1], c.dat.shape[1]
    s_dev = torch.from_numpy(c.s).cuda()
    dobs = torch.zeros((3, nr, c.nt), dtype=torch.float32, device="cuda")
    datum = torch.zeros((3, nd, c.nt), dtype=torch.float32, device="cuda")
    print("modeling_batch 3 shots", flush=True)
    gpu_daw.modeling_batch(c.vp, c.npml, c.h, c.ofs, c.nt, c.dt, s_dev, src, c.rec, dobs, dat_tab=c.dat,
                           datum_dev=datum)
    dobs = dobs.cpu().numpy()
    datum = datum.cpu().numpy()
    for s in range(3):
        sis, wf = O.forward(c.vp, c.npml, c.h, c.ofs, c.nt, c.dt, c.s, src[:, s].copy(), c.rec, want_wavefield=True)
        assert np.array_equal(dobs[s], sis)
        # on-the-fly datum gather == subwavefieldresample of the stored wavefield
        assert np.array_equal(datum[s], O.resample(wf, c.npml, c.dat, c.ofs))


def test_redatum_batch_pipeline(gpu_daw):
    """Mode 1 (forward, residual, MINRES, source extension) + mode 2 datum
    gather for 3 shots at once vs the oracle pipeline per shot (FP32 and FP64),
    and batched == one-shot-at-a-time bit for bit."""
    import torch
    c = Case(nt=250)
    ns, its, mu = 3, 4, 1e-2
    src = _shots(c, ns)
    nr, nd = c.rec.shape[1], c.dat.shape[1]
    dobs_np = np.stack([O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, src[:, s].copy(), c.rec) for s in range(ns)])
    s_dev = torch.from_numpy(c.s).cuda()
    dobs = torch.from_numpy(dobs_np).cuda()
    outs = {}
    for mb in (0, 1):
        datum = torch.zeros((ns, nd, c.nt), dtype=torch.float32, device="cuda")
        datae = torch.zeros((ns, nr, c.nt), dtype=torch.float32, device="cuda")
        dde = torch.zeros((ns, nr, c.nt), dtype=torch.float32, device="cuda")
        print("redatum_batch max_batch", mb, flush=True)
        hist = gpu_daw.redatum_batch(c.vp0, c.npml, c.h, 1, c.nt, c.dt, s_dev, src, c.rec, c.dat, dobs, mu, its,
                                     datum_dev=datum, datae_dev=datae, deltade_dev=dde, max_batch=mb)
        outs[mb] = (datum.cpu().numpy(), datae.cpu().numpy(), dde.cpu().numpy(), hist)
    for k in range(3):
        assert np.array_equal(outs[0][k], outs[1][k])
    datum, datae, dde, hist = outs[0]
    for s in range(ns):
        row = src[:, s].copy()
        for prec, tol in (("f32", 1e-4), ("f64", 1e-4)):
            dr, wfr = O.forward(c.vp0, c.npml, c.h, 1, c.nt, c.dt, c.s, row, c.rec, want_wavefield=True, prec=prec)
            res = np.ascontiguousarray((dobs_np[s].astype(dr.dtype) - dr)[:, ::-1])
            x, h, info = O.linearminres(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, mu, its, prec=prec)
            esc, wfe = O.esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, x, c.rec, want_wavefield=True, prec=prec)
            ref_datum = O.resample(wfr + wfe, c.npml, c.dat, 1, prec=prec)
            e = (rel_l2(dde[s], x), rel_l2(datae[s], dr + esc), rel_l2(datum[s], ref_datum))
            print(s, prec, "deltade/datae/datum rel", e, flush=True)
            assert max(e) <= tol, (s, prec, e)
            assert np.allclose(hist[s], h, rtol=1e-4)


@pytest.mark.parametrize("nt,dt,ntr", [(300, 1e-3, 7), (3000, 8e-4, 7), (6000, 5e-4, 7), (9000, 4e-4, 7),
                                        (12000, 2.5e-4, 9), (20000, 2.5e-4, 5), (40000, 1e-4, 3), (65536, 1e-4, 2)])
def test_filter_device(gpu_daw, nt, dt, ntr):
    """filmartin2d on device traces against the oracle (the reference's NR
    FFT restated), within 1e-6: FFT lengths up to 8192 in LDS, 16384-65536
    as the four-step transform through HBM (odd trace counts leave the last
    pair with one trace)."""
    import torch
    rng = np.random.default_rng(nt)
    d = rng.standard_normal((ntr, nt)).astype(np.float32)
    band = (2.0, 25.0, 1.0, 5.0)
    x = torch.from_numpy(d).cuda()
    y = torch.empty_like(x)
    gpu_daw.filmartin2d_dev(x, y, dt, *band)
    ref = O.filmartin2d(d, dt, *band)
    assert rel_l2(y.cpu().numpy(), ref) < 1e-6
    gpu_daw.filmartin2d_dev(x, x, dt, *band)          # in place
    assert rel_l2(x.cpu().numpy(), ref) < 1e-6


def test_filmartin2d_dropin_golden(gpu_daw):
    """The host-array drop-in against the reference's own filmartin2d output
    (golden fixture written by the reference build), within 1e-6."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_small.npz"), allow_pickle=False)
    d = gold["filt_in"]
    assert rel_l2(gpu_daw.filmartin2d(d, 1e-3, 3.0, 30.0, 2.0, 10.0), gold["filt_out"]) < 1e-6
    assert rel_l2(gpu_daw.filmartin2d(d, 8e-4, 2.0, 700.0, 1.0, 5.0), gold["filt_out2"]) < 1e-6


@pytest.mark.parametrize("nt", [2500, 40000])
def test_filmartin2d_dropin(gpu_daw, nt):
    """The host-array drop-in filmartin2d runs on the GPU too, within 1e-6 of the oracle."""
    rng = np.random.default_rng(nt + 1)
    d = rng.standard_normal((5, nt)).astype(np.float32)
    band = (3.0, 30.0, 2.0, 10.0)
    got = gpu_daw.filmartin2d(d, 1e-3 if nt < 10000 else 2e-4, *band)
    ref = O.filmartin2d(d, 1e-3 if nt < 10000 else 2e-4, *band)
    assert rel_l2(got, ref) < 1e-6
