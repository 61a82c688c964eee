"""The reference's own CPU-runnable configuration (121 x 401, h = 10 m,
20-cell PML, nt = 2500, 301 receivers, 1 shot, 30 MINRES iterations, datum
at z = 400 m) through the batched GPU pipeline, twice:

1. MINRES in the reference's arithmetic (set_minres_arith("ref32"): REAL dot
   products summed in dot_product's and datanorm's order, REAL scalars and
   vector updates) against the FP32 oracle, the reference's own CPU path
   restated operation for operation (pinned bit-exact against the reference
   build by tests/test_oracle.py).  The propagators are bit-exact and so
   is the solve: deltade and the MINRES history must equal the oracle's bit
   for bit.  The gathers are within 1e-6 (measured 1.3e-7): the GPU records
   u^r and u^e at the datum on the fly and adds the two gathers, the
   reference adds the wavefields and then resamples.
2. The default MINRES (FP64 dot products and scalars) against the FP64
   oracle of the plain method, within 1e-4 (measured: deltade 2.3e-5,
   gathers 7.0e-6, history 8.1e-6 max relative).

Quantities: deltade (the MINRES solution), the redatumed gathers of
u^r + u^e at the datum (unfiltered; the oracle resamples the summed
wavefields as subwavefieldresample does, the GPU records each part on the
fly), the MINRES history qrnorm/beta1 (relative L2 and max relative).
"""
import importlib

import numpy as np
import pytest

import oracle as O
from cases import rel_l2

pytestmark = pytest.mark.gpu

BAR = 1e-4


def _oracle_pipeline(w, c0, h, dt, s, src, rec, dat, dobs, prec):
    dr, wfr = O.forward(c0, w.npml, h, w.ofs, w.nt, dt, s, src[:, 0].copy(), rec, want_wavefield=True, prec=prec)
    if prec == "f32":
        res = O.residuals(dobs, dr)
    else:
        res = np.ascontiguousarray((dobs.astype(np.float64) - dr)[:, ::-1])
    x, hist, info = O.linearminres(c0, w.npml, h, w.ofs, w.nt, dt, res, rec, w.mu, w.maxiter, prec=prec)
    _, wfe = O.esource(c0, w.npml, h, w.ofs, w.nt, dt, x, rec, want_wavefield=True, prec=prec)
    wfr += wfe
    del wfe
    datum = O.resample(wfr, w.npml, dat, w.ofs, prec=prec)
    return x, np.asarray(hist, np.float64), datum


def _errors(got, want):
    (x, hist, datum), (xo, ho, do) = got, want
    return {"deltade_rel_l2": rel_l2(x, xo), "datum_rel_l2": rel_l2(datum, do),
            "hist_rel_l2": rel_l2(hist, ho), "hist_max_rel": float(np.max(np.abs(hist / ho - 1))),
            "deltade_bitexact": bool(np.array_equal(np.asarray(x, np.float32), np.asarray(xo, np.float32))),
            "hist_bitexact": bool(np.array_equal(np.asarray(hist, np.float32), np.asarray(ho, np.float32)))}


@pytest.mark.timeout(600)
def test_c1_pipeline_both_arithmetics(gpu_daw):
    import torch
    syn = importlib.import_module("daw_amd.synth")
    w = syn.workload(1)
    assert (w.nz, w.nx, w.nt, w.maxiter) == (121, 401, 2500, 30)
    vp, c0 = syn.models(w)
    h, dt = np.float32(w.h), np.float32(w.dt)
    src = O.acqui_table(w.src_z, w.src_x, h, w.npml, w.ofs)
    rec = O.acqui_table(w.rec_z, w.rec_x, h, w.npml, w.ofs)
    dat = O.acqui_table(w.dat_z, w.dat_x, h, w.npml, w.ofs)
    s = syn.ricker(w.nt, float(dt), w.f0, float(h))
    print("C1 oracle forward (observed data)", flush=True)
    dobs = O.forward(vp, w.npml, h, w.ofs, w.nt, dt, s, src[:, 0].copy(), rec)
    dobs = syn.add_noise(dobs, 2001)

    def gpu(arith):
        datum = torch.zeros((1, dat.shape[1], w.nt), dtype=torch.float32, device="cuda")
        dde = torch.zeros((1, rec.shape[1], w.nt), dtype=torch.float32, device="cuda")
        gpu_daw.set_minres_arith(arith)
        try:
            hist = gpu_daw.redatum_batch(c0, w.npml, h, w.ofs, w.nt, dt, torch.from_numpy(s).cuda(), src, rec, dat,
                                         torch.from_numpy(dobs[None]).cuda(), w.mu, w.maxiter, datum_dev=datum,
                                         deltade_dev=dde)
        finally:
            gpu_daw.set_minres_arith("f64")
        return dde.cpu().numpy()[0], hist[0], datum.cpu().numpy()[0]

    print("C1 GPU redatum_batch, reference arithmetic", flush=True)
    g32 = gpu("ref32")
    print("C1 GPU redatum_batch, FP64 MINRES", flush=True)
    g64 = gpu("f64")
    print("C1 FP32 oracle pipeline", flush=True)
    o32 = _oracle_pipeline(w, c0, h, dt, s, src, rec, dat, dobs, "f32")
    e32 = _errors(g32, o32)
    print("C1 ref32 GPU vs FP32 oracle:", e32, flush=True)
    print("C1 FP64 oracle pipeline", flush=True)
    o64 = _oracle_pipeline(w, c0, h, dt, s, src, rec, dat, dobs, "f64")
    e64 = _errors(g64, o64)
    print("C1 f64 GPU vs FP64 oracle:", e64, flush=True)
    print("C1 ref32 GPU vs FP64 oracle:", _errors(g32, o64), flush=True)
    print("C1 FP32 oracle vs FP64 oracle:", _errors(o32, o64), flush=True)
    assert e32["deltade_bitexact"] and e32["hist_bitexact"], e32
    assert e32["datum_rel_l2"] <= 1e-6, e32
    assert max(e64["deltade_rel_l2"], e64["datum_rel_l2"], e64["hist_rel_l2"], e64["hist_max_rel"]) <= BAR, e64
