"""GPU tests at the workloads of the larger configurations.

Configuration 3 (351 x 1401, h = 6.25 m, 20 Hz, dt = 0.5 ms, faulted model
with a high-velocity body; shots on cropped subdomains of 480-960 columns
with one receiver per column and nt 4000-6000), sampled shots of its
128-shot line at their full sizes:
  - the forward sweep of the uneven batch (datae with no MINRES iteration is
    datar) bit for bit against the FP32 oracle run on each shot's cropped
    model, receivers and nt -- what the reference computes for that shot;
  - batch split and shot order change nothing, bit for bit;
  - the smallest shot's pipeline (3 MINRES iterations): in the reference's
    arithmetic bit-exact against the FP32 oracle, by default within 1e-4 of
    the FP64 oracle.
Configuration 4 (the 201 x 801, nt = 3000 workload with 16-bit snapshots):
  - the Hessian-type operator with BF16 and FP16 snapshots bit for bit
    against the oracle's operator with the same codec, on one shot's residual;
  - one shot's redatuming (30 MINRES iterations) with BF16 and FP16 snapshots,
    and with 16-bit MINRES Krylov vectors as well, against the FP64 oracle of
    the plain method (bounds at COMBOS; every error printed).
Configuration 5 (801 x 3201, h = 3.125 m, nt = 12000, BF16 snapshots):
  - the admission of shots by the HBM budget: the largest shots of the line
    at full size run with the default budget and with a budget that admits
    one shot at a time, bit-identical;
  - the forward sweep of one cropped shot at truncated nt bit-exact against
    the FP32 oracle;
  - the band-pass at nt = 12000 (FFT length 16384) on device traces within
    1e-6 of the oracle.
"""
import importlib

import numpy as np
import pytest

import oracle as O
from cases import rel_l2

pytestmark = pytest.mark.gpu


def _syn():
    return importlib.import_module("daw_amd.synth")


class Line:
    """Sampled shots of an uneven configuration's line, with tables in
    subdomain coordinates (subacqui on each shot's subdomain)."""

    def __init__(self, cfg, idx, nt_cap=None):
        syn = _syn()
        w = syn.uneven_workload(cfg)
        self.w = w
        self.vp, self.c0 = syn.uneven_models(w)
        self.h, self.dt = np.float32(w.h), np.float32(w.dt)
        self.idx = np.asarray(idx)
        self.x0, self.n2s = w.x0[self.idx], w.n2s[self.idx]
        self.nts = w.nts[self.idx] if nt_cap is None else np.minimum(w.nts[self.idx], nt_cap).astype(np.int32)
        self.s = syn.ricker(int(self.nts.max()), float(self.dt), w.f0, float(self.h))
        self.src, self.recs, self.dats = [], [], []
        for k, i in enumerate(self.idx):
            rx = np.arange(self.n2s[k]) * w.rec_dx
            self.src.append(O.acqui_table([w.src_z], [w.src_x[i]], self.h, w.npml, w.ofs)[:, 0])
            self.recs.append(O.acqui_table(np.full(rx.size, w.rec_z), rx, self.h, w.npml, w.ofs))
            dx = rx[::4]
            self.dats.append(O.acqui_table(np.full(dx.size, w.dat_z), dx, self.h, w.npml, w.ofs))
        self.src = np.stack(self.src, axis=1)

    def crop(self, model, k):
        return np.ascontiguousarray(model[self.x0[k]:self.x0[k] + self.n2s[k]])

    def dobs(self, daw):
        """Observed data of every sampled shot in the true model (GPU forward, bit-exact with the oracle's)."""
        import torch
        s_dev = torch.from_numpy(self.s).cuda()
        out = []
        for k in range(len(self.idx)):
            nt = int(self.nts[k])
            d = torch.zeros((1, self.recs[k].shape[1], nt), dtype=torch.float32, device="cuda")
            daw.modeling_batch(self.crop(self.vp, k), self.w.npml, self.h, self.w.ofs, nt, self.dt, s_dev[:nt],
                               self.src[:, k:k + 1].copy(), self.recs[k], d)
            out.append(d.reshape(-1))
        return torch.cat(out)

    def run(self, daw, dobs, maxiter, order=None, max_batch=0, budget=0.0):
        """redatum_uneven over the sampled shots (in `order`); returns per-shot
        (datae, deltade, datum) host arrays and the MINRES histories, in the
        line's order."""
        import torch
        order = np.arange(len(self.idx)) if order is None else np.asarray(order)
        nv = [self.recs[k].shape[1] * int(self.nts[k]) for k in range(len(self.idx))]
        nd = [self.dats[k].shape[1] * int(self.nts[k]) for k in range(len(self.idx))]
        offs = np.concatenate([[0], np.cumsum(nv)])
        dob = torch.cat([dobs[offs[k]:offs[k + 1]] for k in order])
        s_dev = torch.from_numpy(self.s).cuda()
        datae = torch.zeros(sum(nv), dtype=torch.float32, device="cuda")
        dde = torch.zeros(sum(nv), dtype=torch.float32, device="cuda")
        datum = torch.zeros(sum(nd), dtype=torch.float32, device="cuda")
        hist = daw.redatum_uneven(self.c0, self.w.npml, self.h, self.w.ofs, self.dt, s_dev, self.x0[order],
                                  self.n2s[order], self.nts[order], self.src[:, order].copy(),
                                  [self.recs[k] for k in order], dob, self.w.mu, maxiter,
                                  dat_tabs=[self.dats[k] for k in order], datum_dev=datum, datae_dev=datae,
                                  deltade_dev=dde, max_batch=max_batch, hbm_budget_bytes=budget)
        datae, dde, datum = datae.cpu().numpy(), dde.cpu().numpy(), datum.cpu().numpy()
        res, ro, do = {}, 0, 0
        for j, k in enumerate(order):
            nr, ndk, nt = self.recs[k].shape[1], self.dats[k].shape[1], int(self.nts[k])
            res[k] = (datae[ro:ro + nr * nt].reshape(nr, nt), dde[ro:ro + nr * nt].reshape(nr, nt),
                      datum[do:do + ndk * nt].reshape(ndk, nt), hist[j])
            ro += nr * nt
            do += ndk * nt
        return [res[k] for k in range(len(self.idx))]


def _c3_sample():
    w = _syn().uneven_workload(3)
    cost = w.n2s.astype(np.int64) * w.nts
    lo, hi = int(np.argmin(cost)), int(np.argmax(cost))
    mid = int(np.argsort(cost)[len(cost) // 2])
    return [lo, mid, hi], w


@pytest.mark.timeout(600)
def test_c3_forward_bitexact(gpu_daw):
    idx, w = _c3_sample()
    L = Line(3, idx)
    print("C3 shots", list(idx), "columns", list(L.n2s), "nt", list(L.nts), flush=True)
    dobs = L.dobs(gpu_daw)
    out = L.run(gpu_daw, dobs, 0)
    for k in range(len(idx)):
        nt = int(L.nts[k])
        print(f"C3 oracle forward, shot {idx[k]}: 351x{L.n2s[k]}, nt={nt}", flush=True)
        ref = O.forward(L.crop(L.c0, k), w.npml, L.h, w.ofs, nt, L.dt, L.s[:nt], L.src[:, k].copy(), L.recs[k])
        assert np.array_equal(out[k][0], ref), (idx[k], rel_l2(out[k][0], ref))


@pytest.mark.timeout(600)
def test_c3_batch_split_and_order(gpu_daw):
    w = _syn().uneven_workload(3)
    idx = np.round(np.linspace(0, w.nshots - 1, 6)).astype(int)
    L = Line(3, idx)
    dobs = L.dobs(gpu_daw)
    print("C3 six shots, one sub-batch", flush=True)
    a = L.run(gpu_daw, dobs, 2)
    print("C3 six shots, sub-batches of two", flush=True)
    b = L.run(gpu_daw, dobs, 2, max_batch=2)
    print("C3 six shots, reversed, sub-batches of three", flush=True)
    c = L.run(gpu_daw, dobs, 2, order=np.arange(6)[::-1], max_batch=3)
    for k in range(6):
        for other in (b, c):
            for j in range(3):
                assert np.array_equal(a[k][j], other[k][j]), (k, j)
            assert np.array_equal(a[k][3], other[k][3])


@pytest.mark.timeout(900)
def test_c3_smallest_shot_pipeline(gpu_daw):
    idx, w = _c3_sample()
    L = Line(3, idx[:1])
    its = 3
    dobs = L.dobs(gpu_daw)
    gpu_daw.set_minres_arith("ref32")
    try:
        g32 = L.run(gpu_daw, dobs, its)[0]
    finally:
        gpu_daw.set_minres_arith("f64")
    g64 = L.run(gpu_daw, dobs, its)[0]
    nt = int(L.nts[0])
    c0 = L.crop(L.c0, 0)
    d = dobs.cpu().numpy().reshape(L.recs[0].shape[1], nt)
    for prec, got in (("f32", g32), ("f64", g64)):
        print(f"C3 {prec} oracle pipeline, 351x{L.n2s[0]}, nt={nt}", flush=True)
        dr, wfr = O.forward(c0, w.npml, L.h, w.ofs, nt, L.dt, L.s[:nt], L.src[:, 0].copy(), L.recs[0],
                            want_wavefield=True, prec=prec)
        res = O.residuals(d, dr) if prec == "f32" else np.ascontiguousarray((d.astype(np.float64) - dr)[:, ::-1])
        x, hh, _ = O.linearminres(c0, w.npml, L.h, w.ofs, nt, L.dt, res, L.recs[0], w.mu, its, prec=prec)
        _, wfe = O.esource(c0, w.npml, L.h, w.ofs, nt, L.dt, x, L.recs[0], want_wavefield=True, prec=prec)
        wfr += wfe
        del wfe
        ref = O.resample(wfr, w.npml, L.dats[0], w.ofs, prec=prec)
        del wfr
        e = (rel_l2(got[1], x), rel_l2(got[2], ref), float(np.max(np.abs(got[3] / hh - 1))))
        print(f"C3 smallest shot, {prec}: deltade / datum rel L2, max rel history", e, flush=True)
        if prec == "f32":
            assert np.array_equal(got[1], x) and np.array_equal(got[3].astype(np.float32), hh.astype(np.float32))
            assert e[1] <= 1e-6
        else:
            assert max(e) <= 1e-4


def _c2_shot():
    syn = _syn()
    w = syn.workload(2)
    vp, c0 = syn.models(w)
    h, dt = np.float32(w.h), np.float32(w.dt)
    src = O.acqui_table(w.src_z, w.src_x, h, w.npml, w.ofs)[:, 31:32].copy()      # a shot mid-line
    rec = O.acqui_table(w.rec_z, w.rec_x, h, w.npml, w.ofs)
    dat = O.acqui_table(w.dat_z, w.dat_x, h, w.npml, w.ofs)
    s = syn.ricker(w.nt, float(dt), w.f0, float(h))
    return w, vp, c0, h, dt, src, rec, dat, s


@pytest.fixture
def snap16(gpu_daw):
    yield gpu_daw
    gpu_daw.set_snapshot_format("fp32")
    gpu_daw.set_krylov_format("fp32")


@pytest.mark.timeout(600)
def test_c4_hessian_16bit_bitexact(snap16):
    w, vp, c0, h, dt, src, rec, dat, s = _c2_shot()
    dobs = O.forward(vp, w.npml, h, w.ofs, w.nt, dt, s, src[:, 0].copy(), rec)
    res = O.residuals(dobs, O.forward(c0, w.npml, h, w.ofs, w.nt, dt, s, src[:, 0].copy(), rec))
    for snap in ("bf16", "fp16"):
        print("C4 Hessian-type operator,", snap, "snapshots, 201x801 nt=3000", flush=True)
        snap16.set_snapshot_format(snap)
        a = snap16.submodeling_adjoint_esource_data(c0, w.npml, h, w.ofs, w.nt, dt, res, rec)
        b = O.esource_data(c0, w.npml, h, w.ofs, w.nt, dt, res, rec, snap=snap)
        assert np.array_equal(a, b), (snap, rel_l2(a, b))


# (snapshots, Krylov vectors) -> bound on the errors against the FP64 oracle.
# FP16 snapshots with FP32 Krylov vectors meet the 1e-4 bar (measured 2.3e-5);
# BF16 snapshots (8 significant bits) give 1.8e-4.  16-bit Krylov vectors
# perturb the Lanczos vectors themselves, and the error grows with the
# iterations: measured 1.5e-3 (gathers) and 1.0e-3 (history) for FP16 at 30
# iterations; 1.2e-2 and 7.0e-2 for BF16.  They stay an option, not the default.
COMBOS = {("fp16", "fp32"): 1e-4, ("bf16", "fp32"): 2.5e-4, ("fp16", "fp16"): 3e-3, ("bf16", "bf16"): 0.1}


@pytest.mark.timeout(900)
def test_c4_pipeline_16bit_vs_f64(snap16):
    import torch
    w, vp, c0, h, dt, src, rec, dat, s = _c2_shot()
    dobs = O.forward(vp, w.npml, h, w.ofs, w.nt, dt, s, src[:, 0].copy(), rec)
    dobs = _syn().add_noise(dobs, 2002)
    got = {}
    for snap, kry in COMBOS:
        print("C4 redatum_batch, one shot,", snap, "snapshots,", kry, "Krylov vectors, 30 MINRES iterations",
              flush=True)
        snap16.set_snapshot_format(snap)
        snap16.set_krylov_format(kry)
        datum = torch.zeros((1, dat.shape[1], w.nt), dtype=torch.float32, device="cuda")
        hist = snap16.redatum_batch(c0, w.npml, h, w.ofs, w.nt, dt, torch.from_numpy(s).cuda(), src, rec, dat,
                                    torch.from_numpy(dobs[None]).cuda(), w.mu, w.maxiter, datum_dev=datum)
        got[(snap, kry)] = (datum.cpu().numpy()[0], hist[0])
    print("C4 FP64 oracle pipeline", flush=True)
    dr, wfr = O.forward(c0, w.npml, h, w.ofs, w.nt, dt, s, src[:, 0].copy(), rec, want_wavefield=True, prec="f64")
    res = np.ascontiguousarray((dobs.astype(np.float64) - dr)[:, ::-1])
    x, hh, _ = O.linearminres(c0, w.npml, h, w.ofs, w.nt, dt, res, rec, w.mu, w.maxiter, prec="f64")
    _, wfe = O.esource(c0, w.npml, h, w.ofs, w.nt, dt, x, rec, want_wavefield=True, prec="f64")
    wfr += wfe
    del wfe
    ref = O.resample(wfr, w.npml, dat, w.ofs, prec="f64")
    for combo in COMBOS:
        d, hist = got[combo]
        e = (rel_l2(d, ref), rel_l2(hist, hh), float(np.max(np.abs(hist / hh - 1))))
        print(f"C4 {combo[0]} snapshots, {combo[1]} Krylov vectors: datum rel L2, history rel L2, "
              f"max rel history vs FP64", e, flush=True)
        assert max(e) <= COMBOS[combo], (combo, e)


@pytest.mark.timeout(600)
def test_c5_admission_by_hbm_budget(gpu_daw):
    import torch
    w = _syn().uneven_workload(5)
    idx = np.argsort(w.n2s)[-3:]                  # the widest subdomains of the line
    L = Line(5, idx)
    dobs = torch.zeros(sum(r.shape[1] * int(nt) for r, nt in zip(L.recs, L.nts)), dtype=torch.float32, device="cuda")
    gpu_daw.set_snapshot_format("bf16")
    try:
        print("C5 three widest shots, 801x", list(L.n2s), "nt=12000, BF16 snapshots, default budget", flush=True)
        a = L.run(gpu_daw, dobs, 1)
        nb_a = gpu_daw.stats()["max_batch_used"]
        # the default budget (85 % of free HBM) must admit what fits: a shot's
        # BF16 snapshot store is ~2*3*801*n2s*12000 bytes
        per_shot = 6.0 * 801 * float(L.n2s.max()) * 12000
        free, total = torch.cuda.mem_get_info()
        print("C5 admitted", nb_a, "shots per sub-batch; per-shot snapshots %.1f GB; free %.1f GB" %
              (per_shot / 1e9, free / 1e9), flush=True)
        assert 1 <= nb_a <= 3 and nb_a * per_shot <= 0.85 * total
        print("C5 budget for one shot at a time", flush=True)
        b = L.run(gpu_daw, dobs, 1, budget=1.3 * per_shot)
        assert gpu_daw.stats()["max_batch_used"] == 1
    finally:
        gpu_daw.set_snapshot_format("fp32")
    for k in range(3):
        for j in range(4):
            assert np.array_equal(a[k][j], b[k][j]), (k, j)


@pytest.mark.timeout(600)
def test_c5_forward_bitexact_truncated(gpu_daw):
    w = _syn().uneven_workload(5)
    k0 = int(np.argsort(w.n2s)[len(w.n2s) // 2])
    L = Line(5, [k0], nt_cap=2000)
    dobs = L.dobs(gpu_daw)
    out = L.run(gpu_daw, dobs, 0)
    print(f"C5 oracle forward, 801x{L.n2s[0]}, nt=2000", flush=True)
    ref = O.forward(L.crop(L.c0, 0), w.npml, L.h, w.ofs, 2000, L.dt, L.s[:2000], L.src[:, 0].copy(), L.recs[0])
    assert np.array_equal(out[0][0], ref), rel_l2(out[0][0], ref)


def test_c5_bandpass_nt12000(gpu_daw):
    import torch
    rng = np.random.default_rng(12)
    d = rng.standard_normal((401, 12000)).astype(np.float32)
    band = (3.0, 40.0, 1.0, 5.0)
    dev = torch.from_numpy(d).cuda()
    gpu_daw.filmartin2d_dev(dev, dev, 2.5e-4, *band)
    got = dev.cpu().numpy()
    ref = O.filmartin2d(d, 2.5e-4, *band)
    e = rel_l2(got, ref)
    print("C5 band-pass nt=12000 (FFT 16384) rel L2 vs oracle", e, flush=True)
    assert e <= 1e-6
