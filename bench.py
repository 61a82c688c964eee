"""Benchmark: batched DAW redatuming on one MI355X.

Workload (default): 201 x 801 grid, h = 8 m, 20-cell PML, nt = 3000,
dt = 0.8 ms, 10 Hz Ricker, 64 shots from x = 400 m at 50 m spacing, a fixed
401-receiver spread, 30 MINRES iterations, datum at z = 400 m -- the
reference's redatuming workflow (sub_redatum_DAW.f90 modes 1 and 2) for all
64 shots.  Observed data are modelled beforehand in the true model with 1 %
seeded noise (untimed input synthesis, the reference's mode 0).

One step = one redatuming pass over one batch of up to 24 shots advanced
together (the launch batch the library admits at this size: 3 shots per XCD,
wave state within the last-level cache; the 64 FP32 snapshot stores would
need 371 GB): the 64-shot line is batches of 24, 24 and 16 shots, which the
steps walk in turn.  Per batch: background forward, residual, fac, 30 MINRES iterations
(each one back-propagation + re-injection sweep pair), the source-extension
sweep pair, the datum gathers and their band-pass (filmartin2d, as mode 2
writes dataobs_redatum), with inputs resident in HBM.  The line also reports the plain
one-shot-at-a-time GPU path, the roofline of the dominant step kernel (HIP
events on the library's stream) and the reference's own CPU path timed on
this host's cores (one shot per core, truncated nt, scaled).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
T_PROC0 = time.perf_counter()     # the time budget counts from process start (first torch import included)
sys.path.insert(0, ROOT)


def _pkg():
    import __graft_entry__ as ge
    daw = ge._load_pkg()
    import importlib
    return daw, importlib.import_module("daw_amd.synth")


def _log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def setup(daw, syn, cfg, shots=None, nt=None, maxiter=None):
    import torch
    w = syn.workload(cfg)
    if nt:
        w.nt = nt
    if maxiter is not None:
        w.maxiter = maxiter
    vp, c0 = syn.models(w)
    h, dt = np.float32(w.h), np.float32(w.dt)
    nsh = w.nshots if shots is None else shots
    src = daw.subacqui(w.src_z[:nsh], w.src_x[:nsh], h, w.npml, w.ofs)
    rec = daw.subacqui(w.rec_z, w.rec_x, h, w.npml, w.ofs)
    dat = daw.subacqui(w.dat_z, w.dat_x, h, w.npml, w.ofs)
    s = syn.ricker(w.nt, float(dt), w.f0, float(h))
    s_dev = torch.from_numpy(s).cuda()
    dobs = torch.zeros((nsh, rec.shape[1], w.nt), dtype=torch.float32, device="cuda")
    dnew = torch.zeros((nsh, dat.shape[1], w.nt), dtype=torch.float32, device="cuda")
    daw.modeling_batch(vp, w.npml, h, w.ofs, w.nt, dt, s_dev, src, rec, dobs, dat_tab=dat, datum_dev=dnew)
    # data recorded at the datum in the true model, band-passed like the redatumed gathers
    dnew2 = dnew.view(-1, w.nt)
    daw.filmartin2d_dev(dnew2, dnew2, float(dt), *w.band)
    g = torch.Generator(device="cuda")
    g.manual_seed(2000 + cfg)
    rms = float(torch.sqrt(torch.mean(dobs.double() ** 2)))
    dobs += 0.01 * rms * torch.randn(dobs.shape, generator=g, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    return dict(w=w, vp=vp, c0=c0, h=h, dt=dt, src=src, rec=rec, dat=dat, s=s, s_dev=s_dev, dobs=dobs, dnew=dnew,
                nsh=nsh)


def run_redatum(daw, P, shots=None, max_batch=0, hbm_budget=0.0):
    import torch
    w = P["w"]
    nsh = P["nsh"] if shots is None else shots
    datum = torch.zeros((nsh, P["dat"].shape[1], w.nt), dtype=torch.float32, device="cuda")
    hist = daw.redatum_batch(P["c0"], w.npml, P["h"], w.ofs, w.nt, P["dt"], P["s_dev"], P["src"][:, :nsh],
                             P["rec"], P["dat"], P["dobs"][:nsh].contiguous(), w.mu, w.maxiter, 0.0,
                             datum_dev=datum, max_batch=max_batch, hbm_budget_bytes=hbm_budget, band=w.band)
    return datum, hist


def setup_uneven(daw, syn, cfg, shots=None, maxiter=None):
    """Configurations 3 and 5: shots on their own subdomains, receiver sets and
    nt.  A sample of `shots` shots spread over the line can stand in for all."""
    import torch
    w = syn.uneven_workload(cfg)
    if maxiter is not None:
        w.maxiter = maxiter
    idx = np.arange(w.nshots) if not shots or shots >= w.nshots else \
        np.round(np.linspace(0, w.nshots - 1, shots)).astype(int)
    vp, c0 = syn.uneven_models(w)
    h, dt = np.float32(w.h), np.float32(w.dt)
    x0, n2s, nts = w.x0[idx], w.n2s[idx], w.nts[idx]
    s = syn.ricker(int(nts.max()), float(dt), w.f0, float(h))
    s_dev = torch.from_numpy(s).cuda()
    src, recs, dats, dobs, dnew = [], [], [], [], []
    for k, i in enumerate(idx):
        rx = np.arange(n2s[k]) * w.rec_dx
        src.append(daw.subacqui([w.src_z], [w.src_x[i]], h, w.npml, w.ofs)[:, 0])
        recs.append(daw.subacqui(np.full(rx.size, w.rec_z), rx, h, w.npml, w.ofs))
        dx = rx[::4]
        dats.append(daw.subacqui(np.full(dx.size, w.dat_z), dx, h, w.npml, w.ofs))
        # observed data in the true model of this subdomain, and the data at the datum
        nt = int(nts[k])
        d = torch.zeros((1, recs[-1].shape[1], nt), dtype=torch.float32, device="cuda")
        dd = torch.zeros((1, dats[-1].shape[1], nt), dtype=torch.float32, device="cuda")
        daw.modeling_batch(np.ascontiguousarray(vp[x0[k]:x0[k] + n2s[k]]), w.npml, h, w.ofs, nt, dt, s_dev[:nt],
                           src[-1][:, None], recs[-1], d, dat_tab=dats[-1], datum_dev=dd)
        dobs.append(d.reshape(-1))
        dnew.append(dd.reshape(-1))
    dobs = torch.cat(dobs)
    g = torch.Generator(device="cuda")
    g.manual_seed(2000 + cfg)
    rms = float(torch.sqrt(torch.mean(dobs.double() ** 2)))
    dobs += 0.01 * rms * torch.randn(dobs.shape, generator=g, device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    return dict(w=w, c0=c0, h=h, dt=dt, x0=x0, n2s=n2s, nts=nts, src=np.stack(src, axis=1), recs=recs, dats=dats,
                s_dev=s_dev, dobs=dobs, dnew=torch.cat(dnew), nsh=len(idx))


def run_uneven(daw, P, max_batch=0, hbm_budget=0.0):
    import torch
    w = P["w"]
    nd = int(sum(d.shape[1] * int(nt) for d, nt in zip(P["dats"], P["nts"])))
    datum = torch.zeros(nd, dtype=torch.float32, device="cuda")
    hist = daw.redatum_uneven(P["c0"], w.npml, P["h"], w.ofs, P["dt"], P["s_dev"], P["x0"], P["n2s"], P["nts"],
                              P["src"], P["recs"], P["dobs"], w.mu, w.maxiter, 0.0, dat_tabs=P["dats"],
                              datum_dev=datum, max_batch=max_batch, hbm_budget_bytes=hbm_budget)
    return datum, hist


def main_uneven(args, daw, syn):
    """Timed redatuming of an uneven configuration (3 or 5)."""
    import torch
    _log(f"setup: uneven configuration {args.config}")
    P = setup_uneven(daw, syn, args.config, shots=args.shots, maxiter=args.maxiter)
    w = P["w"]
    daw.set_snapshot_format(args.snapshots if args.snapshots != "fp32" else w.snapshots)
    snaps = args.snapshots if args.snapshots != "fp32" else w.snapshots
    budget = args.hbm_budget_gb * 1e9
    daw.reset_stats()
    daw.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _log(f"timed step: {P['nsh']} shots")
    datum, hist = run_uneven(daw, P, max_batch=args.max_batch, hbm_budget=budget)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    daw.set_timing(False)
    st = daw.stats()
    n1e = w.nz + 2 * w.npml
    n2e = P["n2s"] + 2 * w.npml
    snap_b = 12.0 if snaps == "fp32" else 6.0 + 12.0 / 64.0
    # algorithmic bytes of the Hessian sweeps: per shot and step, state r+w and
    # kappa over the extended subdomain plus the snapshots of its interior
    per_step = 36.0 * n1e * n2e + snap_b * w.nz * P["n2s"]
    hess_bytes = float(np.sum(per_step * P["nts"])) * 2 * w.maxiter
    hess_ms = st["sweep_ms"][1] + st["sweep_ms"][2]
    achieved = hess_bytes / (hess_ms * 1e-3) / 1e9 if hess_ms > 0 else None
    err = float(torch.linalg.norm(datum - P["dnew"]) / torch.linalg.norm(P["dnew"]))
    out = {
        "metric": "shots redatumed/hour + stencil GPts/s on 1 MI355X, uneven batch",
        "value": P["nsh"] / elapsed * 3600.0,
        "unit": "shots/hour",
        "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": elapsed * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded faulted layered model, background sigma 10, Ricker, 1% seeded noise",
        "config": {"workload": w.name, "shots_sampled": P["nsh"], "maxiter": w.maxiter, "snapshots": snaps,
                   "subdomain_columns": [int(P["n2s"].min()), int(P["n2s"].max())],
                   "nt": [int(P["nts"].min()), int(P["nts"].max())],
                   "max_shots_per_launch": st["max_batch_used"], "tile": list(daw.get_tile()),
                   "steps_per_launch": daw.get_steps()},
        "gpts_per_s": st["point_updates"] / elapsed / 1e9,
        "redatum_rel_l2_vs_true_datum_data": err,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0 if achieved else None, "traffic": None,
                     "kernel": "k_step<1>/<2> (Hessian sweeps, average over the run)"},
        "cpu_baseline": None,
    }
    _emit(out)


def _cpu_worker(path):
    """One MPI rank's work in the reference's layout (para_range.f90:11-22:
    a shot per rank): the reference's own per-shot chain of mode 1 + mode 2
    (sub_redatum_DAW.f90:82-155, :230-240) -- submodeling, subresiduals,
    sublinearminres, submodeling_adjoint_esource, the wavefield sum,
    subwavefieldresample and filmartin2d -- on one shot, timed after the
    process-wide kwsinc2d set-up.  Runs in a child process that never touches
    the GPU; prints one JSON object."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    z = np.load(path, allow_pickle=False)
    c0, s, src, rec, dat, dobs = z["c0"], z["s"], z["src"], z["rec"], z["dat"], z["dobs"]
    npml, ofs, nt, maxiter = (int(z[k]) for k in ("npml", "ofs", "nt", "maxiter"))
    h, dt, mu = np.float32(z["h"]), np.float32(z["dt"]), np.float32(z["mu"])
    band = tuple(float(v) for v in z["band"])
    kind = "reference" if O.ref_available() else "port"
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)            # the reference's progress text (kwsinc.f90:224, minresModule.f90:512)
    try:
        if kind == "reference":
            O.ref_init(ofs)
            t0 = time.perf_counter()
            dr, wfr = O.ref_forward(c0, npml, h, ofs, nt, dt, s, src, rec)
            res = O.ref_residuals(dobs, dr)
            x = O.ref_linearminres(c0, npml, h, ofs, nt, dt, res, rec, mu, maxiter)
            _, wfe = O.ref_esource(c0, npml, h, ofs, nt, dt, x, rec)
            wfr += wfe
            O.ref_filmartin2d(O.ref_resample(wfr, npml, dat, ofs), dt, *band)
        else:
            t0 = time.perf_counter()
            dr, wfr = O.forward(c0, npml, h, ofs, nt, dt, s, src, rec, want_wavefield=True)
            res = O.residuals(dobs, dr)
            x, _, _ = O.linearminres(c0, npml, h, ofs, nt, dt, res, rec, mu, maxiter)
            _, wfe = O.esource(c0, npml, h, ofs, nt, dt, x, rec, want_wavefield=True)
            wfr += wfe
            O.filmartin2d(O.resample(wfr, npml, dat, ofs), dt, *band)
        t = time.perf_counter() - t0
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
    _emit({"seconds": t, "kind": kind})


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(P, nt_sample, cores=None, timeout=240):
    """The reference's CPU path on this host's cores, one shot per core (the
    para_range layout of sub_redatum_DAW.f90:30-34): `cores` child processes
    each run one shot's full mode-1 + mode-2 chain (7 + 2*itn sweeps) with the
    reference built from its own sources (oracle/_ref), truncated to the first
    nt_sample time steps of the shot's observed data, all at once.  Per-shot
    seconds scale linearly in nt to the full trace length (every sweep and
    every MINRES vector operation is O(nt)); aggregate shots/hour = cores *
    3600 / per-shot seconds."""
    import resource
    import subprocess
    import tempfile
    w = P["w"]
    if cores is None:
        try:
            avail = len(os.sched_getaffinity(0))
        except AttributeError:
            avail = os.cpu_count() or 1
        cores = max(1, min(16, avail))
    nsh = P["nsh"]
    idx = [int(round(i)) for i in np.linspace(0, nsh - 1, cores)]
    d = P["dobs"]
    dobs = d[:, :, :nt_sample].cpu().numpy() if hasattr(d, "cpu") else np.ascontiguousarray(d[:, :, :nt_sample])
    work = tempfile.mkdtemp(prefix="daw_cpu_", dir=os.environ.get("TMPDIR", "/tmp"))
    hard = resource.getrlimit(resource.RLIMIT_STACK)[1]

    def big_stack():     # MINRES keeps its work vectors on the stack (minresModule.f90:29)
        resource.setrlimit(resource.RLIMIT_STACK, (hard, hard))

    env = dict(os.environ, OMP_NUM_THREADS="1")
    procs = []
    try:
        for k, s in enumerate(idx):
            path = os.path.join(work, f"shot{k}.npz")
            np.savez(path, c0=P["c0"], s=P["s"][:nt_sample], src=np.ascontiguousarray(P["src"][:, s]),
                     rec=P["rec"], dat=P["dat"], dobs=dobs[s], npml=w.npml, ofs=w.ofs, nt=nt_sample,
                     maxiter=w.maxiter, h=P["h"], dt=P["dt"], mu=np.float32(w.mu), band=np.array(w.band))
            procs.append(subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", path],
                                          stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, env=env,
                                          preexec_fn=big_stack, text=True))
        t_wall0 = time.perf_counter()
        outs = []
        for pr in procs:
            out, _ = pr.communicate(timeout=max(1.0, timeout - (time.perf_counter() - t_wall0)))
            if pr.returncode != 0:
                raise RuntimeError(f"cpu worker exited with {pr.returncode}")
            outs.append(json.loads(out.strip().splitlines()[-1]))
        wall = time.perf_counter() - t_wall0
    finally:
        for pr in procs:
            if pr.poll() is None:
                pr.kill()
                pr.wait()
        import shutil
        shutil.rmtree(work, ignore_errors=True)
    secs = np.array([o["seconds"] for o in outs])
    scale = w.nt / nt_sample
    t_shot = float(np.mean(secs)) * scale
    n1e, n2e = w.nz + 2 * w.npml, w.nx + 2 * w.npml
    sweeps = 7 + 2 * w.maxiter
    return dict(value=round(cores * 3600.0 / t_shot, 2), unit="shots/hour", cores=cores, kind=outs[0]["kind"],
                cpu_model=_cpu_model(), per_shot_s=round(t_shot, 1),
                sample=f"{cores} processes at once, one shot each (shots {idx[0] + 1}..{idx[-1] + 1} spread over the "
                       f"line), the reference's full per-shot chain ({sweeps} sweeps, {w.maxiter} MINRES its, "
                       f"resample + filmartin2d) truncated to nt={nt_sample} of {w.nt}: "
                       f"{float(np.min(secs)):.1f}-{float(np.max(secs)):.1f} s per shot ({wall:.1f} s wall), "
                       f"scaled x{scale:g} to the full nt = {t_shot:.0f} s per shot",
                gpts_per_s=round(cores * n1e * n2e * sweeps * nt_sample / float(np.mean(secs)) / 1e9, 3))


def f64_check(P, datum, hist, nshots):
    """Per-shot relative L2 error of the redatumed gathers and of the MINRES
    residual history against the FP64 oracle of the plain method; the
    oracle's gathers go through the same band-pass (filmartin2d) as the
    timed path's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    w = P["w"]
    O.set_progress(True)
    dobs = P["dobs"].cpu().numpy()
    e_dat, e_hist = [], []
    for s in range(nshots):
        _log(f"fp64 oracle shot {s + 1}/{nshots}")
        row = P["src"][:, s].copy()
        dr, wfr = O.forward(P["c0"], w.npml, P["h"], w.ofs, w.nt, P["dt"], P["s"], row, P["rec"],
                            want_wavefield=True, prec="f64")
        res = np.ascontiguousarray((dobs[s].astype(np.float64) - dr)[:, ::-1])
        _log(f"fp64 oracle shot {s + 1}: MINRES")
        x, h, _ = O.linearminres(P["c0"], w.npml, P["h"], w.ofs, w.nt, P["dt"], res, P["rec"], w.mu, w.maxiter,
                                 prec="f64")
        _log(f"fp64 oracle shot {s + 1}: source extension")
        esc, wfe = O.esource(P["c0"], w.npml, P["h"], w.ofs, w.nt, P["dt"], x, P["rec"], want_wavefield=True,
                             prec="f64")
        ref = O.resample(wfr + wfe, w.npml, P["dat"], w.ofs, prec="f64")
        del wfr, wfe
        ref = O.filmartin2d(ref.astype(np.float32), float(P["dt"]), *w.band)
        got = datum[s].cpu().numpy()
        e_dat.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
        e_hist.append(float(np.max(np.abs(hist[s][:len(h)] / h - 1))))
    return {"f64_shots": nshots, "datum_rel_l2_vs_f64": e_dat, "minres_hist_max_rel_diff_vs_f64": e_hist}


def measured_traffic(key):
    """HBM bytes per launch of a kernel configuration, measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE runs (calibrated on a device copy) and recorded in
    profiles/hbm_traffic.json; None when this configuration was not measured."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(key, {}).get("bytes_per_launch")


def calibrate_copy(nbytes=1 << 30):
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(10):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 10
    _emit({"calibration": "copy", "bytes_read": nbytes, "bytes_written": nbytes, "launches": 13,
           "copy_GBps_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9})


_JSON_OUT = None


def main_workflow(args, syn):
    """The reference's workflow through dawredatum, files in and out: mode 0
    (observed data, batched), then modes 1+2 fused by --datum over all shots,
    and the plain route (mode 1 with oio = 1 writing the wavefields, then
    mode 2 reading them back) on a bounded sample of shots.  Wall-clock time
    of the CLI processes, including process start, file reads and writes."""
    import shutil
    import subprocess
    import tempfile
    cli = os.path.join(ROOT, "model-based-seismic-data-redatuming-with-daws_amd", "dawredatum")
    if not os.path.exists(cli):
        raise SystemExit("dawredatum is not built: run __graft_entry__.build()")
    w = syn.workload(args.config)
    if args.nt:
        w.nt = args.nt
    if args.maxiter is not None:
        w.maxiter = args.maxiter
    nsh = w.nshots if args.shots is None else args.shots
    kp = max(0, min(args.workflow_plain_shots, nsh))
    vp, c0 = syn.models(w)
    raw = syn.ricker(w.nt, w.dt, w.f0, 1.0)          # the wavelet file holds it before the division by h**2
    work = tempfile.mkdtemp(prefix="dawredatum_", dir=os.environ.get("TMPDIR", "/tmp"))

    def acq(name, ns, rz, rx):
        pts = [(w.src_z[i], w.src_x[i]) for i in range(ns)] + list(zip(rz, rx))
        with open(os.path.join(work, name), "w") as f:
            f.write(f"{ns} {len(rz)}\n" + "".join(f"{float(z)!r} {float(x)!r}\n" for z, x in pts))

    def run(mode, acq_name, oio, *extra):
        b = w.band
        with open(os.path.join(work, "redatum.par"), "w") as f:
            f.write(f"{mode}\nvtrue.bin vinit.bin\n{acq_name}\nwavelet.bin\n{b[0]} {b[1]} {b[2]} {b[3]}\n"
                    f"{w.nz} {w.nx} {w.npml} {w.h!r} {w.nt} {w.dt!r} {w.dt!r}\n0 0\n{w.ofs}\n{w.mu!r}\n{oio}\n"
                    f"0.0 {w.maxiter}\n")
        _log(f"dawredatum mode {mode} oio {oio} {' '.join(extra)}")
        t0 = time.perf_counter()
        r = subprocess.run([cli, "--quiet", *extra], cwd=work, capture_output=True, text=True)
        el = time.perf_counter() - t0
        if r.returncode != 0:
            raise SystemExit(f"dawredatum mode {mode} failed: {r.stderr[-2000:]}")
        _log(f"  {el:.2f} s")
        return el

    try:
        vp.tofile(os.path.join(work, "vtrue.bin"))
        c0.tofile(os.path.join(work, "vinit.bin"))
        raw.tofile(os.path.join(work, "wavelet.bin"))
        acq("acq_org.txt", nsh, w.rec_z, w.rec_x)
        acq("acq_new.txt", nsh, w.dat_z, w.dat_x)
        t_model = run(0, "acq_org.txt", 0)
        for i in range(1, nsh + 1):
            os.replace(os.path.join(work, f"dataobs_{i:05d}.bin"), os.path.join(work, f"dataobs_org_{i:05d}.bin"))
        # the observed data's first cpu_nt samples, for the CPU leg (dataobs(nt,nrec) per shot)
        nts = min(args.cpu_nt, w.nt)
        dobs_cpu = np.stack([np.fromfile(os.path.join(work, f"dataobs_org_{i:05d}.bin"), np.float32)
                             .reshape(len(w.rec_x), w.nt)[:, :nts] for i in range(1, nsh + 1)])
        t_fused = run(1, "acq_org.txt", 0, "--datum", "acq_new.txt", "--batch", str(args.workflow_batch))
        fused = {i: np.fromfile(os.path.join(work, f"dataobs_redatum_{i:05d}.bin"), np.float32) for i in range(1, kp + 1)}
        plain = None
        if kp > 0:
            acq("acq_org.txt", kp, w.rec_z, w.rec_x)
            acq("acq_new.txt", kp, w.dat_z, w.dat_x)
            t1 = run(1, "acq_org.txt", 1)
            wf_bytes = os.path.getsize(os.path.join(work, "wavefield_exd00001.bin"))
            t2 = run(2, "acq_new.txt", 1)
            diff = 0.0
            for i in range(1, kp + 1):
                a = np.fromfile(os.path.join(work, f"dataobs_redatum_{i:05d}.bin"), np.float32).astype(np.float64)
                b = fused[i].astype(np.float64)
                diff = max(diff, float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)))
            plain = {"shots": kp, "mode1_s": round(t1, 3), "mode2_s": round(t2, 3),
                     "shots_per_hour": round(3600.0 * kp / (t1 + t2), 1),
                     "wavefield_file_bytes_per_shot": 2 * wf_bytes,
                     "rel_l2_vs_fused": diff}
    finally:
        shutil.rmtree(work, ignore_errors=True)
    value = 3600.0 * nsh / t_fused
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        P = dict(w=w, c0=c0, h=np.float32(w.h), dt=np.float32(w.dt), nsh=nsh, dobs=dobs_cpu,
                 s=syn.ricker(w.nt, w.dt, w.f0, w.h),
                 src=O.acqui_table(w.src_z[:nsh], w.src_x[:nsh], w.h, w.npml, w.ofs),
                 rec=O.acqui_table(w.rec_z, w.rec_x, w.h, w.npml, w.ofs),
                 dat=O.acqui_table(w.dat_z, w.dat_x, w.h, w.npml, w.ofs))
        _log("reference CPU path on the host cores")
        cpu = cpu_baseline(P, nts, timeout=args.cpu_timeout)
    _emit({"metric": "redatumed shots/hour through dawredatum (modes 1+2 fused by --datum, files in and out)",
           "value": round(value, 1), "unit": "shots/h", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": round(1000.0 * t_fused, 1), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": w.name, "shots": nsh, "maxiter": w.maxiter, "batch": args.workflow_batch},
           "modeling_mode0_s": round(t_model, 3),
           "plain_sample": plain, "cpu_baseline": cpu})


def main_resample(args, daw):
    """dt -> dto resampling (subresample, subresample_hicks.f90:25-83) of the
    redatumed gathers of the 201x801 workload: 64 shots x 401 receivers,
    nt = 3000 at 0.8 ms to 4 ms, traces resident in HBM.  Algorithmic bytes:
    each input sample read once, each output sample written once."""
    import torch
    ntr, n1in, d1in, d1out = 64 * 401, 3000, 8e-4, 4e-3
    n1out = int(np.floor(float(np.float32(np.float32(n1in - 1) * np.float32(d1in)) / np.float32(d1out)) + 0.5)) + 1
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    xin = torch.randn((ntr, n1in), generator=g, device="cuda", dtype=torch.float32)
    xout = torch.empty((ntr, n1out), device="cuda", dtype=torch.float32)
    for _ in range(max(args.warmup, 2)):
        daw.subresample_dev(xin, xout, d1in, d1out)
    reps = max(args.steps, 20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        daw.subresample_dev(xin, xout, d1in, d1out)
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / reps
    nbytes = 4.0 * ntr * (n1in + n1out)
    # the reference's own subresample (oracle/_ref) on a bounded sample of the traces, 1 core
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    ns = 4000
    xs = xin[:ns].cpu().numpy()
    ys = np.zeros((ns, n1out), np.float32)
    R = O.ref() if O.ref_available() else None
    kind = "reference" if R is not None and hasattr(R, "dawref_subresample") else "port"
    t0 = time.perf_counter()
    if kind == "reference":
        R.dawref_subresample(xs, ys, n1in, ns, n1out, d1in, d1out)
    else:
        O.subresample(xs, d1in, d1out, n1out)
    tc = time.perf_counter() - t0
    cpu = {"value": round(ns / tc, 1), "unit": "traces/s", "cores": 1, "kind": kind,
           "sample": f"{ns} of the {ntr} traces in {tc:.3f} s"}
    _emit({"metric": "dt->dto resampling of redatumed gathers (subresample)", "value": round(ntr / per, 1),
           "unit": "traces/s", "n_gpus": 1, "steps": reps, "warmup": max(args.warmup, 2),
           "ms_per_step": round(1000 * per, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"{ntr} traces, {n1in} samples at {d1in} s -> {n1out} at {d1out} s"},
           "roofline_call": {"bound": "hbm", "achieved": round(nbytes / per / 1e9, 1), "peak": 8000.0,
                             "unit": "GB/s", "frac": round(nbytes / per / 8e12, 3), "traffic": None,
                             "note": "whole call (host plan and table upload, launch, sync); "
                                     "kernel alone in the rocprofv3 summary"},
           "cpu_baseline": cpu})


def _emit(obj):
    """The one JSON line, on the process's original stdout."""
    out = _JSON_OUT if _JSON_OUT is not None else sys.stdout
    out.write(json.dumps(obj) + "\n")
    out.flush()


def main():
    # stdout carries the JSON line only: text that native libraries write to
    # file descriptor 1 (the reference's Fortran runtime flushes its buffered
    # output when the process exits) goes to stderr instead
    global _JSON_OUT
    sys.stdout.flush()
    _JSON_OUT = os.fdopen(os.dup(1), "w")
    os.dup2(2, 1)
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="timed steps (batches of --batch shots)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=24, help="shots per step (advanced together)")
    ap.add_argument("--time-budget", type=float, default=540.0,
                    help="seconds after which the plain-path and CPU legs are left out")
    ap.add_argument("--cpu-timeout", type=float, default=180.0)
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--shots", type=int, default=None)
    ap.add_argument("--nt", type=int, default=None)
    ap.add_argument("--maxiter", type=int, default=None)
    ap.add_argument("--plain-shots", type=int, default=1)
    ap.add_argument("--cpu-nt", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--hbm-budget-gb", type=float, default=0.0)
    ap.add_argument("--max-batch", type=int, default=0)
    ap.add_argument("--tile-rows", type=int, default=0, help="tile rows, 0 for the autotuner")
    ap.add_argument("--tile-cols", type=int, default=64)
    ap.add_argument("--launch-steps", type=int, default=0, help="time steps per sweep launch: 1, 2, 0 for the autotuner")
    ap.add_argument("--snapshots", choices=["fp32", "bf16", "fp16"], default="fp32",
                    help="storage of the Hessian operator's wavefield snapshots")
    ap.add_argument("--graph", action="store_true", help="replay the sweeps as captured hipGraphs")
    ap.add_argument("--rec-batch", type=int, default=1, choices=[0, 1],
                    help="1: Hicks records as batched launches every few steps; 0: record blocks in the step launches")
    ap.add_argument("--f64-shots", type=int, default=0,
                    help="also run the FP64 oracle of the plain method on this many shots (slow, CPU)")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--workflow", action="store_true",
                    help="time the reference's file workflow through the dawredatum CLI instead")
    ap.add_argument("--workflow-plain-shots", type=int, default=1,
                    help="shots of the plain route (wavefield files) timed beside the fused one")
    ap.add_argument("--workflow-batch", type=int, default=64, help="dawredatum --batch")
    ap.add_argument("--resample", action="store_true",
                    help="time the dt -> dto time resampler on the workload's gathers instead")
    args = ap.parse_args()
    if args.cpu_worker:
        return _cpu_worker(args.cpu_worker)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); shots are independent and would "
              "shard across GPUs without any collective; this program drives one GPU.", file=sys.stderr)
        sys.exit(2)
    if args.workflow:
        return main_workflow(args, _pkg()[1])
    import torch
    if args.calibrate_copy:
        calibrate_copy()
        return
    daw, syn = _pkg()
    if args.resample:
        return main_resample(args, daw)
    daw.set_steps(args.launch_steps)
    daw.set_tile(args.tile_rows, args.tile_cols)
    daw.set_snapshot_format(args.snapshots)
    daw.set_graph(args.graph)
    daw.set_record_batch(args.rec_batch)
    if args.config in (3, 5):
        return main_uneven(args, daw, syn)
    t_start = T_PROC0
    _log("setup: modelling observed data")
    P = setup(daw, syn, args.config, shots=args.shots, nt=args.nt, maxiter=args.maxiter)
    w = P["w"]
    nsh = P["nsh"]
    budget = args.hbm_budget_gb * 1e9
    # one step = one redatuming pass over one batch of shots advanced together
    # (the launch batch: 16 shots at 201x801, whose wave state stays in the
    # last-level cache); steps walk the batches of the shot line in turn
    bsz = max(1, min(args.batch, nsh))
    nbat = (nsh + bsz - 1) // bsz
    batches = []
    for b in range(nbat):
        a, e = b * bsz, min(nsh, (b + 1) * bsz)
        batches.append(dict(src=np.ascontiguousarray(P["src"][:, a:e]), dobs=P["dobs"][a:e],
                            datum=torch.zeros((e - a, P["dat"].shape[1], w.nt), dtype=torch.float32, device="cuda"),
                            shots=(a, e)))

    def step(k):
        b = batches[k % nbat]
        return daw.redatum_batch(P["c0"], w.npml, P["h"], w.ofs, w.nt, P["dt"], P["s_dev"], b["src"], P["rec"],
                                 P["dat"], b["dobs"], w.mu, w.maxiter, 0.0, datum_dev=b["datum"], max_batch=bsz,
                                 hbm_budget_bytes=budget, band=w.band)

    # first-use warm-up of the library (kernel loading, stream creation) and
    # of the launch autotuner for every batch size of the line: short runs
    # (nt = 256, one MINRES iteration) of each size; the tuner's choice
    # depends on the batch's shots and grid, and its probes are 192 steps
    for ns in sorted({b["shots"][1] - b["shots"][0] for b in batches}):
        Q = setup(daw, syn, args.config, shots=ns, nt=256, maxiter=1)
        run_redatum(daw, Q, max_batch=ns)
    for i in range(args.warmup):
        _log(f"warmup step {i + 1}/{args.warmup}")
        step(i)
    daw.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hists = {}
    per_step = []        # (shots, sweep stats, launch configuration) of every timed step
    for i in range(args.steps):
        b = batches[i % nbat]
        _log(f"timed step {i + 1}/{args.steps}: shots {b['shots'][0] + 1}..{b['shots'][1]}")
        daw.reset_stats()
        ts = time.perf_counter()
        hists[i % nbat] = step(i)
        per_step.append((b["shots"][1] - b["shots"][0], daw.stats(), [daw.get_launch(m) for m in range(3)]))
        _log(f"  {time.perf_counter() - ts:.2f} s, of which sweeps {sum(per_step[-1][1]['sweep_ms']) / 1e3:.2f} s")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    daw.set_timing(False)
    shots_timed = sum(p[0] for p in per_step)
    shots_per_hour = shots_timed / elapsed * 3600.0
    gpts = sum(p[1]["point_updates"] for p in per_step) / elapsed / 1e9
    n1e, n2e = w.nz + 2 * w.npml, w.nx + 2 * w.npml
    # roofline of the full batches (bsz shots per launch): their sweep times,
    # from HIP events around the sweeps on the library's stream
    full = [p for p in per_step if p[0] == bsz] or per_step
    shots_per_launch = full[0][0]
    st = {"sweep_ms": [sum(p[1]["sweep_ms"][m] for p in full) for m in range(3)],
          "sweep_steps": [sum(p[1]["sweep_steps"][m] for p in full) for m in range(3)],
          "record_ms": [sum(p[1]["record_ms"][m] for p in full) for m in range(3)]}
    # the step kernels' own time: the sweeps' events less the batched record
    # launches they contain (their own events)
    step_ms = [st["sweep_ms"][m] - st["record_ms"][m] for m in range(3)]
    launch = {m: tuple(full[-1][2][i]) for i, m in enumerate(("forward", "back-propagation", "re-injection"))}
    # dominant kernel: that of the back-propagation or the re-injection sweep;
    # one launch advances all shots of a batch by launch_steps time steps
    dom = max((1, 2), key=lambda m: step_ms[m])
    dcfg = launch[("forward", "back-propagation", "re-injection")[dom]]
    launch_steps = dcfg[0]
    march = dcfg[2] in (52, 58)
    kname = ("k_march2_reinject" if dcfg[2] == 52 else "k_march" if march
             else "k_step2" if launch_steps == 2 else "k_step")
    modes = ({2: f"{kname} ({dcfg[1]}-row segments, two steps per launch)"} if dcfg[2] == 52 else
             {1: f"{kname}<1> (back-propagation)", 2: f"{kname}<2> (re-injection)"})
    avg_ms = step_ms[dom] / max(st["sweep_steps"][dom], 1) * launch_steps   # per launch
    # algorithmic bytes of one launch per shot: per point-update (the unit),
    # the wave state read and written (px, pz, vx, vz: 32 B) and kappa read
    # (4 B) per extended point, plus the step's snapshots of the interior (3
    # fields, 4 B each, or 2 B each + one 4 B scale per 64 columns), times the
    # steps the launch advances (a two-step launch moves the state through HBM
    # once: its measured traffic is lower than this)
    snap_b = 12.0 if args.snapshots == "fp32" else 6.0 + 12.0 / 64.0
    bytes_per_shot_launch = launch_steps * (36.0 * n1e * n2e + snap_b * w.nz * w.nx)
    alg_bytes = bytes_per_shot_launch * shots_per_launch
    achieved = alg_bytes / (avg_ms * 1e-3) / 1e9
    snap_code = {"fp32": 0, "bf16": 1, "fp16": 2}[args.snapshots]
    targs = (None if dcfg[2] == 52                                                      # k_march2_reinject
             else f"{dom},{dcfg[1]}" if march                                           # k_march<MODE, ROWS>
             else f"{dom},{snap_code},{dcfg[1]},{dcfg[2] // 64}" if launch_steps == 2     # k_step2<MODE, SNAP, TZ, NC>
             else f"{dom},{dcfg[1]},{dcfg[2] // 64},{snap_code}")                     # k_step<MODE, TZ, TXW, SNAP>
    traffic_key = (f"{kname}<{targs}>" if targs else f"{kname} rows {dcfg[1]}") + \
        f" {w.nz}x{w.nx} {int(shots_per_launch)} shots per launch"
    traffic = measured_traffic(traffic_key)
    datum = torch.cat([b["datum"] for b in batches])
    hist = np.concatenate([hists[b] for b in sorted(hists)]) if hists else None
    # quality: redatumed gathers vs data modelled at the datum in the true model (both band-passed)
    done = [b for b in range(nbat) if b in hists]
    dn = torch.cat([P["dnew"][batches[b]["shots"][0]:batches[b]["shots"][1]] for b in done])
    dg = torch.cat([batches[b]["datum"] for b in done])
    err = float(torch.linalg.norm(dg - dn) / torch.linalg.norm(dn))
    _log(f"timed: {shots_timed} shots in {elapsed:.1f} s = {shots_per_hour:.0f} shots/hour; "
         f"dominant {modes[dom]} {avg_ms * 1e3:.2f} us per launch = {achieved / 8000.0:.3f} of 8 TB/s")

    def time_left():
        return args.time_budget - (time.perf_counter() - t_start)

    # plain path: one shot at a time (the batch of shot 1, redatumed alone)
    plain = diff = None
    if args.plain_shots > 0 and time_left() > 60:
        b0 = batches[0]
        kp = min(args.plain_shots, b0["shots"][1] - b0["shots"][0])
        daw.reset_stats()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _log(f"plain path: {kp} shots one at a time")
        dp = torch.zeros((kp, P["dat"].shape[1], w.nt), dtype=torch.float32, device="cuda")
        daw.redatum_batch(P["c0"], w.npml, P["h"], w.ofs, w.nt, P["dt"], P["s_dev"],
                          np.ascontiguousarray(b0["src"][:, :kp]), P["rec"], P["dat"], b0["dobs"][:kp], w.mu,
                          w.maxiter, 0.0, datum_dev=dp, max_batch=1, band=w.band)
        torch.cuda.synchronize()
        tp = time.perf_counter() - t1
        plain = kp / tp * 3600.0
        ref_b = b0["datum"][:kp]
        diff = float(torch.max(torch.abs(dp - ref_b)) / torch.max(torch.abs(ref_b)))
    # mixed precision: the same batch with FP32 snapshots (untimed) and, on
    # request, the FP64 oracle of the plain method
    mixed = None
    if args.snapshots != "fp32" and time_left() > 60:
        daw.set_snapshot_format("fp32")
        _log("fp32-snapshot reference run of the first batch")
        b0 = batches[0]
        d32 = torch.zeros_like(b0["datum"])
        h32 = daw.redatum_batch(P["c0"], w.npml, P["h"], w.ofs, w.nt, P["dt"], P["s_dev"], b0["src"], P["rec"],
                                P["dat"], b0["dobs"], w.mu, w.maxiter, 0.0, datum_dev=d32, max_batch=bsz,
                                band=w.band)
        daw.set_snapshot_format(args.snapshots)
        per = [float(torch.linalg.norm(b0["datum"][s] - d32[s]) / torch.linalg.norm(d32[s]))
               for s in range(d32.shape[0])]
        mixed = {"snapshots": args.snapshots,
                 "datum_rel_l2_vs_fp32_max": max(per), "datum_rel_l2_vs_fp32_mean": float(np.mean(per)),
                 "minres_hist_max_rel_diff_vs_fp32": float(np.max(np.abs(hists[0] / h32 - 1)))}
    if args.f64_shots > 0:
        mixed = dict(mixed or {}, **f64_check(P, datum, hist, args.f64_shots))
    cpu = None
    if not args.no_cpu and time_left() > args.cpu_timeout + 20:
        _log(f"cpu baseline: reference path on the host cores, nt={args.cpu_nt}")
        try:
            cpu = cpu_baseline(P, args.cpu_nt, timeout=args.cpu_timeout)
        except Exception as exc:          # the GPU line stands without it
            _log(f"cpu baseline failed: {exc}")
            cpu = {"value": None, "error": str(exc)}
    out = {
        "metric": "shots redatumed/hour + stencil GPts/s on 1 MI355X, batched vs plain at equal error",
        "value": shots_per_hour,
        "unit": "shots/hour",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded layered model (sigma 3 smoothing), background sigma 10, Ricker, 1% seeded noise",
        "config": {"workload": f"{w.name}, {w.maxiter} MINRES its, datum z=400 m; one step = the redatuming of one "
                               f"batch of {bsz} shots advanced together (steps walk the {nbat} batches of the line)",
                   "shots": nsh, "shots_per_step": bsz, "shots_timed": shots_timed, "nt": w.nt,
                   "grid": [w.nz, w.nx], "h": w.h, "npml": w.npml, "receivers": int(P["rec"].shape[1]),
                   "shots_per_launch": shots_per_launch, "snapshots": args.snapshots, "graph": args.graph, "record_batch": bool(args.rec_batch),
                   "launch": {m: {"steps": v[0], "tile": [v[1], v[2]]} for m, v in launch.items()}},
        "gpts_per_s": gpts,
        "sweep_us_per_step": {k: st["sweep_ms"][i] * 1e3 / max(st["sweep_steps"][i], 1)
                              for i, k in enumerate(("forward", "back-propagation", "re-injection"))},
        "record_us_per_step": {k: st["record_ms"][i] * 1e3 / max(st["sweep_steps"][i], 1)
                               for i, k in enumerate(("forward", "back-propagation", "re-injection"))},
        "sweep_us_per_step_shots": shots_per_launch,
        "plain_shots_per_hour": plain,
        "batched_over_plain": (shots_per_hour / plain) if plain else None,
        "batched_vs_plain_max_rel_diff": diff,
        "redatum_rel_l2_vs_true_datum_data": err,
        "minres_final_rel_residual_mean": float(np.mean(hist[:, -1])) if hist is not None else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": traffic, "traffic_key": traffic_key,
                     "kernel": modes[dom],
                     "avg_launch_ms": avg_ms, "algorithmic_bytes_per_launch": alg_bytes,
                     "steps_per_launch": launch_steps,
                     "sweep_frac": alg_bytes / (st["sweep_ms"][dom] / max(st["sweep_steps"][dom], 1) * launch_steps
                                                * 1e-3) / 1e9 / 8000.0},
        "cpu_baseline": cpu,
        "wall_s": round(time.perf_counter() - t_start, 1),
    }
    if mixed:
        out["precision_check"] = mixed
    _emit(out)


if __name__ == "__main__":
    main()
