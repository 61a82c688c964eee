"""bench.py's host-side parts on the CPU: the one-GPU guard and the CPU
baseline leg (the reference's per-shot chain in child processes, one shot
per process), on a small workload."""
import os
import subprocess
import sys
from types import SimpleNamespace

import numpy as np

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_cpu_baseline_leg_small():
    sys.path.insert(0, ROOT)
    import bench
    nz, nx, npml, nt, nsh = 30, 60, 10, 160, 3
    h, dt = np.float32(10.0), np.float32(1e-3)
    w = SimpleNamespace(nz=nz, nx=nx, npml=npml, nt=nt, maxiter=2, mu=1e-2, ofs=1, band=(2.0, 30.0, 1.0, 5.0))
    vp = (1500.0 + 20.0 * np.arange(nz)[None, :] + 3.0 * np.arange(nx)[:, None]).astype(np.float32)
    c0 = (0.5 * (vp + vp.mean())).astype(np.float32)
    t = np.arange(nt) * dt - 0.1
    s = (((1 - 2 * (np.pi * 15 * t) ** 2) * np.exp(-(np.pi * 15 * t) ** 2)) / h ** 2).astype(np.float32)
    src = O.acqui_table(np.full(nsh, 10.0), np.array([150.0, 300.0, 450.0]), h, npml, 1)
    rec = O.acqui_table(np.full(21, 15.0), np.linspace(50.0, 550.0, 21), h, npml, 1)
    dat = O.acqui_table(np.full(21, 150.0), np.linspace(50.0, 550.0, 21), h, npml, 1)
    dobs = np.stack([O.forward(vp, npml, h, 1, nt, dt, s, src[:, k].copy(), rec) for k in range(nsh)])
    P = dict(w=w, c0=c0, h=h, dt=dt, nsh=nsh, dobs=dobs, s=s, src=src, rec=rec, dat=dat)
    out = bench.cpu_baseline(P, 80, cores=2, timeout=240)
    assert out["cores"] == 2 and out["kind"] in ("reference", "port")
    assert out["value"] > 0 and out["per_shot_s"] > 0
    assert "nt=80 of 160" in out["sample"]
