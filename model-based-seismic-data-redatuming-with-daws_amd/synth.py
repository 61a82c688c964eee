"""Seeded synthetic inputs for the DAW redatuming workloads.

The reference ships no generator (its demo calls external tools,
demo/Marmousi/input.sh:18-53), so the workloads are defined here:

* velocity: a 20-cell 1500 m/s water layer over 8 flat layers at
  linspace(1800, 4200, 8) m/s, each shifted by U(-200, 200), Gaussian
  smoothed with sigma = 3 cells; optional dipping faults and a Gaussian
  4500 m/s body for the faulted family;
* background: the true model smoothed with sigma = 10 cells;
* wavelet: Ricker (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - 1.5/f,
  handed to the propagators divided by h**2 (main_redatum_DAW.f90:111);
* noise: Gaussian with sigma = 0.01 * rms(dobs).

Model seed 1000 + config, noise seed 2000 + config (PCG64).
"""
from dataclasses import dataclass, field

import numpy as np


def _smooth(a, sigma):
    if sigma <= 0:
        return a
    try:
        from scipy.ndimage import gaussian_filter
        return gaussian_filter(a, sigma, mode="nearest")
    except ImportError:   # separable numpy fallback for the input generator only
        r = int(4 * sigma + 0.5)
        k = np.exp(-0.5 * (np.arange(-r, r + 1) / sigma) ** 2)
        k /= k.sum()
        out = a
        for ax in (0, 1):
            pad = [(0, 0), (0, 0)]
            pad[ax] = (r, r)
            b = np.pad(out, pad, mode="edge")
            out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="valid"), ax, b)
        return out


def layered_model(nz, nx, seed, water_cells=20, faults=0, body=False, h=10.0):
    """vp as (nx, nz) float32, i.e. the reference's vp(n1=nz, n2=nx)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    v = np.empty((nz, nx), np.float64)
    v[:water_cells] = 1500.0
    speeds = np.linspace(1800.0, 4200.0, 8) + rng.uniform(-200.0, 200.0, 8)
    edges = np.linspace(water_cells, nz, 9).round().astype(int)
    z = np.arange(nz)[:, None] * np.ones((1, nx))
    for k in range(8):
        v[edges[k]:edges[k + 1]] = speeds[k]
    if faults:
        xx = np.arange(nx)[None, :] * np.ones((nz, 1))
        for _ in range(faults):
            x0 = rng.uniform(0.2, 0.8) * nx
            dip = np.tan(np.deg2rad(rng.uniform(50.0, 80.0))) * rng.choice([-1.0, 1.0])
            throw = rng.uniform(20.0, 80.0) / h
            side = (xx - x0) * dip > (z - water_cells)
            shifted = np.clip(np.round(z - throw).astype(int), 0, nz - 1)
            v2 = np.take_along_axis(v, shifted, axis=0)
            mask = side & (z >= water_cells)
            v = np.where(mask, v2, v)
        v[:water_cells] = 1500.0
    if body:
        cz, cx = rng.uniform(0.5, 0.8) * nz, rng.uniform(0.3, 0.7) * nx
        rz, rx = 0.08 * nz, 0.06 * nx
        g = np.exp(-(((z - cz) / rz) ** 2 + ((np.arange(nx)[None, :] - cx) / rx) ** 2))
        v = v + g * (4500.0 - v) * (z >= water_cells)
    v = _smooth(v, 3.0)
    v[:water_cells] = 1500.0
    return np.ascontiguousarray(v.T, dtype=np.float32)


def background(vp, sigma=10.0):
    return np.ascontiguousarray(_smooth(vp.astype(np.float64), sigma), dtype=np.float32)


def ricker(nt, dt, f, h):
    t = np.arange(nt, dtype=np.float64) * dt - 1.5 / f
    a = (np.pi * f * t) ** 2
    w = ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)
    return (w / np.float32(h) ** 2).astype(np.float32)


def add_noise(d, seed, level=0.01):
    rng = np.random.Generator(np.random.PCG64(seed))
    rms = float(np.sqrt(np.mean(d.astype(np.float64) ** 2)))
    return (d + rng.standard_normal(d.shape) * level * rms).astype(np.float32)


@dataclass
class Workload:
    name: str
    nz: int
    nx: int
    h: float
    npml: int
    nt: int
    dt: float
    f0: float
    src_z: np.ndarray
    src_x: np.ndarray
    rec_z: np.ndarray
    rec_x: np.ndarray
    dat_z: np.ndarray
    dat_x: np.ndarray
    maxiter: int = 30
    mu: float = 1e-2
    ofs: int = 1
    seed: int = 0
    faults: int = 0
    body: bool = False
    band: tuple = field(default=(2.0, 25.0, 1.0, 5.0))   # flow fhig llow lhig (Hz)

    @property
    def nshots(self):
        return int(self.src_x.size)


def workload(cfg):
    """The configurations of the benchmark table (1-based like the table)."""
    if cfg == 1:
        rx = 2000.0 + (np.arange(301) - 150) * 12.5
        return Workload("C1 121x401 h=10 nt=2500 1 shot 301 rec", 121, 401, 10.0, 20, 2500, 1e-3, 8.0,
                        np.array([10.0]), np.array([2000.0]), np.full(301, 15.0), rx, np.full(301, 400.0), rx.copy(),
                        seed=1001, band=(2.0, 20.0, 1.0, 5.0))
    if cfg == 2:
        sx = 400.0 + 50.0 * np.arange(64)
        rx = 16.0 * np.arange(401)
        return Workload("C2 201x801 h=8 nt=3000 64 shots 401 rec", 201, 801, 8.0, 20, 3000, 8e-4, 10.0,
                        np.full(64, 8.0), sx, np.full(401, 12.0), rx, np.full(401, 400.0), rx.copy(),
                        seed=1002, band=(2.0, 25.0, 1.0, 5.0))
    raise ValueError(f"unknown workload {cfg}")


@dataclass
class UnevenWorkload:
    """Shots on their own subdomains (aperture), receiver sets and nt."""
    name: str
    nz: int
    nx: int
    h: float
    npml: int
    dt: float
    f0: float
    x0: np.ndarray          # first column of each shot's subdomain
    n2s: np.ndarray         # subdomain widths (columns)
    nts: np.ndarray         # time steps per shot
    src_x: np.ndarray       # source x in subdomain coordinates (m)
    src_z: float
    rec_z: float
    rec_dx: float           # receiver spacing (m): one per rec_dx over the subdomain
    dat_z: float
    maxiter: int = 30
    mu: float = 1e-2
    ofs: int = 1
    seed: int = 0
    faults: int = 0
    body: bool = False
    snapshots: str = "fp32"
    band: tuple = field(default=(2.0, 40.0, 1.0, 5.0))

    @property
    def nshots(self):
        return int(self.x0.size)


def uneven_workload(cfg):
    """Configurations with per-shot apertures, receiver counts and nt: 3 (351 x 1401,
    h = 6.25 m, 128 shots, 3-6 km apertures, nt 4000-6000, FP32) and 5 (801 x 3201,
    h = 3.125 m, 256 shots at 37.5 m, 400-1600 receivers, nt = 12000, 16-bit
    snapshots)."""
    rng = np.random.Generator(np.random.PCG64(3000 + cfg))
    if cfg == 3:
        nz, nx, h, ns = 351, 1401, 6.25, 128
        ap = rng.uniform(3000.0, 6000.0, ns)
        n2s = np.round(ap / h).astype(np.int32)                       # 480-960 columns, one receiver each
        sx = np.linspace(0.05, 0.95, ns) * (nx - 1) * h
        nts = rng.integers(4000, 6001, ns).astype(np.int32)
        name, dt, f0, dx, snaps = "C3 351x1401 h=6.25 128 shots uneven apertures 3-6 km nt 4000-6000", 5e-4, 20.0, h, "fp32"
        seed, faults, body = 1003, 3, True
    elif cfg == 5:
        nz, nx, h, ns = 801, 3201, 3.125, 256
        nrec = rng.integers(400, 1601, ns)
        n2s = nrec.astype(np.int32)                                  # one receiver per column
        sx = 400.0 + 37.5 * np.arange(ns)
        nts = np.full(ns, 12000, np.int32)
        name, dt, f0, dx, snaps = "C5 801x3201 h=3.125 256 shots 400-1600 rec nt=12000", 2.5e-4, 30.0, h, "bf16"
        seed, faults, body = 1005, 3, False
    else:
        raise ValueError(f"unknown uneven workload {cfg}")
    # subdomain centred on the source, clipped to the model
    x0 = np.clip(np.round(sx / h).astype(np.int64) - n2s // 2, 0, nx - n2s).astype(np.int32)
    src_local = sx - x0 * h
    return UnevenWorkload(name, nz, nx, h, 20, dt, f0, x0, n2s, nts, src_local, 10.0, 15.0, dx, 400.0,
                          seed=seed, faults=faults, body=body, snapshots=snaps)


def uneven_models(w):
    vp = layered_model(w.nz, w.nx, w.seed, water_cells=min(20, w.nz // 4), faults=w.faults, body=w.body, h=w.h)
    return vp, background(vp)


def small_workload(nz=40, nx=90, nt=300, nshots=2, nrec=41, h=10.0, dt=1e-3, f0=15.0, maxiter=4, seed=7):
    """A small instance of the same family for parity tests."""
    sx = np.linspace(0.3, 0.7, nshots) * (nx - 1) * h
    rx = np.linspace(0.05, 0.95, nrec) * (nx - 1) * h
    return Workload(f"small {nz}x{nx} nt={nt}", nz, nx, h, 10, nt, dt, f0, np.full(nshots, 10.0), sx,
                    np.full(nrec, 15.0), rx, np.full(nrec, 0.5 * (nz - 1) * h), rx.copy(), maxiter=maxiter,
                    seed=seed, band=(2.0, 30.0, 1.0, 5.0))


def models(w):
    vp = layered_model(w.nz, w.nx, w.seed, water_cells=min(20, w.nz // 4), faults=w.faults, body=w.body, h=w.h)
    return vp, background(vp)
