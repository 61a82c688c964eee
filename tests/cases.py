"""Small seeded problems shared by the parity tests."""
import numpy as np

import oracle as O


class Case:
    def __init__(self, nz=40, nx=90, npml=10, h=10.0, nt=300, dt=1e-3, f0=15.0, ofs=1, nrec=41, rec_z=15.0,
                 src_z=10.0, seed=3, dat_z=None):
        rng = np.random.Generator(np.random.PCG64(seed))
        self.n1, self.n2, self.npml, self.h, self.nt, self.dt, self.ofs = nz, nx, npml, np.float32(h), nt, \
            np.float32(dt), ofs
        z = np.arange(nz)[None, :]
        x = np.arange(nx)[:, None]
        vp = 1500.0 + 25.0 * z + 4.0 * x + 30.0 * rng.standard_normal((nx, nz))
        vp[:, :5] = 1500.0
        self.vp = vp.astype(np.float32)                      # (n2, n1)
        self.vp0 = (0.5 * (self.vp + self.vp.mean())).astype(np.float32)
        t = np.arange(nt) * dt - 1.5 / f0
        a = (np.pi * f0 * t) ** 2
        self.s = (((1 - 2 * a) * np.exp(-a)) / h ** 2).astype(np.float32)
        xmax = (nx - 1) * h
        self.src = O.acqui_table([src_z], [0.45 * xmax], h, npml, ofs)[:, 0].copy()
        self.rx = np.linspace(0.05, 0.95, nrec) * xmax
        self.rec = O.acqui_table(np.full(nrec, rec_z), self.rx, h, npml, ofs)
        dz = 0.5 * (nz - 1) * h if dat_z is None else dat_z
        self.dat = O.acqui_table(np.full(nrec, dz), self.rx, h, npml, ofs)
        self.rng = rng


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
