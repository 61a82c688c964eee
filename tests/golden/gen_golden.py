"""Write the golden fixtures of tests/golden/ from the reference's own code.

Runs only where oracle/_ref/libdawref.so was built from the reference
sources (oracle/build_ref.sh).  Every fixture is data: inputs and the
reference's outputs on small seeded cases, stored in one .npz (no pickles).

    python tests/golden/gen_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

OUT = os.path.join(HERE, "ref_small.npz")


def tiny_case(n1=24, n2=44, npml=10, nt=120, nrec=13, seed=11):
    rng = np.random.Generator(np.random.PCG64(seed))
    h, dt = np.float32(10.0), np.float32(1e-3)
    vp = (1500.0 + 30.0 * np.arange(n1)[None, :] + 5.0 * np.arange(n2)[:, None]
          + 20.0 * rng.standard_normal((n2, n1))).astype(np.float32)
    t = np.arange(nt) * dt - 0.1
    s = (((1 - 2 * (np.pi * 15 * t) ** 2) * np.exp(-(np.pi * 15 * t) ** 2)) / h ** 2).astype(np.float32)
    return rng, vp, h, dt, npml, nt, s


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref/libdawref.so is not built; run oracle/build_ref.sh first")
    fx = {}
    fx["hicks1d"] = O.ref_hicks1d()
    rng = np.random.Generator(np.random.PCG64(5))
    for ofs in (1, 0):
        O.ref_init(ofs)
        idx = np.stack([rng.integers(1, 101, 40), rng.integers(1, 101, 40), rng.integers(0, 11, 40)], 1)
        idx[0] = (51, 51, 0)
        idx[1] = (1, 100, 10)
        fx[f"hicks2d_idx_ofs{ofs}"] = idx.astype(np.int32)
        fx[f"hicks2d_ofs{ofs}"] = np.stack([O.ref_hicks2d_tile(a, b, p, ofs) for a, b, p in idx])
    fx["bessi0_x"] = np.array([0.0, 0.5, 3.7, 3.75, 5.0, 10.0, 14.18], np.float32)
    fx["bessi0_y"] = np.array([O.ref_bessi0(v) for v in fx["bessi0_x"]], np.float32)

    # acquisition of the 121 x 401, h = 10 m workload: subacqui on a real file
    h, npml = np.float32(10.0), 20
    rx = (2000.0 + (np.arange(301) - 150) * 12.5).astype(np.float32)
    pts_z = np.concatenate([[10.0], np.full(301, 15.0), np.full(301, 400.0),
                            rng.uniform(0.0, 200.0, 200)]).astype(np.float32)
    pts_x = np.concatenate([[2000.0], rx, rx, rng.uniform(0.0, 4000.0, 200)]).astype(np.float32)
    with tempfile.TemporaryDirectory() as d:
        cwd = os.getcwd()
        os.chdir(d)
        try:
            with open("acq.txt", "w") as f:
                f.write(f"1 {pts_z.size - 1}\n")
                for z, x in zip(pts_z, pts_x):
                    f.write(f"{float(z)!r} {float(x)!r}\n")
            for ofs in (1, 0):
                st = np.zeros((5, 1), np.int32)
                rt = np.zeros((5, pts_z.size - 1), np.int32)
                name = b"acq.txt"
                O.ref().dawref_subacqui(name, len(name), h, npml, ofs, 1, pts_z.size - 1, st, rt)
                fx[f"acqui_tab_ofs{ofs}"] = np.concatenate([st, rt], 1)
        finally:
            os.chdir(cwd)
    fx["acqui_z"], fx["acqui_x"], fx["acqui_h"], fx["acqui_npml"] = pts_z, pts_x, np.float32(h), np.int32(npml)

    # tiny propagation cases
    rng, vp, h, dt, npml, nt, s = tiny_case()
    n2, n1 = vp.shape
    xmax = (n2 - 1) * float(h)
    src = O.acqui_table([10.0], [0.45 * xmax], h, npml, 1)[:, 0].copy()
    rxs = np.linspace(0.05, 0.95, 13) * xmax
    rec = O.acqui_table(np.full(13, 15.0), rxs, h, npml, 1)
    dat = O.acqui_table(np.full(13, 120.0), rxs, h, npml, 1)
    O.ref_init(1)
    fx.update(vp=vp, h=h, dt=dt, npml=np.int32(npml), nt=np.int32(nt), s=s, src=src, rec=rec, dat=dat)
    sis, wf = O.ref_forward(vp, npml, h, 1, nt, dt, s, src, rec)
    fx["fwd_sis"], fx["fwd_wf"] = sis, wf
    vp0 = (0.5 * (vp + vp.mean())).astype(np.float32)
    fx["vp0"] = vp0
    dcal, _ = O.ref_forward(vp0, npml, h, 1, nt, dt, s, src, rec)
    res = O.ref_residuals(sis, dcal)
    fx["res"] = res
    fx["esd_sis"] = O.ref_esource_data(vp0, npml, h, 1, nt, dt, res, rec)
    esc, wfe = O.ref_esource(vp0, npml, h, 1, nt, dt, res, rec)
    fx["es_sis"], fx["es_wf_slice"] = esc, wfe[::10]
    fx["resample_sis"] = O.ref_resample(wf, npml, dat, 1)
    fx["lmr_deltade"] = O.ref_linearminres(vp0, npml, h, 1, nt, dt, res, rec, 1e-2, 5)
    rn = []
    for k in range(1, 6):
        _, info = O.ref_minres(vp0, npml, h, 1, nt, dt, res, rec, 1e-2, k)
        rn.append([info["rnorm"], info["fac"], info["Anorm"], info["Acond"], info["ynorm"]])
    fx["minres_scalars"] = np.array(rn, np.float32)
    d = rng.standard_normal((4, 300)).astype(np.float32)
    fx["filt_in"] = d
    fx["filt_out"] = O.ref_filmartin2d(d, 1e-3, 3.0, 30.0, 2.0, 10.0)
    fx["filt_out2"] = O.ref_filmartin2d(d, 8e-4, 2.0, 700.0, 1.0, 5.0)   # fhig above Nyquist
    np.savez_compressed(OUT, **fx)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
