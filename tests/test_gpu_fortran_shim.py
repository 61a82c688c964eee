"""The Fortran drop-in: a Fortran program calls submodeling,
submodeling_adjoint_esource[_data], subwavefieldresample, sublinearminres,
filmartin2d and the kwsinc module with the reference's argument lists
(tests/fortran/shim_driver.f90), linked against libdaw through the
iso_c_binding shim; its outputs are checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "shim_driver")


def test_fortran_driver_through_shim(gpu_daw, tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("tests/fortran/_build/shim_driver is not built (__graft_entry__.build())")
    c = Case(nt=200)
    nr, nd = c.rec.shape[1], c.dat.shape[1]
    res = c.rng.standard_normal((nr, c.nt)).astype(np.float32)
    mu = np.float32(1e-2)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([c.n1, c.n2, c.npml, c.nt, nr, nd, c.ofs], np.int32).tofile(f)
        np.array([c.h, c.dt, mu], np.float32).tofile(f)
        for a in (c.vp, c.vp0, c.s):
            np.ascontiguousarray(a, np.float32).tofile(f)
        for a in (c.src, c.rec, c.dat):
            np.ascontiguousarray(a, np.int32).tofile(f)
        res.tofile(f)
    print("running the Fortran driver", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(tmp_path / "out.bin", "rb").read()
    ints = np.frombuffer(raw[:20], np.int32)
    f = np.frombuffer(raw[20:], np.float32)
    nt, n1, n2 = c.nt, c.n1, c.n2
    d2 = np.float32(2.5) * c.dt
    ntout = int(np.floor(float(np.float32(np.float32(nt - 1) * c.dt) / d2) + 0.5)) + 1
    sizes = [nr * nt, n1 * n2 * nt, nr * nt, nr * nt, n1 * n2 * nt, nd * nt, nr * nt, nr * nt, nr * ntout]
    parts, o = [], 0
    for n in sizes:
        parts.append(f[o:o + n])
        o += n
    assert o == f.size
    sis, wf, sis2, sis3, wfp, sisd, filt, dde, rsmp = parts
    ix, ia, nf = O.subsourcekws(15.0, 123.4, c.h, c.ofs)
    assert list(ints) == list(ix) + list(ia) + [nf]
    osis, owf = O.forward(c.vp, c.npml, c.h, c.ofs, nt, c.dt, c.s, c.src, c.rec, want_wavefield=True)
    assert np.array_equal(sis.reshape(nr, nt), osis)
    assert np.array_equal(wf.reshape(nt, n2, n1), owf)
    assert np.array_equal(sis2.reshape(nr, nt), O.esource_data(c.vp0, c.npml, c.h, c.ofs, nt, c.dt, res, c.rec))
    esc, ewf = O.esource(c.vp0, c.npml, c.h, c.ofs, nt, c.dt, res, c.rec, want_wavefield=True)
    assert np.array_equal(sis3.reshape(nr, nt), esc)
    assert np.array_equal(wfp.reshape(nt, n2, n1), ewf)
    assert np.array_equal(sisd.reshape(nd, nt), O.resample(owf, c.npml, c.dat, c.ofs))
    assert rel_l2(filt.reshape(nr, nt), O.filmartin2d(osis, c.dt, 3.0, 30.0, 2.0, 10.0)) < 1e-6
    assert np.array_equal(rsmp.reshape(nr, ntout), O.subresample(osis, c.dt, d2, ntout))
    x64, _, info = O.linearminres(c.vp0, c.npml, c.h, c.ofs, nt, c.dt, res, c.rec, mu, 4, prec="f64")
    assert rel_l2(dde.reshape(nr, nt), x64) < 1e-4
    # grid%oio = 1: the MINRES log the reference's sublinearminres opens (:51-61)
    log = (tmp_path / "MINRES_00001.txt").read_text().splitlines()
    assert log[2].startswith(" Enter MINRES.") and f"{nr * nt:7d}" in log[3]
    assert [int(l.split()[0]) for l in log if len(l.split()) == 5 and l.split()[0].isdigit()] == [1, 2, 3, 4]
    assert f"istop ={info['istop']:3d}" in "\n".join(log)
