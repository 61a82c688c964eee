"""The drop-in, linked against the reference's own driver (CPU only, no GPU
work): the unmodified main_redatum_DAW.f90, sub_redatum_DAW.f90, subacqui.f90,
subreadpar.f90, subresiduals.f90, subname.f90 and para_range.f90 of the
reference tree are compiled with flang in a temporary directory, against the
modules kwsinc and minresModule of fortran/daw_shim.f90 and, for
sublinearminres, fortran/daw_shim_minres.f90 compiled against the driver's own
`parameters` module -- the object list of the reference's src/Makefile:24-25
with its hot-path objects (kwsinc.o, submodeling2.o, minresModule.o,
sublinearminres.o, subwavefieldresample.o, libfourier.o) replaced by the shim
and -ldaw.  MPI is a serial stub of our own: a `module mpi` that includes the
reference's vendored mpif.h, and the six routines the driver calls.

The test asserts that the executable links, that every hot-path routine the
driver calls resolves to the shim, and that the shim's calls resolve to
libdaw.so.  Nothing built here travels anywhere; the test is skipped where the
reference tree or flang is absent.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "model-based-seismic-data-redatuming-with-daws_amd")
REF_SRC = os.path.join(os.environ.get("DAW_REFERENCE_ROOT", "/root/reference"), "src")
FC = os.environ.get("FC", "flang")

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_SRC) or shutil.which(FC) is None
                                or not os.path.exists(os.path.join(PKG, "libdaw.so")),
                                reason="needs the reference tree, flang and a built libdaw.so")

DRIVER = ["main_redatum_DAW.f90", "sub_redatum_DAW.f90", "subacqui.f90", "subreadpar.f90", "subresiduals.f90",
          "subname.f90", "para_range.f90"]
HOT_PATH = ["kwsinc.f90", "submodeling2.f90", "minresModule.f90", "sublinearminres.f90", "subwavefieldresample.f90",
            "libfourier.f90"]

MPI_MODULE = """module mpi
  include 'mpif.h'
end module mpi
"""

MPI_STUBS = r"""
#include <time.h>
void mpi_init_(int *ierr) { *ierr = 0; }
void mpi_finalize_(int *ierr) { *ierr = 0; }
void mpi_comm_size_(int *comm, int *size, int *ierr) { (void)comm; *size = 1; *ierr = 0; }
void mpi_comm_rank_(int *comm, int *rank, int *ierr) { (void)comm; *rank = 0; *ierr = 0; }
void mpi_bcast_(void *buf, int *count, int *type, int *root, int *comm, int *ierr)
{
    (void)buf; (void)count; (void)type; (void)root; (void)comm;
    *ierr = 0;
}
double mpi_wtime_(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}
"""


def _run(cmd, cwd):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    assert r.returncode == 0, f"{' '.join(cmd)}\n{r.stdout[-3000:]}\n{r.stderr[-3000:]}"
    return r.stdout


def _symbols(path, cwd, flags):
    out = _run(["nm", *flags, path], cwd)
    return {line.split()[-1]: line.split()[-2] for line in out.splitlines() if len(line.split()) >= 2}


def test_reference_driver_links_against_the_shim(tmp_path):
    d = str(tmp_path)
    ff = ["-O2", "-cpp", "-w", "-I", d, "-I", REF_SRC, "-module-dir", d]
    with open(os.path.join(d, "mpi_module.f90"), "w") as f:
        f.write(MPI_MODULE)
    with open(os.path.join(d, "mpi_stubs.c"), "w") as f:
        f.write(MPI_STUBS)
    objs = []

    def fc(src, name):
        obj = os.path.join(d, name)
        _run([FC, *ff, "-c", src, "-o", obj], d)
        objs.append(obj)

    fc(os.path.join(d, "mpi_module.f90"), "mpi_module.o")
    fc(os.path.join(PKG, "fortran", "daw_shim.f90"), "daw_shim.o")            # modules kwsinc, minresmodule
    fc(os.path.join(REF_SRC, "main_redatum_DAW.f90"), "main_redatum_DAW.o")   # module parameters + the program
    fc(os.path.join(PKG, "fortran", "daw_shim_minres.f90"), "daw_shim_minres.o")
    for src in DRIVER[1:]:
        fc(os.path.join(REF_SRC, src), src.replace(".f90", ".o"))
    stub = os.path.join(d, "mpi_stubs.o")
    _run(["gcc", "-O2", "-c", os.path.join(d, "mpi_stubs.c"), "-o", stub], d)
    exe = os.path.join(d, "main_redatum_DAW")
    _run([FC, "-O2", *objs, stub, "-L", PKG, "-ldaw", "-Wl,-rpath," + PKG, "-o", exe], d)
    assert os.path.exists(exe)
    # no object of the reference's hot path took part
    for src in HOT_PATH:
        assert not os.path.exists(os.path.join(d, src.replace(".f90", ".o")))
    # the driver's hot-path calls resolve to the shim (defined in the executable) ...
    syms = _symbols(exe, d, [])
    for name in ("submodeling_", "submodeling_adjoint_esource_", "submodeling_adjoint_esource_data_",
                 "subwavefieldresample_", "sublinearminres_"):
        assert syms.get(name) == "T", (name, syms.get(name))
    # ... and the shim's to libdaw.so (undefined here, bound at run time)
    undef = _symbols(exe, d, ["-u"])
    for name in ("daw_submodeling", "daw_submodeling_adjoint_esource", "daw_submodeling_adjoint_esource_data",
                 "daw_subwavefieldresample", "daw_sublinearminres", "daw_filmartin2d", "daw_kwsinc2d"):
        assert name in undef, name
    ldd = _run(["ldd", exe], d)
    assert os.path.join(PKG, "libdaw.so") in ldd or "libdaw.so" in ldd
