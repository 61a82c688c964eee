"""Build libdaw.so (HIP kernels for gfx950 + C-ABI) in-tree with hipcc.

    python model-based-seismic-data-redatuming-with-daws_amd/build.py

The shared library lands next to this file; include/daw.h declares its
exports.  -ffp-contract=off keeps every single-precision expression rounded
as written (the reference's arithmetic); packed FP32 instructions are
disabled for the device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdaw.so")
ARCH = os.environ.get("DAW_OFFLOAD_ARCH", "gfx950")

SOURCES = ["daw_host.cpp", "daw_io.cpp", "daw_kernels.hip", "daw_engine.hip"]


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", "hipcc"):
        if os.path.exists(c) or c == "hipcc":
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    flags = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wno-unused-function",
             f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}"]
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(bdir, src + ".o")
        cmd = [hipcc()] + flags
        if src.endswith(".hip"):
            cmd += ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "-munsafe-fp-atomics"]
        cmd += ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}"] + objs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


CLI = os.path.join(HERE, "dawredatum")


def build_cli(verbose=False):
    """dawredatum (csrc/dawredatum.cpp): the reference driver's modes over
    libdaw.so, a host-only C++ program linked against the library and the HIP
    runtime (it only allocates and copies device buffers)."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    tmp = CLI + ".tmp"
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-ffp-contract=off", "-Wall", "-D__HIP_PLATFORM_AMD__",
           f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(rocm, 'include')}",
           os.path.join(CSRC, "dawredatum.cpp"), "-L", HERE, "-ldaw", "-L", os.path.join(rocm, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{os.path.join(rocm, 'lib')}", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, CLI)
    return CLI


def build_fortran(verbose=False):
    """The iso_c_binding drop-in (fortran/daw_shim.f90 -> fortran/_build/
    libdaw_fortran.a + .mod files) and the test driver that calls it the way
    the reference's driver calls the reference (tests/fortran/)."""
    fc = os.environ.get("FC", "flang")
    fdir = os.path.join(HERE, "fortran", "_build")
    tsrc = os.path.join(ROOT, "tests", "fortran")
    tdir = os.path.join(tsrc, "_build")
    os.makedirs(fdir, exist_ok=True)
    os.makedirs(tdir, exist_ok=True)

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    ff = ["-O2", "-fPIC"]
    run([fc] + ff + ["-module-dir", fdir, "-c", os.path.join(HERE, "fortran", "daw_shim.f90"),
                     "-o", os.path.join(fdir, "daw_shim.o")])
    run(["ar", "rcs", os.path.join(fdir, "libdaw_fortran.a"), os.path.join(fdir, "daw_shim.o")])
    run([fc] + ff + ["-module-dir", tdir, "-c", os.path.join(tsrc, "parameters_test.f90"),
                     "-o", os.path.join(tdir, "parameters_test.o")])
    run([fc] + ff + ["-I", fdir, "-I", tdir, "-module-dir", tdir, "-c",
                     os.path.join(HERE, "fortran", "daw_shim_minres.f90"), "-o", os.path.join(tdir, "daw_shim_minres.o")])
    rpath = "-Wl,-rpath,$ORIGIN/../../../" + os.path.basename(HERE)
    run([fc] + ff + ["-I", fdir, "-I", tdir, os.path.join(tsrc, "shim_driver.f90"),
                     os.path.join(tdir, "parameters_test.o"), os.path.join(tdir, "daw_shim_minres.o"),
                     os.path.join(fdir, "libdaw_fortran.a"), "-L", HERE, "-ldaw", rpath,
                     "-o", os.path.join(tdir, "shim_driver")])
    return os.path.join(tdir, "shim_driver")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cli(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
