"""MI355X DAW redatuming: Python host mirror of the reference's operator API.

Each function below stands for the Fortran routine of the same name in the
reference (paths relative to its tree) and calls the HIP library libdaw.so
through its C-ABI (include/daw.h).  There is no CPU fallback: importing works
anywhere, but every compute call raises if libdaw.so is missing or no GPU is
visible.

Array conventions follow the reference's Fortran arrays (first index fastest)
seen from numpy in C order:
    vp(n1,n2)            -> ndarray (n2, n1)  float32
    sis(nt,nr)           -> ndarray (nr, nt)  float32
    wavefield(n1,n2,nt)  -> ndarray (nt, n2, n1)
    src_tab/rec_tab(n,5) -> ndarray (5, n)    int32
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdaw.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

_lib = None
_hicks_ofs = None

EXPORTS = [
    "daw_last_error", "daw_version", "daw_kwsinc2d", "daw_deallocatekws2d", "daw_hicks1d", "daw_hicks2d_tile",
    "daw_subsourcekws", "daw_subacqui", "daw_submodeling", "daw_submodeling_adjoint_esource",
    "daw_submodeling_adjoint_esource_data", "daw_subwavefieldresample", "daw_subresiduals",
    "daw_sublinearminres", "daw_filmartin2d", "daw_redatum_batch", "daw_set_timing", "daw_reset_stats",
    "daw_get_stats", "daw_modeling_batch", "daw_set_tile", "daw_get_tile", "daw_filmartin2d_dev",
    "daw_submodeling_transpose", "daw_set_hessian", "daw_set_snapshot_format",
    "daw_set_graph", "daw_set_steps", "daw_get_steps", "daw_redatum_uneven",
    "daw_get_launch", "daw_sublinearminres_log", "daw_write_minres_log", "daw_readpar", "daw_read_acqui",
    "daw_subresample", "daw_subresample_dev", "daw_set_minres_arith",
    "daw_subresample_dev_noalloc", "daw_set_krylov_format", "daw_set_record_batch",
]


class DawError(RuntimeError):
    pass


class RedatumArgs(C.Structure):
    _fields_ = [
        ("c0", C.c_void_p), ("n1", C.c_int), ("n2", C.c_int), ("npml", C.c_int),
        ("h", C.c_float), ("dt", C.c_float),
        ("ofs", C.c_int), ("nt", C.c_int), ("type_src", C.c_int),
        ("s_dev", C.c_void_p), ("nshots", C.c_int), ("nsrc_total", C.c_int), ("src_tab", C.c_void_p),
        ("nrec", C.c_int), ("rec_tab", C.c_void_p), ("ndat", C.c_int), ("dat_tab", C.c_void_p),
        ("dobs_dev", C.c_void_p), ("mu", C.c_float), ("maxiter", C.c_int), ("rtol", C.c_float),
        ("datum_dev", C.c_void_p), ("datae_dev", C.c_void_p), ("deltade_dev", C.c_void_p),
        ("hist_host", C.c_void_p), ("max_batch", C.c_int), ("hbm_budget_bytes", C.c_double),
        ("filter", C.c_int), ("flow", C.c_float), ("fhig", C.c_float), ("llow", C.c_float), ("lhig", C.c_float),
        ("minres_log_host", C.c_void_p), ("method", C.c_int),
    ]


class UnevenArgs(C.Structure):
    _fields_ = [
        ("c0", C.c_void_p), ("n1", C.c_int), ("n2", C.c_int), ("npml", C.c_int),
        ("h", C.c_float), ("dt", C.c_float), ("ofs", C.c_int), ("type_src", C.c_int),
        ("s_dev", C.c_void_p), ("nshots", C.c_int),
        ("x0", C.c_void_p), ("n2s", C.c_void_p), ("nts", C.c_void_p), ("src_tab", C.c_void_p),
        ("nrecs", C.c_void_p), ("rec_tabs", C.c_void_p), ("ndats", C.c_void_p), ("dat_tabs", C.c_void_p),
        ("dobs_dev", C.c_void_p), ("mu", C.c_float), ("maxiter", C.c_int), ("rtol", C.c_float),
        ("datum_dev", C.c_void_p), ("datae_dev", C.c_void_p), ("deltade_dev", C.c_void_p),
        ("hist_host", C.c_void_p), ("max_batch", C.c_int), ("hbm_budget_bytes", C.c_double),
        ("filter", C.c_int), ("flow", C.c_float), ("fhig", C.c_float), ("llow", C.c_float), ("lhig", C.c_float),
        ("minres_log_host", C.c_void_p), ("method", C.c_int),
    ]


class ModelingArgs(C.Structure):
    _fields_ = [
        ("vp", C.c_void_p), ("n1", C.c_int), ("n2", C.c_int), ("npml", C.c_int),
        ("h", C.c_float), ("dt", C.c_float),
        ("ofs", C.c_int), ("nt", C.c_int), ("type_src", C.c_int), ("type_rec", C.c_int),
        ("s_dev", C.c_void_p), ("nshots", C.c_int), ("nsrc_total", C.c_int), ("src_tab", C.c_void_p),
        ("nrec", C.c_int), ("rec_tab", C.c_void_p), ("ndat", C.c_int), ("dat_tab", C.c_void_p),
        ("dobs_dev", C.c_void_p), ("datum_dev", C.c_void_p), ("max_batch", C.c_int),
    ]


class Par(C.Structure):
    """daw_par: the fields of redatum.par (subreadpar.f90:9-21)."""
    _fields_ = [
        ("mode", C.c_int), ("name_true", C.c_char * 160), ("name_init", C.c_char * 160),
        ("name_acqui", C.c_char * 160), ("name_src", C.c_char * 160),
        ("flow", C.c_float), ("fhig", C.c_float), ("llow", C.c_float), ("lhig", C.c_float),
        ("n1", C.c_int), ("n2", C.c_int), ("npml", C.c_int), ("h", C.c_float), ("nt", C.c_int),
        ("dt", C.c_float), ("dto", C.c_float), ("type_src", C.c_int), ("type_rec", C.c_int), ("ofs", C.c_int),
        ("mu", C.c_float), ("oio", C.c_int), ("error", C.c_float), ("maxiter", C.c_float),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("sweep_ms", C.c_double * 3), ("sweep_steps", C.c_longlong * 3), ("batch_ms", C.c_double),
        ("point_updates", C.c_longlong), ("step_launches", C.c_longlong), ("max_batch_used", C.c_int),
        ("record_ms", C.c_double * 3),
    ]


def _gpu_runtime_first():
    """PyTorch ships its own HIP runtime; libdaw.so links the system one.  Both
    work in one process (the batched entry points take PyTorch's device
    buffers) when PyTorch's runtime has opened the GPU first; the other order
    leaves PyTorch without a device.  So open it before loading the library."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def lib():
    """Load libdaw.so (raises DawError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DawError("libdaw.so is not built: run __graft_entry__.build() or "
                           "python model-based-seismic-data-redatuming-with-daws_amd/build.py")
        _gpu_runtime_first()
        L = C.CDLL(LIB_PATH)
        L.daw_last_error.restype = C.c_char_p
        L.daw_hicks1d.argtypes = [_f32p]
        L.daw_hicks2d_tile.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f32p]
        L.daw_subsourcekws.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, _i32p, _i32p, _i32p]
        L.daw_subacqui.argtypes = [_f32p, _f32p, C.c_int, C.c_float, C.c_int, C.c_int, _i32p]
        L.daw_submodeling.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float,
                                      _f32p, C.c_int, C.c_int, _i32p, _i32p, C.c_int, _f32p, C.c_void_p]
        L.daw_submodeling_adjoint_esource.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                                      C.c_float, _f32p, C.c_int, C.c_int, _i32p, _f32p, C.c_void_p]
        L.daw_submodeling_adjoint_esource_data.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int,
                                                           C.c_int, C.c_float, _f32p, C.c_int, C.c_int, _i32p, _f32p]
        L.daw_subwavefieldresample.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, _i32p, C.c_int, _f32p,
                                               _f32p]
        L.daw_subresiduals.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int]
        L.daw_sublinearminres.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float,
                                          C.c_int, C.c_int, _i32p, C.c_float, C.c_float, C.c_float, _f32p, _f32p,
                                          _f64p, _f64p]
        L.daw_filmartin2d.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_float, _f32p]
        L.daw_redatum_batch.argtypes = [C.POINTER(RedatumArgs)]
        L.daw_redatum_uneven.argtypes = [C.POINTER(UnevenArgs)]
        L.daw_get_stats.argtypes = [C.POINTER(Stats)]
        L.daw_modeling_batch.argtypes = [C.POINTER(ModelingArgs)]
        L.daw_set_timing.argtypes = [C.c_int]
        L.daw_sublinearminres_log.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                              C.c_float, C.c_int, C.c_int, _i32p, C.c_float, C.c_float, C.c_float,
                                              _f32p, _f32p, C.c_char_p]
        L.daw_subresample.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]
        L.daw_subresample_dev.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]
        L.daw_subresample_dev_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_float,
                                                  C.c_float]
        L.daw_set_minres_arith.argtypes = [C.c_int]
        L.daw_set_krylov_format.argtypes = [C.c_int]
        L.daw_write_minres_log.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_float, _f64p]
        L.daw_readpar.argtypes = [C.c_char_p, C.POINTER(Par)]
        L.daw_read_acqui.argtypes = [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_void_p, C.c_int,
                                     C.c_void_p, C.c_int]
        L.daw_submodeling_transpose.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int,
                                                C.c_float, _f32p, _i32p, _i32p, C.c_int, _f32p]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise DawError(lib().daw_last_error().decode(errors="replace") or f"daw error {rc}")


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ------------------------------------------------------------------ kwsinc module (kwsinc.f90)

def kwsinc2d(ofs):
    """kwsinc2d(ifreesurf), kwsinc.f90:212 (after allocatekws2d :64)."""
    global _hicks_ofs
    _check(lib().daw_kwsinc2d(int(ofs)))
    _hicks_ofs = int(ofs)


def deallocatekws2d():
    """deallocatekws2d, kwsinc.f90:73."""
    global _hicks_ofs
    _check(lib().daw_deallocatekws2d())
    _hicks_ofs = None


def _ensure_hicks(ofs):
    if _hicks_ofs != ofs:
        kwsinc2d(ofs)


def hicks1d():
    """coef_hicks1D_monopole(-10:10,1:100) (kwsinc.f90:94-179) as (100, 21)."""
    out = np.zeros((100, 21), np.float32)
    _check(lib().daw_hicks1d(out))
    return out


def hicks2d_tile(ia1, ia2, ip, ofs):
    """coef_hicks2D_monopole(:,:,ia1,ia2,ip) (kwsinc.f90:262-364) as (21 [n2], 21 [n1])."""
    out = np.zeros((21, 21), np.float32)
    _check(lib().daw_hicks2d_tile(int(ia1), int(ia2), int(ip), int(ofs), out))
    return out


def subsourcekws(z, x, h, ofs):
    """subsourcekws (kwsinc.f90:582-606): (ix(2), ialpha(2), nalpha_frees)."""
    ix = np.zeros(2, np.int32)
    ia = np.zeros(2, np.int32)
    nf = np.zeros(1, np.int32)
    _check(lib().daw_subsourcekws(float(z), float(x), float(h), int(ofs), ix, ia, nf))
    return ix, ia, int(nf[0])


def subacqui(z, x, h, npml, ofs):
    """src_tab / rec_tab of subacqui.f90:12-62 for points (z[i], x[i]) -> (5, n) int32."""
    z = _f32(np.atleast_1d(z))
    x = _f32(np.atleast_1d(x))
    tab = np.zeros((5, z.size), np.int32)
    _check(lib().daw_subacqui(z, x, z.size, float(h), int(npml), int(ofs), tab))
    return tab


# ------------------------------------------------------------------ propagators (submodeling2.f90)

def submodeling(vp, npml, h, ofs, nt, dt, s, src_row, rec_tab, type_src=0, type_rec=0, want_wavefield=False):
    """submodeling (submodeling2.f90:11-288).  s is the wavelet already
    divided by h**2 (main_redatum_DAW.f90:111).  Returns sis (nr, nt) and,
    when asked, wavefield (nt, n2, n1)."""
    _ensure_hicks(ofs)
    vp = _f32(vp)
    n2, n1 = vp.shape
    rec_tab = _i32(rec_tab)
    nr = rec_tab.shape[1]
    sis = np.zeros((max(nr, 1), nt), np.float32)
    wf = np.zeros((nt, n2, n1), np.float32) if want_wavefield else None
    _check(lib().daw_submodeling(vp, n1, n2, int(npml), float(h), int(ofs), int(nt), float(dt), _f32(s),
                                 int(type_src), int(type_rec), _i32(src_row), rec_tab, nr, sis,
                                 None if wf is None else wf.ctypes.data))
    sis = sis[:nr]
    return (sis, wf) if want_wavefield else sis


def submodeling_transpose(vp, npml, h, ofs, nt, dt, d, src_row, rec_tab):
    """Exact discrete adjoint of submodeling with respect to the source
    wavelet: g (nt,) = F^T d for data d (nr, nt) (no reference counterpart)."""
    _ensure_hicks(ofs)
    vp = _f32(vp)
    n2, n1 = vp.shape
    rec_tab = _i32(rec_tab)
    g = np.zeros(nt, np.float32)
    _check(lib().daw_submodeling_transpose(vp, n1, n2, int(npml), float(h), int(ofs), int(nt), float(dt), _f32(d),
                                           _i32(src_row), rec_tab, rec_tab.shape[1], g))
    return g


def set_hessian(exact):
    """0: the reference's Hessian-type operator; 1: exact (symmetric) variant."""
    _check(lib().daw_set_hessian(1 if exact else 0))


def set_krylov_format(fmt):
    """Storage of the MINRES Krylov vectors r1, r2, w, w1, w2 in the FP64
    MINRES: "fp32" (default), "bf16" or "fp16" (per-256-element power-of-two
    scales; dot products and scalars stay FP64)."""
    if fmt not in SNAPSHOT_FORMATS:
        raise ValueError(f"Krylov vector format {fmt!r}: one of {sorted(SNAPSHOT_FORMATS)}")
    _check(lib().daw_set_krylov_format(SNAPSHOT_FORMATS[fmt]))


MINRES_ARITH = {"f64": 0, "ref32": 1}


def set_minres_arith(arith):
    """MINRES arithmetic of the solve: "f64" (default: FP64 dot products and
    Lanczos/Givens scalars over FP32 vectors) or "ref32" (the reference's
    own FP32 arithmetic: sequential FP32 dot products in its summation
    order, FP32 scalars and vector updates, minresModule.f90:10-360,
    sublinearminres.f90:64-72,117-133)."""
    if arith not in MINRES_ARITH:
        raise ValueError(f"minres arithmetic {arith!r}: one of {sorted(MINRES_ARITH)}")
    _check(lib().daw_set_minres_arith(MINRES_ARITH[arith]))


def set_steps(steps):
    """Time steps per launch of the sweep kernels: 1, 2 or 0 (autotuned)."""
    _check(lib().daw_set_steps(int(steps)))


def get_steps():
    return int(lib().daw_get_steps())


def get_launch(mode):
    """(steps per launch, tile rows, tile columns) the autotuner chose for
    sweep mode 0 (forward), 1 (back-propagation) or 2 (re-injection)."""
    st, r, c = C.c_int(), C.c_int(), C.c_int()
    _check(lib().daw_get_launch(int(mode), C.byref(st), C.byref(r), C.byref(c)))
    return st.value, r.value, c.value


def set_graph(on):
    """Replay full sweeps as captured hipGraphs (True) or launch step by step."""
    _check(lib().daw_set_graph(1 if on else 0))


SNAPSHOT_FORMATS = {"fp32": 0, "bf16": 1, "fp16": 2}


def set_snapshot_format(fmt):
    """Snapshot store of the Hessian-type operator between its passes:
    "fp32" (the reference's, default), "bf16" or "fp16" (16-bit values with a
    power-of-two scale per 64-column segment of each snapshot row)."""
    code = SNAPSHOT_FORMATS[fmt] if isinstance(fmt, str) else int(fmt)
    _check(lib().daw_set_snapshot_format(code))


def submodeling_adjoint_esource(vp, npml, h, ofs, nt, dt, res, rec_tab, type_rec=0, want_wavefield=False):
    """submodeling_adjoint_esource (submodeling2.f90:523-878)."""
    _ensure_hicks(ofs)
    vp = _f32(vp)
    n2, n1 = vp.shape
    rec_tab = _i32(rec_tab)
    nr = rec_tab.shape[1]
    sis = np.zeros((nr, nt), np.float32)
    wf = np.zeros((nt, n2, n1), np.float32) if want_wavefield else None
    _check(lib().daw_submodeling_adjoint_esource(vp, n1, n2, int(npml), float(h), int(ofs), int(nt), float(dt),
                                                 _f32(res), int(type_rec), nr, rec_tab, sis,
                                                 None if wf is None else wf.ctypes.data))
    return (sis, wf) if want_wavefield else sis


def submodeling_adjoint_esource_data(vp, npml, h, ofs, nt, dt, res, rec_tab, type_rec=0):
    """submodeling_adjoint_esource_data (submodeling2.f90:880-1217)."""
    _ensure_hicks(ofs)
    vp = _f32(vp)
    n2, n1 = vp.shape
    rec_tab = _i32(rec_tab)
    nr = rec_tab.shape[1]
    sis = np.zeros((nr, nt), np.float32)
    _check(lib().daw_submodeling_adjoint_esource_data(vp, n1, n2, int(npml), float(h), int(ofs), int(nt),
                                                      float(dt), _f32(res), int(type_rec), nr, rec_tab, sis))
    return sis


def subwavefieldresample(wavefield, npml, rec_tab, ofs, dt=0.0):
    """subwavefieldresample (subwavefieldresample.f90:1-41): wavefield (nt, n2, n1) -> sis (nr, nt)."""
    _ensure_hicks(ofs)
    wf = _f32(wavefield)
    nt, n2, n1 = wf.shape
    rec_tab = _i32(rec_tab)
    nr = rec_tab.shape[1]
    sis = np.zeros((nr, nt), np.float32)
    _check(lib().daw_subwavefieldresample(n1, n2, int(npml), nt, float(dt), rec_tab, nr, sis, wf))
    return sis


def subresiduals(dobs, dcal):
    """subresiduals (subresiduals.f90:6-18): time-reversed dobs - dcal."""
    dobs = _f32(dobs)
    dcal = _f32(dcal)
    nr, nt = dobs.shape
    res = np.zeros_like(dobs)
    _check(lib().daw_subresiduals(dobs, dcal, res, nt, nr))
    return res


def sublinearminres(c0, npml, h, ofs, nt, dt, deltadr, rec_tab, mu, maxiter, error=0.0, type_rec=0):
    """sublinearminres (sublinearminres.f90:7-86): returns (deltade (nr, nt),
    hist (qrnorm/beta1 per iteration), info dict)."""
    _ensure_hicks(ofs)
    c0 = _f32(c0)
    n2, n1 = c0.shape
    rec_tab = _i32(rec_tab)
    nr = rec_tab.shape[1]
    x = np.zeros((nr, nt), np.float32)
    hist = np.zeros(max(int(maxiter), 1), np.float64)
    info = np.zeros(9, np.float64)
    _check(lib().daw_sublinearminres(c0, n1, n2, int(npml), float(h), int(ofs), int(nt), float(dt), int(type_rec),
                                     nr, rec_tab, float(mu), float(maxiter), float(error), _f32(deltadr), x, hist,
                                     info))
    d = dict(fac=info[0], Anorm=info[1], Acond=info[2], rnorm=info[3], Arnorm=info[4], ynorm=info[5],
             beta1=info[6], istop=int(info[7]), itn=int(info[8]))
    return x, hist[:max(d["itn"], 0)], d


def sublinearminres_log(c0, npml, h, ofs, nt, dt, deltadr, rec_tab, mu, maxiter, error, log_path, type_rec=0):
    """sublinearminres with grid%oio = 1 (sublinearminres.f90:51-61): the
    solve, with MINRES's log written to log_path.  Returns deltade (nr, nt)."""
    _ensure_hicks(ofs)
    c0 = _f32(c0)
    n2, n1 = c0.shape
    rec_tab = _i32(rec_tab)
    nr = rec_tab.shape[1]
    x = np.zeros((nr, nt), np.float32)
    _check(lib().daw_sublinearminres_log(c0, n1, n2, int(npml), float(h), int(ofs), int(nt), float(dt),
                                         int(type_rec), nr, rec_tab, float(mu), float(maxiter), float(error),
                                         _f32(deltadr), x, os.fsencode(log_path)))
    return x


def write_minres_log(path, n, itnlim, rtol, block):
    """The MINRES log text of one solve from its log block (daw_write_minres_log)."""
    _check(lib().daw_write_minres_log(os.fsencode(path), int(n), int(itnlim), float(rtol),
                                      np.ascontiguousarray(block, dtype=np.float64)))


def readpar(path="redatum.par"):
    """subreadpar (subreadpar.f90:9-21): the parameter file as a dict."""
    p = Par()
    _check(lib().daw_readpar(os.fsencode(path), C.byref(p)))
    d = {k: getattr(p, k) for k, _ in Par._fields_}
    for k in ("name_true", "name_init", "name_acqui", "name_src"):
        d[k] = d[k].decode()
    return d


def read_acqui(path):
    """The acquisition file (main_redatum_DAW.f90:120-123, subacqui.f90:29-57):
    source and receiver coordinates as (n, 2) float32 arrays of (z, x)."""
    ns, nr = C.c_int(), C.c_int()
    _check(lib().daw_read_acqui(os.fsencode(path), C.byref(ns), C.byref(nr), None, 0, None, 0))
    src = np.zeros((ns.value, 2), np.float32)
    rec = np.zeros((nr.value, 2), np.float32)
    _check(lib().daw_read_acqui(os.fsencode(path), C.byref(ns), C.byref(nr), src.ctypes.data, ns.value,
                                rec.ctypes.data, nr.value))
    return src, rec


def subresample(xin, d1in, d1out, n1out=None):
    """subresample (subresample_hicks.f90:25-83): traces (n2, n1in) sampled at
    d1in resampled to (n2, n1out) at d1out.  n1out defaults to the count
    sub_redatum_DAW.f90:24 gives, nint((n1in-1)*d1in/d1out)+1."""
    xin = _f32(xin)
    n2, n1in = xin.shape
    if n1out is None:
        q = np.float32(np.float32(n1in - 1) * np.float32(d1in)) / np.float32(d1out)
        n1out = int(np.floor(float(q) + 0.5)) + 1     # NINT of the REAL quotient
    out = np.zeros((n2, int(n1out)), np.float32)
    _check(lib().daw_subresample(xin, out, n1in, n2, int(n1out), float(d1in), float(d1out)))
    return out


def subresample_dev(xin_dev, xout_dev, d1in, d1out):
    """subresample on contiguous float32 cuda tensors (n2, n1in) -> (n2, n1out)."""
    import torch
    for t in (xin_dev, xout_dev):
        if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.dim() != 2:
            raise ValueError("subresample_dev takes contiguous 2-D float32 cuda tensors")
    if xin_dev.shape[0] != xout_dev.shape[0]:
        raise ValueError("subresample_dev: trace counts differ")
    torch.cuda.synchronize()
    _check(lib().daw_subresample_dev(xin_dev.data_ptr(), xout_dev.data_ptr(), xin_dev.shape[1], xin_dev.shape[0],
                                     xout_dev.shape[1], float(d1in), float(d1out)))


def filmartin2d(din, dt, flow, fhig, llow, lhig):
    """filmartin2d (minresModule.f90:486-614) on traces (nr, nt)."""
    din = _f32(din)
    nr, nt = din.shape
    out = np.zeros_like(din)
    _check(lib().daw_filmartin2d(din, nr, nt, float(dt), float(flow), float(fhig), float(llow), float(lhig), out))
    return out


# ------------------------------------------------------------------ batched redatuming

def redatum_batch(c0, npml, h, ofs, nt, dt, s_dev, src_tab, rec_tab, dat_tab, dobs_dev, mu, maxiter, rtol=0.0,
                  datum_dev=None, datae_dev=None, deltade_dev=None, want_hist=True, max_batch=0,
                  hbm_budget_bytes=0.0, type_src=0, band=None, method=1, minres_log=None):
    """Mode-1 DAW plus the mode-2 datum gather of sub_redatum_DAW.f90:82-155,
    :238 for all shots of src_tab, advanced together on the GPU.

    s_dev, dobs_dev and the outputs are torch device tensors (float32):
    dobs_dev [nshots, nrec, nt]; datum_dev [nshots, ndat, nt] receives the
    redatumed gathers, band-passed as mode 2 does when band = (flow, fhig,
    llow, lhig) is given.  Returns the MINRES histories
    (nshots, maxiter) when want_hist.  method=3 runs the Symes variant of
    mode 3 (:157-228) instead of mode 1; minres_log, a float64 array of
    (nshots, 9 + 5*maxiter), receives the MINRES log blocks."""
    import torch
    _ensure_hicks(ofs)
    c0 = _f32(c0)
    n2, n1 = c0.shape
    src_tab = _i32(src_tab)
    rec_tab = _i32(rec_tab)
    dat_tab = _i32(dat_tab) if dat_tab is not None else None
    nshots = src_tab.shape[1]
    nrec = rec_tab.shape[1]
    ndat = 0 if dat_tab is None else dat_tab.shape[1]
    for t, shape in ((dobs_dev, (nshots, nrec, nt)), (datum_dev, (nshots, ndat, nt)),
                     (datae_dev, (nshots, nrec, nt)), (deltade_dev, (nshots, nrec, nt))):
        if t is not None:
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or tuple(t.shape) != shape:
                raise ValueError(f"device tensor must be contiguous float32 cuda of shape {shape}")
    if s_dev is not None and (s_dev.dtype != torch.float32 or not s_dev.is_cuda or s_dev.numel() != nt):
        raise ValueError("s_dev must be a float32 cuda tensor of nt samples")
    if minres_log is not None and (minres_log.dtype != np.float64 or not minres_log.flags.c_contiguous or
                                   minres_log.shape != (nshots, 9 + 5 * max(int(maxiter), 1))):
        raise ValueError("minres_log must be a C-contiguous float64 array of (nshots, 9 + 5*maxiter)")
    hist = np.zeros((nshots, max(int(maxiter), 1)), np.float64) if want_hist else None
    a = RedatumArgs()
    a.c0 = c0.ctypes.data
    a.n1, a.n2, a.npml = n1, n2, int(npml)
    a.h, a.dt = float(h), float(dt)
    a.ofs, a.nt, a.type_src = int(ofs), int(nt), int(type_src)
    a.s_dev = s_dev.data_ptr() if s_dev is not None else None
    a.nshots, a.nsrc_total, a.src_tab = nshots, nshots, src_tab.ctypes.data
    a.nrec, a.rec_tab = nrec, rec_tab.ctypes.data
    a.ndat, a.dat_tab = ndat, (dat_tab.ctypes.data if dat_tab is not None else None)
    a.dobs_dev = dobs_dev.data_ptr()
    a.mu, a.maxiter, a.rtol = float(mu), int(maxiter), float(rtol)
    a.datum_dev = datum_dev.data_ptr() if datum_dev is not None else None
    a.datae_dev = datae_dev.data_ptr() if datae_dev is not None else None
    a.deltade_dev = deltade_dev.data_ptr() if deltade_dev is not None else None
    a.hist_host = hist.ctypes.data if hist is not None else None
    a.max_batch = int(max_batch)
    a.hbm_budget_bytes = float(hbm_budget_bytes)
    if band is not None:
        a.filter = 1
        a.flow, a.fhig, a.llow, a.lhig = (float(v) for v in band)
    a.minres_log_host = minres_log.ctypes.data if minres_log is not None else None
    a.method = int(method)
    torch.cuda.synchronize()
    _check(lib().daw_redatum_batch(C.byref(a)))
    return hist


def redatum_uneven(c0, npml, h, ofs, dt, s_dev, x0, n2s, nts, src_tab, rec_tabs, dobs_dev, mu, maxiter, rtol=0.0,
                   dat_tabs=None, datum_dev=None, datae_dev=None, deltade_dev=None, want_hist=True, max_batch=0,
                   hbm_budget_bytes=0.0, type_src=0, band=None, method=1, minres_log=None):
    """Redatuming of shots on their own subdomains (columns x0[s] .. x0[s]+n2s[s]-1
    of c0), receivers and nt, advanced together.  src_tab (5, nshots) and the
    per-shot tables rec_tabs[s] (5, nrec_s) / dat_tabs[s] (5, ndat_s) are in
    subdomain coordinates (subacqui on the subdomain).  dobs_dev and the
    outputs are flat float32 cuda tensors holding shot s's [nrec_s, nt_s]
    (datum: [ndat_s, nt_s]) block after shot s-1's.  Returns the MINRES
    histories (nshots, maxiter) when want_hist.  method=3 runs the Symes
    variant (sub_redatum_DAW.f90:157-228); minres_log, float64 (nshots, 9 +
    5*maxiter), receives the MINRES log blocks."""
    import torch
    _ensure_hicks(ofs)
    c0 = _f32(c0)
    n2, n1 = c0.shape
    ns = len(x0)
    x0 = _i32(x0)
    n2s = _i32(n2s)
    nts = _i32(nts)
    src_tab = _i32(src_tab)
    rec_tabs = [_i32(t) for t in rec_tabs]
    nrecs = _i32([t.shape[1] for t in rec_tabs])
    rec_all = _i32(np.concatenate([t.ravel() for t in rec_tabs]))
    if dat_tabs is not None:
        dat_tabs = [_i32(t) for t in dat_tabs]
        ndats = _i32([t.shape[1] for t in dat_tabs])
        dat_all = _i32(np.concatenate([t.ravel() for t in dat_tabs]))
    nv = int(np.sum(nrecs.astype(np.int64) * nts))
    nd = int(np.sum(ndats.astype(np.int64) * nts)) if dat_tabs is not None else 0
    for t, n in ((dobs_dev, nv), (datum_dev, nd), (datae_dev, nv), (deltade_dev, nv)):
        if t is not None and (t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.numel() != n):
            raise ValueError(f"device tensor must be a contiguous float32 cuda tensor of {n} values")
    if s_dev.dtype != torch.float32 or not s_dev.is_cuda or s_dev.numel() < int(nts.max()):
        raise ValueError("s_dev must be a float32 cuda tensor of max(nts) samples")
    hist = np.zeros((ns, max(int(maxiter), 1)), np.float64) if want_hist else None
    a = UnevenArgs()
    a.c0 = c0.ctypes.data
    a.n1, a.n2, a.npml = n1, n2, int(npml)
    a.h, a.dt = float(h), float(dt)
    a.ofs, a.type_src = int(ofs), int(type_src)
    a.s_dev = s_dev.data_ptr()
    a.nshots = ns
    a.x0, a.n2s, a.nts, a.src_tab = x0.ctypes.data, n2s.ctypes.data, nts.ctypes.data, src_tab.ctypes.data
    a.nrecs, a.rec_tabs = nrecs.ctypes.data, rec_all.ctypes.data
    if dat_tabs is not None:
        a.ndats, a.dat_tabs = ndats.ctypes.data, dat_all.ctypes.data
    a.dobs_dev = dobs_dev.data_ptr()
    a.mu, a.maxiter, a.rtol = float(mu), int(maxiter), float(rtol)
    a.datum_dev = datum_dev.data_ptr() if datum_dev is not None else None
    a.datae_dev = datae_dev.data_ptr() if datae_dev is not None else None
    a.deltade_dev = deltade_dev.data_ptr() if deltade_dev is not None else None
    a.hist_host = hist.ctypes.data if hist is not None else None
    a.max_batch = int(max_batch)
    a.hbm_budget_bytes = float(hbm_budget_bytes)
    if band is not None:
        a.filter = 1
        a.flow, a.fhig, a.llow, a.lhig = (float(b) for b in band)
    if minres_log is not None and (minres_log.dtype != np.float64 or not minres_log.flags.c_contiguous or
                                   minres_log.shape != (ns, 9 + 5 * max(int(maxiter), 1))):
        raise ValueError("minres_log must be a C-contiguous float64 array of (nshots, 9 + 5*maxiter)")
    a.minres_log_host = minres_log.ctypes.data if minres_log is not None else None
    a.method = int(method)
    _check(lib().daw_redatum_uneven(C.byref(a)))
    return hist


def modeling_batch(vp, npml, h, ofs, nt, dt, s_dev, src_tab, rec_tab, dobs_dev, dat_tab=None, datum_dev=None,
                   max_batch=0, type_src=0, type_rec=0):
    """Mode-0 forward modelling (sub_redatum_DAW.f90:55-65) of every shot in
    src_tab at once: dobs_dev [nshots, nrec, nt] and, with datum receivers,
    datum_dev [nshots, ndat, nt] (torch float32 device tensors)."""
    import torch
    _ensure_hicks(ofs)
    vp = _f32(vp)
    n2, n1 = vp.shape
    src_tab = _i32(src_tab)
    rec_tab = _i32(rec_tab)
    dat_tab = _i32(dat_tab) if dat_tab is not None else None
    nshots = src_tab.shape[1]
    nrec = rec_tab.shape[1]
    ndat = 0 if dat_tab is None else dat_tab.shape[1]
    for t, shape in ((dobs_dev, (nshots, nrec, nt)), (datum_dev, (nshots, ndat, nt))):
        if t is not None and (t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous()
                              or tuple(t.shape) != shape):
            raise ValueError(f"device tensor must be contiguous float32 cuda of shape {shape}")
    a = ModelingArgs()
    a.vp = vp.ctypes.data
    a.n1, a.n2, a.npml = n1, n2, int(npml)
    a.h, a.dt = float(h), float(dt)
    a.ofs, a.nt, a.type_src, a.type_rec = int(ofs), int(nt), int(type_src), int(type_rec)
    a.s_dev = s_dev.data_ptr()
    a.nshots, a.nsrc_total, a.src_tab = nshots, nshots, src_tab.ctypes.data
    a.nrec, a.rec_tab = nrec, rec_tab.ctypes.data
    a.ndat, a.dat_tab = ndat, (dat_tab.ctypes.data if dat_tab is not None else None)
    a.dobs_dev = dobs_dev.data_ptr()
    a.datum_dev = datum_dev.data_ptr() if datum_dev is not None else None
    a.max_batch = int(max_batch)
    torch.cuda.synchronize()
    _check(lib().daw_modeling_batch(C.byref(a)))


def set_timing(on):
    _check(lib().daw_set_timing(1 if on else 0))


def filmartin2d_dev(din_dev, dout_dev, dt, flow, fhig, llow, lhig):
    """filmartin2d on torch device traces [nr, nt] (dout may be din)."""
    import torch
    nr, nt = din_dev.shape
    torch.cuda.synchronize()
    _check(lib().daw_filmartin2d_dev(C.c_void_p(din_dev.data_ptr()), nr, nt, C.c_float(dt), C.c_float(flow),
                                     C.c_float(fhig), C.c_float(llow), C.c_float(lhig),
                                     C.c_void_p(dout_dev.data_ptr())))


def set_tile(rows, cols=64):
    """Step-kernel tile rows x cols (8x128, 16x64, 16x128, 20x64, 24x64, 32x64,
    32x128, 64x64; rows x 58 for the column march, rows = 8, 12, 16, 24, 32;
    rows x 52 for the two-step march of the re-injection, rows = 16, 24, 32,
    40, 48), or rows = 0 for the autotuner (default)."""
    _check(lib().daw_set_tile(int(rows), int(cols)))


def set_record_batch(on):
    """True (default): Hicks records as batched launches every few steps;
    False: record blocks inside the step launches.  Same results."""
    _check(lib().daw_set_record_batch(1 if on else 0))


def get_tile():
    r, c = C.c_int(), C.c_int()
    _check(lib().daw_get_tile(C.byref(r), C.byref(c)))
    return r.value, c.value


def reset_stats():
    _check(lib().daw_reset_stats())


def stats():
    st = Stats()
    _check(lib().daw_get_stats(C.byref(st)))
    return dict(sweep_ms=list(st.sweep_ms), sweep_steps=list(st.sweep_steps), batch_ms=st.batch_ms,
                point_updates=st.point_updates, step_launches=st.step_launches, max_batch_used=st.max_batch_used,
                record_ms=list(st.record_ms))
