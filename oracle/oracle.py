"""ctypes wrappers over the CPU oracle (oracle/liboracle.so) and, when it was
built, the reference's own path (oracle/_ref/libdawref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU baseline,
never as the product path.

All arrays follow the reference's Fortran layout (first index fastest): a
(n1,n2) model is a numpy array of shape (n2, n1) in C order, sis(nt,nr) is
shape (nr, nt), wavefield(n1,n2,nt) is shape (nt, n2, n1), index tables
(n,5) are shape (5, n) int32.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdawref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.oracle_bessi0.restype = C.c_float
        L.oracle_bessi0.argtypes = [C.c_float]
        L.oracle_hicks1d.argtypes = [_f32p]
        L.oracle_hicks2d_tile.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f32p]
        L.oracle_subsourcekws.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, _i32p, _i32p, _i32p]
        L.oracle_acqui_table.argtypes = [_f32p, _f32p, C.c_int, C.c_float, C.c_int, C.c_int, _i32p]
        L.oracle_residuals.argtypes = [_f32p, _f32p, _f32p, C.c_int, C.c_int]
        L.oracle_realft2s.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
        L.oracle_filmartin2d.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                         C.c_float, C.c_float, _f32p]
        L.oracle_filmartin2d.restype = C.c_int
        L.oracle_filmartin_design.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                              C.c_float, _f64p, _i32p]
        L.oracle_filmartin_design.restype = C.c_int
        L.oracle_set_progress.argtypes = [C.c_int]
        L.oracle_snap_scale.argtypes = [C.c_float]
        L.oracle_snap_scale.restype = C.c_float
        L.oracle_snap_round_array.argtypes = [_f32p, C.c_int, C.c_int]
        L.snap_quantize_f32.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        for sfx, rp in (("f32", _f32p), ("f64", _f64p)):
            f = getattr(L, "oracle_forward_" + sfx)
            f.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float,
                          _f32p, C.c_int, C.c_int, _i32p, _i32p, C.c_int, rp, C.c_void_p]
            f.restype = C.c_int
            f = getattr(L, "oracle_esource_data_" + sfx)
            f.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float,
                          rp, C.c_int, C.c_int, _i32p, rp]
            f.restype = C.c_int
            f = getattr(L, "oracle_esource_data_snap_" + sfx)
            f.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float,
                          rp, C.c_int, C.c_int, _i32p, rp, C.c_int]
            f.restype = C.c_int
            f = getattr(L, "oracle_esource_" + sfx)
            f.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float,
                          rp, C.c_int, C.c_int, _i32p, rp, C.c_void_p, C.c_void_p]
            f.restype = C.c_int
            f = getattr(L, "oracle_resample_" + sfx)
            f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, C.c_int, C.c_int, rp, rp]
            f.restype = C.c_int
            f = getattr(L, "oracle_esource_data_exact_" + sfx)
            f.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float, rp, C.c_int,
                          _i32p, rp]
            f.restype = C.c_int
            f = getattr(L, "oracle_forward_adjoint_" + sfx)
            f.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float, rp, _i32p,
                          _i32p, C.c_int, rp]
            f.restype = C.c_int
            f = getattr(L, "oracle_linearminres_" + sfx)
            f.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_int, C.c_float,
                          C.c_int, C.c_int, _i32p, C.c_float, C.c_int, C.c_float, rp, rp, rp, rp, _i32p]
            f.restype = C.c_int
            getattr(L, "oracle_set_minres_trace_" + sfx).argtypes = [C.c_void_p]
        L.oracle_subhicks1d.argtypes = [_f32p]
        L.oracle_subresample.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _rdt(prec):
    return np.float32 if prec == "f32" else np.float64


# ---------------------------------------------------------------- Hicks / geometry

def bessi0(x):
    return np.float32(lib().oracle_bessi0(float(x)))


def hicks1d():
    """coef_hicks1D_monopole(-10:10,1:100) as array (100, 21)."""
    out = np.zeros((100, 21), np.float32)
    lib().oracle_hicks1d(out)
    return out


def hicks2d_tile(ia1, ia2, ip, ofs):
    """coef_hicks2D_monopole(:,:,ia1,ia2,ip) as array (21 [n2], 21 [n1])."""
    out = np.zeros((21, 21), np.float32)
    lib().oracle_hicks2d_tile(int(ia1), int(ia2), int(ip), int(ofs), out)
    return out


def subsourcekws(z, x, h, ofs):
    ix = np.zeros(2, np.int32)
    ia = np.zeros(2, np.int32)
    nf = np.zeros(1, np.int32)
    lib().oracle_subsourcekws(float(z), float(x), float(h), int(ofs), ix, ia, nf)
    return ix, ia, int(nf[0])


def acqui_table(z, x, h, npml, ofs):
    z = _c(z, np.float32)
    x = _c(x, np.float32)
    tab = np.zeros((5, z.size), np.int32)
    lib().oracle_acqui_table(z, x, z.size, float(h), int(npml), int(ofs), tab)
    return tab


def residuals(dobs, dcal):
    dobs = _c(dobs, np.float32)
    dcal = _c(dcal, np.float32)
    nr, nt = dobs.shape
    res = np.zeros_like(dobs)
    lib().oracle_residuals(dobs, dcal, res, nt, nr)
    return res


# ---------------------------------------------------------------- propagators

def forward(vp, npml, h, ofs, nt, dt, s, src_row, rec_tab, type_src=0, type_rec=0, want_wavefield=False,
            prec="f32"):
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    sis = np.zeros((nr, nt), _rdt(prec))
    wf = np.zeros((nt, n2, n1), _rdt(prec)) if want_wavefield else None
    f = getattr(lib(), "oracle_forward_" + prec)
    rc = f(vp, n1, n2, npml, h, ofs, nt, dt, _c(s, np.float32), type_src, type_rec, _c(src_row, np.int32),
           rec_tab, nr, sis, None if wf is None else _vp(wf))
    if rc != 0:
        raise RuntimeError("oracle forward failed")
    return (sis, wf) if want_wavefield else sis


def forward_adjoint(vp, npml, h, ofs, nt, dt, d, src_row, rec_tab, prec="f64"):
    """Exact discrete adjoint of forward (type_src = type_rec = 0): g = F^T d,
    g has nt samples (the gradient with respect to the source wavelet)."""
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    g = np.zeros(nt, _rdt(prec))
    rc = getattr(lib(), "oracle_forward_adjoint_" + prec)(vp, n1, n2, npml, h, ofs, nt, dt, _c(d, _rdt(prec)),
                                                           _c(src_row, np.int32), rec_tab, nr, g)
    if rc != 0:
        raise RuntimeError("oracle forward_adjoint failed")
    return g


SNAP_FORMATS = {"fp32": 0, "bf16": 1, "fp16": 2}


def esource_data(vp, npml, h, ofs, nt, dt, res, rec_tab, type_rec=0, prec="f32", snap="fp32"):
    """submodeling_adjoint_esource_data; snap "bf16"/"fp16" keeps the
    snapshots between the passes in 16-bit with 64-column segment scales."""
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    res = _c(res, _rdt(prec))
    sis = np.zeros((nr, nt), _rdt(prec))
    fmt = SNAP_FORMATS[snap]
    if fmt:
        rc = getattr(lib(), "oracle_esource_data_snap_" + prec)(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr,
                                                                 rec_tab, sis, fmt)
    else:
        rc = getattr(lib(), "oracle_esource_data_" + prec)(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr,
                                                            rec_tab, sis)
    if rc != 0:
        raise RuntimeError("oracle esource_data failed (out of memory?)")
    return sis


def esource_data_exact(vp, npml, h, ofs, nt, dt, res, rec_tab, prec="f64"):
    """Hessian-type operator with pass 1 the exact transpose of pass 2."""
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    sis = np.zeros((nr, nt), _rdt(prec))
    rc = getattr(lib(), "oracle_esource_data_exact_" + prec)(vp, n1, n2, npml, h, ofs, nt, dt, _c(res, _rdt(prec)), nr,
                                                              rec_tab, sis)
    if rc != 0:
        raise RuntimeError("oracle esource_data_exact failed")
    return sis


def esource(vp, npml, h, ofs, nt, dt, res, rec_tab, type_rec=0, want_wavefield=False, prec="f32"):
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    res = _c(res, _rdt(prec))
    sis = np.zeros((nr, nt), _rdt(prec))
    wf = np.zeros((nt, n2, n1), _rdt(prec)) if want_wavefield else None
    cost = np.zeros(1, _rdt(prec))
    rc = getattr(lib(), "oracle_esource_" + prec)(vp, n1, n2, npml, h, ofs, nt, dt, res, type_rec, nr, rec_tab,
                                                   sis, None if wf is None else wf.ctypes.data, _vp(cost))
    if rc != 0:
        raise RuntimeError("oracle esource failed (out of memory?)")
    return (sis, wf) if want_wavefield else sis


def resample(wavefield, npml, rec_tab, ofs, prec="f32"):
    wf = _c(wavefield, _rdt(prec))
    nt, n2, n1 = wf.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    sis = np.zeros((nr, nt), _rdt(prec))
    getattr(lib(), "oracle_resample_" + prec)(n1, n2, npml, nt, rec_tab, nr, ofs, sis, wf)
    return sis


def linearminres(vp, npml, h, ofs, nt, dt, deltadr, rec_tab, mu, itnlim, rtol=0.0, type_rec=0, prec="f32",
                 want_trace=False):
    """sublinearminres: returns (deltade (nr,nt), hist (itn,), info dict), and
    with want_trace also the log trace (itn, 5): x(1), qrnorm/beta1, Anorm,
    Acond, printed (minresModule.f90:308-325)."""
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    b = _c(deltadr, _rdt(prec))
    x = np.zeros((nr, nt), _rdt(prec))
    hist = np.zeros(max(int(itnlim), 1), _rdt(prec))
    out = np.zeros(7, _rdt(prec))
    iout = np.zeros(2, np.int32)
    trace = np.zeros((max(int(itnlim), 1), 5), _rdt(prec)) if want_trace else None
    setter = getattr(lib(), "oracle_set_minres_trace_" + prec)
    setter(trace.ctypes.data if want_trace else None)
    try:
        rc = getattr(lib(), "oracle_linearminres_" + prec)(vp, n1, n2, npml, h, ofs, nt, dt, type_rec, nr, rec_tab,
                                                            mu, int(itnlim), rtol, b, x, hist, out, iout)
    finally:
        setter(None)
    if rc != 0:
        raise RuntimeError("oracle linearminres failed")
    info = dict(fac=out[0], Anorm=out[1], Acond=out[2], rnorm=out[3], Arnorm=out[4], ynorm=out[5],
                beta1=out[6], istop=int(iout[0]), itn=int(iout[1]))
    if want_trace:
        return x, hist[:info["itn"]], info, trace[:info["itn"]]
    return x, hist[:info["itn"]], info


def subhicks1d():
    """The resampler's table (subresample_hicks.f90:87-157, irad 4, nalpha 10) as (10, 9)."""
    out = np.zeros((10, 9), np.float32)
    lib().oracle_subhicks1d(out)
    return out


def subresample(xin, d1in, d1out, n1out):
    """subresample (subresample_hicks.f90:25-83) on traces (n2, n1in)."""
    xin = _c(xin, np.float32)
    n2, n1in = xin.shape
    out = np.zeros((n2, int(n1out)), np.float32)
    lib().oracle_subresample(xin, out, n1in, n2, int(n1out), float(d1in), float(d1out))
    return out


def set_progress(on):
    """Per-iteration progress lines on stderr from the long-running solvers."""
    lib().oracle_set_progress(1 if on else 0)


def snap_round(x, snap):
    """16-bit round trip (ties to even) of float32 values |x| <= 1."""
    x = np.array(x, np.float32).ravel()
    lib().oracle_snap_round_array(x, x.size, SNAP_FORMATS[snap])
    return x


def snap_quantize(a, npml, snap):
    """16-bit round trip of snapshots a[nt][n2][n1] (Fortran a(n1,n2,nt)) with
    one scale per time step, row and 64-column segment of the extended grid."""
    a = np.array(a, np.float32)
    nt, n2, n1 = a.shape
    lib().snap_quantize_f32(a, n1, n2, npml, nt, SNAP_FORMATS[snap])
    return a


def snap_scale(amax):
    return float(lib().oracle_snap_scale(C.c_float(amax)))


# ---------------------------------------------------------------- filter

def realft2s(dre, dim, isign):
    dre = np.array(dre, np.float64)
    dim = np.array(dim, np.float64)
    lib().oracle_realft2s(dre.size, int(isign), dre, dim)
    return dre, dim


def filmartin_design(nt, dt, flow, fhig, llow, lhig):
    dm = np.zeros(65540, np.float64)
    iout = np.zeros(5, np.int32)
    nto = lib().oracle_filmartin_design(nt, dt, flow, fhig, llow, lhig, dm, iout)
    if nto < 0:
        raise ValueError("nt exceeds the filter's FFT length list")
    return nto, dm[1:nto + 2].copy(), iout


def filmartin2d(din, dt, flow, fhig, llow, lhig):
    din = _c(din, np.float32)
    nr, nt = din.shape
    out = np.zeros_like(din)
    rc = lib().oracle_filmartin2d(din, nr, nt, dt, flow, fhig, llow, lhig, out)
    if rc != 0:
        raise ValueError("nt exceeds the filter's FFT length list")
    return out


# ---------------------------------------------------------------- reference build

def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own Fortran path (built from its sources by
    oracle/build_ref.sh), or None when it was not built."""
    global _ref
    if _ref is None and ref_available():
        L = C.CDLL(REF_PATH)
        L.dawref_kwsinc_init.argtypes = [C.c_int]
        L.dawref_hicks1d.argtypes = [_f32p]
        L.dawref_hicks2d.argtypes = [C.c_int, C.c_int, C.c_int, _f32p]
        L.dawref_bessi0.argtypes = [C.c_float, C.POINTER(C.c_float)]
        L.dawref_subsourcekws.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int, _i32p, _i32p, _i32p]
        L.dawref_subacqui.argtypes = [C.c_char_p, C.c_int, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int,
                                      _i32p, _i32p]
        L.dawref_linearminres.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_int,
                                          C.c_int, C.c_int, C.c_float, C.c_int, _f32p, _i32p, _f32p, _f32p]
        if hasattr(L, "dawref_linearminres_log"):
            L.dawref_linearminres_log.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_int,
                                                  C.c_int, C.c_int, C.c_float, C.c_int, _f32p, _i32p, _f32p, _f32p]
        L.dawref_minres.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, C.c_int,
                                    C.c_int, C.c_int, C.c_float, C.c_int, _f32p, _i32p, _f32p, _f32p,
                                    C.POINTER(C.c_float), C.POINTER(C.c_int), C.POINTER(C.c_int), _f32p]
        L.dawref_filmartin2d.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.c_float, _f32p]
        L.dawref_realft2s.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
        if hasattr(L, "dawref_subresample"):
            L.dawref_subhicks1d.argtypes = [_f32p]
            L.dawref_subresample.argtypes = [_f32p, _f32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]
        _ref = L
    return _ref


_ref_ofs = None


def ref_init(ofs):
    """kwsinc2d(ofs) in the reference (allocates its ~580 MB tables once)."""
    global _ref_ofs
    if _ref_ofs != ofs:
        ref().dawref_kwsinc_init(int(ofs))
        _ref_ofs = ofs


def _vp(a):
    return C.c_void_p(a.ctypes.data)


def _ip(v):
    return C.byref(C.c_int(int(v)))


def _fp(v):
    return C.byref(C.c_float(float(v)))


def ref_forward(vp, npml, h, ofs, nt, dt, s, src_row, rec_tab, type_src=0, type_rec=0):
    """submodeling_ (submodeling2.f90:11) -> (sis (nr,nt), wavefield (nt,n2,n1))."""
    ref_init(ofs)
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    sis = np.zeros((nr, nt), np.float32)
    wf = np.zeros((nt, n2, n1), np.float32)
    s = _c(s, np.float32)
    src = _c(src_row, np.int32)
    ref().submodeling_(_ip(1), _vp(vp), _ip(n1), _ip(n2), _ip(npml), _fp(h), _ip(ofs), _ip(nt), _fp(dt),
                       _vp(s), _ip(type_src), _ip(type_rec), _vp(src), _vp(rec_tab), _ip(nr),
                       _vp(sis), _vp(wf))
    return sis, wf


def ref_esource_data(vp, npml, h, ofs, nt, dt, res, rec_tab, type_rec=0):
    ref_init(ofs)
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    res = _c(res, np.float32)
    sis = np.zeros((nr, nt), np.float32)
    ref().submodeling_adjoint_esource_data_(_ip(1), _vp(vp), _ip(n1), _ip(n2), _ip(npml), _fp(h), _ip(ofs),
                                            _ip(nt), _fp(dt), _vp(res), _ip(type_rec), _ip(nr),
                                            _vp(rec_tab), _vp(sis))
    return sis


def ref_esource(vp, npml, h, ofs, nt, dt, res, rec_tab, type_rec=0):
    ref_init(ofs)
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    res = _c(res, np.float32)
    sis = np.zeros((nr, nt), np.float32)
    wf = np.zeros((nt, n2, n1), np.float32)
    ref().submodeling_adjoint_esource_(_ip(1), _vp(vp), _ip(n1), _ip(n2), _ip(npml), _fp(h), _ip(ofs),
                                       _ip(nt), _fp(dt), _vp(res), _ip(type_rec), _ip(nr),
                                       _vp(rec_tab), _vp(sis), _vp(wf))
    return sis, wf


def ref_resample(wavefield, npml, rec_tab, ofs, dt=1e-3):
    ref_init(ofs)
    wf = _c(wavefield, np.float32)
    nt, n2, n1 = wf.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    sis = np.zeros((nr, nt), np.float32)
    ref().subwavefieldresample_(_ip(n1), _ip(n2), _ip(npml), _ip(nt), _fp(dt), _vp(rec_tab), _ip(nr),
                                _vp(sis), _vp(wf))
    return sis


def ref_residuals(dobs, dcal):
    dobs = _c(dobs, np.float32)
    dcal = _c(dcal, np.float32)
    nr, nt = dobs.shape
    res = np.zeros_like(dobs)
    ref().subresiduals_(_ip(1), _vp(dobs), _vp(dcal), _vp(res), _ip(nt), _ip(nr))
    return res


def ref_linearminres(vp, npml, h, ofs, nt, dt, deltadr, rec_tab, mu, maxiter, type_rec=0):
    ref_init(ofs)
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    out = np.zeros((nr, nt), np.float32)
    ref().dawref_linearminres(n1, n2, npml, h, nt, dt, ofs, type_rec, nr, mu, int(maxiter), vp, rec_tab,
                              _c(deltadr, np.float32), out)
    return out


def ref_linearminres_log(vp, npml, h, ofs, nt, dt, deltadr, rec_tab, mu, maxiter, workdir, type_rec=0):
    """The reference's sublinearminres with grid%oio = 1 run in workdir:
    (deltade, text of the MINRES_00001.txt it writes there)."""
    ref_init(ofs)
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    out = np.zeros((nr, nt), np.float32)
    here = os.getcwd()
    os.chdir(workdir)
    try:
        ref().dawref_linearminres_log(n1, n2, npml, h, nt, dt, ofs, type_rec, nr, mu, int(maxiter), vp, rec_tab,
                                      _c(deltadr, np.float32), out)
    finally:
        os.chdir(here)
    with open(os.path.join(workdir, "MINRES_00001.txt")) as f:
        return out, f.read()


def ref_minres(vp, npml, h, ofs, nt, dt, deltadr, rec_tab, mu, itnlim, type_rec=0):
    """MINRES stopped at itnlim: (x (nr,nt), info dict with fac, rnorm, ...)."""
    ref_init(ofs)
    vp = _c(vp, np.float32)
    n2, n1 = vp.shape
    rec_tab = _c(rec_tab, np.int32)
    nr = rec_tab.shape[1]
    x = np.zeros((nr, nt), np.float32)
    fac = C.c_float()
    istop = C.c_int()
    itn = C.c_int()
    scal = np.zeros(5, np.float32)
    ref().dawref_minres(n1, n2, npml, h, nt, dt, ofs, type_rec, nr, mu, int(itnlim), vp, rec_tab,
                        _c(deltadr, np.float32), x, C.byref(fac), C.byref(istop), C.byref(itn), scal)
    info = dict(fac=np.float32(fac.value), istop=istop.value, itn=itn.value, Anorm=scal[0], Acond=scal[1],
                rnorm=scal[2], Arnorm=scal[3], ynorm=scal[4])
    return x, info


def ref_hicks1d():
    out = np.zeros((100, 21), np.float32)
    ref().dawref_hicks1d(out)
    return out


def ref_hicks2d_tile(ia1, ia2, ip, ofs):
    ref_init(ofs)
    out = np.zeros((21, 21), np.float32)
    ref().dawref_hicks2d(int(ia1), int(ia2), int(ip), out)
    return out


def ref_bessi0(x):
    y = C.c_float()
    ref().dawref_bessi0(float(x), C.byref(y))
    return np.float32(y.value)


def ref_subsourcekws(z, x, h, ofs):
    ix = np.zeros(2, np.int32)
    ia = np.zeros(2, np.int32)
    nf = np.zeros(1, np.int32)
    ref().dawref_subsourcekws(float(z), float(x), float(h), int(ofs), ix, ia, nf)
    return ix, ia, int(nf[0])


def ref_filmartin2d(din, dt, flow, fhig, llow, lhig):
    din = _c(din, np.float32)
    nr, nt = din.shape
    out = np.zeros_like(din)
    ref().dawref_filmartin2d(din, nr, nt, dt, flow, fhig, llow, lhig, out)
    return out


def ref_realft2s(dre, dim, isign):
    dre = np.array(dre, np.float64)
    dim = np.array(dim, np.float64)
    ref().dawref_realft2s(dre.size, int(isign), dre, dim)
    return dre, dim
