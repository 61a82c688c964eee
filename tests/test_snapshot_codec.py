"""Pin the oracle's 16-bit snapshot codec (CPU only).

The reference keeps the Hessian-type operator's snapshots in REAL*4
(submodeling2.f90:1060-1087); the BF16/FP16 store has no reference
counterpart, so its oracle is pinned against independent statements of the
same rounding: numpy's IEEE float16 conversion, and BF16 as
round-half-to-even of the significand to 8 bits via frexp/ldexp in float64.
"""
import numpy as np

import oracle as O
from cases import Case, rel_l2


def _bf16_math(x):
    x = np.asarray(x, np.float64)
    m, e = np.frexp(x)                                   # x = m 2^e, 0.5 <= |m| < 1
    return np.ldexp(np.round(np.ldexp(m, 8)), e - 8).astype(np.float32)   # np.round: ties to even


def _fp16_numpy(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def _values(rng):
    mag = 2.0 ** rng.uniform(-30, 0, 20000)
    v = (mag * rng.choice([-1.0, 1.0], mag.size)).astype(np.float32)
    v = np.concatenate([v, rng.uniform(-1, 1, 5000).astype(np.float32)])
    # exact ties between neighbouring FP16 / BF16 values, the FP16 subnormal
    # edge, signed zeros and the top of the range
    k = np.arange(1, 400, dtype=np.float64)
    ties = np.concatenate([(2 * k + 1) * 2.0 ** -25, (2 * k + 1) * 2.0 ** -12, (2 * k + 1) * 2.0 ** -10,
                           1 + 0 * k - 2.0 ** -12, (2 * k + 1) * 2.0 ** -20])
    ties = ties[np.abs(ties) <= 1.0]
    edge = np.array([2.0 ** -14, 2.0 ** -14 - 2.0 ** -25, 2.0 ** -14 - 2.0 ** -26, 2.0 ** -24, 2.0 ** -25,
                     3 * 2.0 ** -26, 2.0 ** -26, 0.0, -0.0, 1.0, 1.0 - 2.0 ** -24, -1.0])
    return np.concatenate([v, ties.astype(np.float32), -ties.astype(np.float32), edge.astype(np.float32)])


def test_fp16_round_matches_numpy():
    v = _values(np.random.default_rng(1))
    got = O.snap_round(v, "fp16")
    want = _fp16_numpy(v)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_bf16_round_matches_frexp_rounding():
    v = _values(np.random.default_rng(2))
    got = O.snap_round(v, "bf16")
    want = _bf16_math(v)
    nz = v != 0
    assert np.array_equal(got[nz], want[nz])
    assert np.array_equal(np.signbit(got[~nz]), np.signbit(v[~nz]))
    assert not np.any(got[~nz])


def test_scale_is_next_power_of_two():
    rng = np.random.default_rng(3)
    a = (2.0 ** rng.uniform(-90, 40, 2000)).astype(np.float32)
    a = np.concatenate([a, np.float32([1.0, 0.5, 2.0 ** -100, 3.0])])
    for x in a:
        s = O.snap_scale(x)
        m, _ = np.frexp(s)
        assert m == 0.5 and x < s <= 2 * x, (x, s)
    assert O.snap_scale(0.0) == 1.0
    assert O.snap_scale(2.0 ** -120) == 2.0 ** -100      # clamp keeps 1/scale finite


def _quantize_numpy(a, npml, snap):
    nt, n2, n1 = a.shape
    out = a.copy()
    seg = (npml + np.arange(n2)) // 64                     # extended column -> 64-column segment
    rnd = _fp16_numpy if snap == "fp16" else _bf16_math
    for k in np.unique(seg):
        cols = np.nonzero(seg == k)[0]
        blk = a[:, cols, :]                                 # (nt, ncols, n1)
        amax = np.abs(blk).max(axis=1)                      # per time step and row
        e = np.frexp(amax.astype(np.float64))[1]
        sc = np.where(amax > 0, np.ldexp(1.0, np.maximum(e, -100)), 1.0).astype(np.float32)
        q = rnd((blk * (np.float32(1) / sc)[:, None, :]).astype(np.float32))
        out[:, cols, :] = q * sc[:, None, :]
    return out


def test_quantize_segments_match_restatement():
    rng = np.random.default_rng(4)
    nt, n2, n1, npml = 5, 150, 7, 10
    a = rng.standard_normal((nt, n2, n1)).astype(np.float32)
    a *= (10.0 ** rng.uniform(-6, 3, (nt, 1, n1))).astype(np.float32)   # row-dependent magnitudes
    a[:, 100:, 2] = 0.0                                                  # an all-zero segment
    for snap in ("bf16", "fp16"):
        got = O.snap_quantize(a, npml, snap)
        assert np.array_equal(got, _quantize_numpy(a, npml, snap)), snap
        assert 0 < rel_l2(got, a) < (4e-3 if snap == "bf16" else 5e-4)


def test_esource_data_16bit_snapshots():
    """The operator with 16-bit snapshots stays within the rounding of the
    stored values of the FP32 and FP64 operators; FP16 (11 significant
    bits) is closer than BF16 (8)."""
    c = Case(nt=200)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    ref64 = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, prec="f64")
    ref32 = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    e = {}
    for snap in ("bf16", "fp16"):
        q = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, snap=snap)
        e[snap] = (rel_l2(q, ref32), rel_l2(q, ref64))
        print(snap, "rel vs f32 / f64", e[snap])
    assert 0 < e["fp16"][0] < e["bf16"][0] < 4e-3
    assert e["fp16"][1] < 5e-4
