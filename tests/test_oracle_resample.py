"""The time resampler subresample (subresample_hicks.f90:25-83) in the oracle
(CPU).  Pinned bit for bit against the reference's own source built by
oracle/build_ref.sh where the reference tree is present, and independently of
it by what the mathematics fixes: the table is a Kaiser-windowed sinc that
sums to ~1 and is 1 at the node for zero shift, a band-limited signal is
interpolated to the window's accuracy, equal sampling is the identity to the
size of the sinc's float residue, and the operator is linear."""
import numpy as np
import pytest

import oracle as O

CASES = [
    # n1in, n2, d1in, d1out, n1out (None: nint((n1in-1)*d1in/d1out)+1)
    (3000, 5, 8e-4, 4e-3, None),        # dt -> dto decimation of the 201x801 workload
    (2500, 3, 1e-3, 2e-3, None),
    (1000, 2, 1e-3, 1e-3, None),        # equal sampling
    (500, 4, 4e-3, 1e-3, None),         # upsampling
    (777, 3, 1.3e-3, 2.9e-3, None),     # incommensurate steps
    (7, 2, 1e-3, 3e-4, 40),             # shorter than the window; output past the input's end
    (300, 1, 2e-3, 2e-3 / 3.0, 899),
]


def _n1out(n1in, d1in, d1out, n1out):
    if n1out is not None:
        return n1out
    q = np.float32(np.float32(n1in - 1) * np.float32(d1in)) / np.float32(d1out)
    return int(np.floor(float(q) + 0.5)) + 1


def _traces(n2, n1in, seed):
    return np.random.default_rng(seed).standard_normal((n2, n1in)).astype(np.float32)


@pytest.mark.skipif(not O.ref_available(), reason="reference tree not built into oracle/_ref")
def test_table_matches_reference_build():
    R = O.ref()
    if not hasattr(R, "dawref_subresample"):
        pytest.skip("oracle/_ref predates the resampler: rerun oracle/build_ref.sh")
    ref = np.zeros((10, 9), np.float32)
    R.dawref_subhicks1d(ref)
    assert np.array_equal(O.subhicks1d(), ref)


def _defined_outputs(n1in, n1out, d1in, d1out):
    """Output samples whose value the reference defines.  subhicks1d never
    writes the table entries with |n + alpha| > irad (subresample_hicks.f90:124:
    n = 4 for alpha > 0, n = -4 for alpha < 0) and subresample allocates the
    table without clearing it (:48), yet sums over j = -4..4 (:71-73): an
    output whose window reaches such an entry adds whatever the heap held
    (seen here: 1e16 in one run of three).  The oracle and the library use 0
    there, the window's value outside its support.  Only the index arithmetic
    is repeated here, to select the samples; the values come from the
    reference build."""
    f = np.float32
    keep = np.zeros(n1out, bool)
    for i1 in range(1, n1out + 1):
        tout = f(i1 - 1) * f(d1out)
        it0 = int(np.floor(float(tout / f(d1in)) + 0.5)) + 1
        alpha = (f(it0 - 1) * f(d1in) - tout) / f(d1in)
        ia = min(int(np.floor(float((f(0.5) - alpha) * f(10.0)) + 0.5)) + 1, 10)
        jmin, jmax = max(1 - it0, -4), min(n1in - it0, 4)
        unset = 4 if ia <= 5 else (-4 if ia >= 7 else None)
        keep[i1 - 1] = unset is None or not (jmin <= unset <= jmax)
    return keep


@pytest.mark.skipif(not O.ref_available(), reason="reference tree not built into oracle/_ref")
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}_{c[2]}_{c[3]}")
def test_subresample_matches_reference_build(case):
    R = O.ref()
    if not hasattr(R, "dawref_subresample"):
        pytest.skip("oracle/_ref predates the resampler: rerun oracle/build_ref.sh")
    n1in, n2, d1in, d1out, n1out = case
    n1out = _n1out(n1in, d1in, d1out, n1out)
    x = _traces(n2, n1in, n1in)
    ref = np.zeros((n2, n1out), np.float32)
    R.dawref_subresample(x, ref, n1in, n2, n1out, d1in, d1out)
    keep = _defined_outputs(n1in, n1out, d1in, d1out)
    assert keep.sum() >= min(n1out, 8)
    assert np.array_equal(O.subresample(x, d1in, d1out, n1out)[:, keep], ref[:, keep])


def test_table_is_a_windowed_sinc():
    c = O.subhicks1d()                   # (ialpha, n + 4)
    alpha = 0.5 - np.arange(10) / 10.0
    # zero shift (ialpha = 6): 1 at the node, the sinc's zeros elsewhere up to
    # the float residue of sin(float(pi) * n)
    assert c[5, 4] == 1.0 and np.abs(np.delete(c[5], 4)).max() < 1e-6
    # every shift sums to ~1 (partition of unity of the windowed sinc)
    assert np.abs(c.sum(axis=1) - 1.0).max() < 2e-2
    # symmetric: shift alpha at node n equals shift -alpha at node -n
    for ia in range(1, 10):
        ib = 10 - ia                     # alpha(ib) = -alpha(ia)
        assert np.allclose(c[ia], c[ib][::-1], atol=1e-6), ia
    # values inside the window support are the Kaiser-windowed sinc
    from numpy import i0, sinc
    for ia in range(10):
        for n in range(-4, 5):
            x = n + alpha[ia]
            if abs(x) <= 4.0:
                want = i0(6.31 * np.sqrt(1.0 - (x / 4.0) ** 2)) / i0(6.31) * sinc(x)
                assert abs(c[ia, n + 4] - want) < 2e-5, (ia, n)


def test_band_limited_signal_is_interpolated():
    d1in, d1out, n1in = 1e-3, 7.3e-4, 2000
    t = np.arange(n1in) * d1in
    f = 10.0
    x = np.sin(2 * np.pi * f * t + 0.3).astype(np.float32)[None, :]
    n1out = int((n1in - 1) * d1in / d1out)
    y = O.subresample(x, d1in, d1out, n1out)[0]
    tout = np.arange(n1out) * np.float32(d1out)
    want = np.sin(2 * np.pi * f * tout + 0.3)
    inner = slice(10, n1out - 10)            # away from the truncated windows at both ends
    # the table holds 10 shifts per sample, so the delay is quantised to
    # +-0.05 sample: a phase error of up to 2 pi f (0.05 d1in); the window adds
    # as much again.  A delay off by a whole sample (a sign slip in alpha)
    # would give 2 pi f d1in, eight times this bound.
    quant = 2 * np.pi * f * 0.05 * d1in
    assert np.abs(y[inner] - want[inner]).max() < 2.5 * quant


def test_equal_sampling_is_identity_and_operator_is_linear():
    x = _traces(3, 500, 1)
    y = _traces(3, 500, 2)
    same = O.subresample(x, 1e-3, 1e-3, 500)
    assert np.abs(same - x).max() < 1e-5 * np.abs(x).max()
    a = np.float32(0.37)
    lhs = O.subresample(a * x + y, 8e-4, 4e-3, 100)
    rhs = a * O.subresample(x, 8e-4, 4e-3, 100) + O.subresample(y, 8e-4, 4e-3, 100)
    assert np.abs(lhs - rhs).max() < 1e-5 * np.abs(rhs).max()
