"""The time resampler on the GPU (daw_subresample, daw_subresample_dev)
against the oracle's subresample, bit for bit: the same REAL index
arithmetic (done once per output sample on the host) and the same sum of at
most nine products per sample, in the same order, with no fused multiply-add."""
import numpy as np
import pytest

import oracle as O
from test_oracle_resample import CASES, _n1out, _traces

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}_{c[2]}_{c[3]}")
def test_subresample_vs_oracle(gpu_daw, case):
    n1in, n2, d1in, d1out, n1out = case
    n1out = _n1out(n1in, d1in, d1out, n1out)
    x = _traces(n2, n1in, 11 * n1in)
    print("subresample", n1in, n2, d1in, d1out, n1out, flush=True)
    got = gpu_daw.subresample(x, d1in, d1out, n1out)
    assert np.array_equal(got, O.subresample(x, d1in, d1out, n1out))
    if case[4] is None:                  # the default length is nto of sub_redatum_DAW.f90:24
        assert gpu_daw.subresample(x, d1in, d1out).shape == (n2, n1out)


def test_subresample_dev_gathers_of_a_batch(gpu_daw):
    """dt -> dto of the redatumed gathers of 64 shots x 401 receivers at the
    201x801 workload's sampling (nt 3000 at 0.8 ms to 4 ms), on device."""
    import torch
    ntr, n1in, d1in, d1out = 64 * 401, 3000, 8e-4, 4e-3
    n1out = _n1out(n1in, d1in, d1out, None)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    xin = torch.randn((ntr, n1in), generator=g, device="cuda", dtype=torch.float32)
    xout = torch.full((ntr, n1out), float("nan"), device="cuda", dtype=torch.float32)
    print("subresample_dev", ntr, n1in, n1out, flush=True)
    gpu_daw.subresample_dev(xin, xout, d1in, d1out)
    rows = np.r_[0:3, ntr // 2, ntr - 3:ntr]
    x = xin[rows].cpu().numpy()
    assert np.array_equal(xout[rows].cpu().numpy(), O.subresample(x, d1in, d1out, n1out))
    assert bool(torch.isfinite(xout).all())
