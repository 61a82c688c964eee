"""daw_subresample_dev_noalloc (the entry point that allocates nothing: its
plan and table live in a fixed device array of the library) against
daw_subresample_dev, called through the C-ABI: bit-identical outputs on two
calls in a row, the second smaller than the first (the array keeps the first
call's contents, which the second must not read), and a size beyond the
array refused with an error before any GPU work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(L, fn, xin, n1out, d1in, d1out):
    import torch
    xout = torch.full((xin.shape[0], n1out), float("nan"), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rc = getattr(L, fn)(C.c_void_p(xin.data_ptr()), C.c_void_p(xout.data_ptr()), xin.shape[1], xin.shape[0], n1out,
                        C.c_float(d1in), C.c_float(d1out))
    assert rc == 0, L.daw_last_error()
    return xout.cpu().numpy()


def test_noalloc_matches_subresample_dev(gpu_daw):
    import torch
    L = gpu_daw.lib()
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    # call 1: 401 traces, 3000 samples at 0.8 ms -> 4 ms (601 outputs);
    # call 2: 37 traces, 900 samples at 1 ms -> 0.6 ms (1499 outputs, up-sampling), then
    # 5 traces 500 -> 97 (smaller than both)
    for ntr, n1in, d1in, d1out, n1out in ((401, 3000, 8e-4, 4e-3, 601), (37, 900, 1e-3, 6e-4, 1499),
                                          (5, 500, 1e-3, 5e-3, 97)):
        print(f"noalloc case: {ntr} traces {n1in} -> {n1out}", flush=True)
        xin = torch.randn((ntr, n1in), generator=g, device="cuda", dtype=torch.float32)
        want = _run(L, "daw_subresample_dev", xin, n1out, d1in, d1out)
        got = _run(L, "daw_subresample_dev_noalloc", xin, n1out, d1in, d1out)
        assert np.isfinite(want).all()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_noalloc_refuses_oversize(gpu_daw):
    import torch
    L = gpu_daw.lib()
    xin = torch.zeros((1, 70000), dtype=torch.float32, device="cuda")
    xout = torch.zeros((1, 65537), dtype=torch.float32, device="cuda")
    rc = L.daw_subresample_dev_noalloc(C.c_void_p(xin.data_ptr()), C.c_void_p(xout.data_ptr()), 70000, 1, 65537,
                                       C.c_float(1e-3), C.c_float(1e-3))
    assert rc != 0 and b"65536" in L.daw_last_error()
