"""Argument checks of the launch knobs and the batched entry points (CPU:
every case fails, or succeeds, before the library touches the GPU)."""
import ctypes as C

import numpy as np


def test_snapshot_format_knob(daw):
    assert daw.lib().daw_set_snapshot_format(3) == -1
    for f in ("bf16", "fp16", "fp32"):
        daw.set_snapshot_format(f)
    assert daw.lib().daw_set_snapshot_format(-1) == -1
    assert b"snapshot format" in daw.lib().daw_last_error()


def test_tile_and_steps_knobs(daw):
    L = daw.lib()
    try:
        for rows, cols in [(12, 64), (16, 96), (8, 64), (64, 128), (-16, 64)]:
            assert L.daw_set_tile(rows, cols) == -1, (rows, cols)
            assert b"no compiled kernel" in L.daw_last_error()
        for rows, cols in [(8, 128), (16, 64), (16, 128), (20, 64), (24, 64), (32, 64), (32, 128), (64, 64)]:
            assert L.daw_set_tile(rows, cols) == 0, (rows, cols)
        assert L.daw_set_steps(3) == -1
        # 64x64 has no two-step kernel: fixing both must be refused either way round
        assert L.daw_set_tile(64, 64) == 0
        assert L.daw_set_steps(2) == -1
        assert L.daw_set_tile(16, 64) == 0
        assert L.daw_set_steps(2) == 0
        assert L.daw_set_tile(64, 64) == -1 and b"two-step" in L.daw_last_error()
        assert L.daw_set_steps(0) == 0
        assert L.daw_set_tile(64, 64) == 0
    finally:
        L.daw_set_steps(0)
        L.daw_set_tile(0, 64)


def test_two_step_march_knobs(daw):
    L = daw.lib()
    try:
        for rows in (8, 20, 56, 64):
            assert L.daw_set_tile(rows, 52) == -1 and b"two-step column march" in L.daw_last_error()
        for rows in (16, 24, 32, 40, 48):
            assert L.daw_set_tile(rows, 52) == 0
            assert daw.get_tile() == (rows, 52)
        assert L.daw_set_steps(1) == -1 and b"two steps per launch" in L.daw_last_error()
        assert L.daw_set_steps(2) == 0
        assert L.daw_set_tile(32, 58) == -1      # the one-step march with two steps per launch
        assert L.daw_set_tile(0, 64) == 0
        assert L.daw_set_steps(1) == 0
        assert L.daw_set_tile(32, 52) == -1 and b"two steps per launch" in L.daw_last_error()
    finally:
        L.daw_set_steps(0)
        L.daw_set_tile(0, 64)


def test_get_launch_mode_range(daw):
    L = daw.lib()
    s, r, c = C.c_int(), C.c_int(), C.c_int()
    for mode in (-1, 3):
        assert L.daw_get_launch(mode, C.byref(s), C.byref(r), C.byref(c)) == -1
    for mode in (0, 1, 2):
        assert L.daw_get_launch(mode, C.byref(s), C.byref(r), C.byref(c)) == 0
        assert s.value in (1, 2) and r.value in (8, 16, 20, 24, 32, 64) and c.value in (64, 128)


def test_batched_entry_points_reject_bad_arguments(daw):
    L = daw.lib()
    assert L.daw_redatum_batch(None) == -1 and b"null" in L.daw_last_error()
    assert L.daw_redatum_uneven(None) == -1 and b"null" in L.daw_last_error()
    assert L.daw_modeling_batch(None) == -1 and b"null" in L.daw_last_error()
    daw.kwsinc2d(1)
    try:
        a = daw.RedatumArgs()
        a.n1, a.n2, a.npml, a.h, a.dt, a.ofs, a.nt = 40, 90, 10, 10.0, 1e-3, 1, 100
        a.nshots = 0
        a.nrec = 1
        rc = L.daw_redatum_batch(C.byref(a))
        assert rc == -1 and b"missing input" in L.daw_last_error()
        a.nshots, a.npml = 1, 2               # sponge thinner than the stencil's reach
        rc = L.daw_redatum_batch(C.byref(a))
        assert rc == -1
    finally:
        daw.deallocatekws2d()


def test_subresample_rejects_bad_arguments(daw):
    L = daw.lib()
    x = np.zeros(10, np.float32)
    y = np.zeros(10, np.float32)
    for n1in, n2, n1out, d1in, d1out in [(0, 1, 5, 1e-3, 1e-3), (10, -1, 5, 1e-3, 1e-3), (10, 1, -5, 1e-3, 1e-3),
                                         (10, 1, 5, 0.0, 1e-3), (10, 1, 5, 1e-3, -1.0)]:
        assert L.daw_subresample(x, y, n1in, n2, n1out, d1in, d1out) == -1
        assert b"subresample" in L.daw_last_error()
    # nothing to do is not an error
    assert L.daw_subresample(x, y, 10, 0, 5, 1e-3, 1e-3) == 0
    assert L.daw_subresample(x, y, 10, 1, 0, 1e-3, 1e-3) == 0


def test_subresample_dev_rejects_overlap(daw):
    L = daw.lib()
    buf = np.zeros(1000, np.float32)
    p = buf.ctypes.data
    # output block inside the input block: refused before any device work
    assert L.daw_subresample_dev(p, p + 400, 100, 4, 50, 1e-3, 2e-3) == -1
    assert b"overlap" in L.daw_last_error()
