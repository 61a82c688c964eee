"""The benchmark workload at its full size (201 x 801 grid, h = 8 m, 20-cell
PML, nt = 3000, 64 shots, 401 receivers), through properties that hold at any
size, so that the oracle's share stays a few single-shot runs:

- the batched forward of all 64 shots is bit-identical to the oracle's FP32
  forward of the same shots (two checked: the first and the last), whatever
  the batch split (all 64 at once, 16 per launch) and shot order;
- the exact transpose passes the dot test <F s, d> = <s, F^T d> at full size;
- a redatuming pass over all 64 shots gives each shot the gathers it gets
  when redatumed alone (batching changes no shot's result beyond the
  band-pass's FFT pairing of traces).
"""
import importlib

import numpy as np
import pytest

import oracle as O
from cases import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2(gpu_daw):
    import torch
    syn = importlib.import_module("daw_amd.synth")
    w = syn.workload(2)
    vp, c0 = syn.models(w)
    h, dt = np.float32(w.h), np.float32(w.dt)
    src = O.acqui_table(w.src_z, w.src_x, h, w.npml, w.ofs)
    rec = O.acqui_table(w.rec_z, w.rec_x, h, w.npml, w.ofs)
    dat = O.acqui_table(w.dat_z, w.dat_x, h, w.npml, w.ofs)
    s = syn.ricker(w.nt, float(dt), w.f0, float(h))
    return dict(w=w, vp=vp, c0=c0, h=h, dt=dt, src=src, rec=rec, dat=dat, s=s, s_dev=torch.from_numpy(s).cuda())


def _model(daw, P, src, max_batch=0):
    import torch
    w = P["w"]
    d = torch.zeros((src.shape[1], P["rec"].shape[1], w.nt), dtype=torch.float32, device="cuda")
    daw.modeling_batch(P["vp"], w.npml, P["h"], w.ofs, w.nt, P["dt"], P["s_dev"], src, P["rec"], d,
                       max_batch=max_batch)
    return d.cpu().numpy()


def test_c2_forward_batches_bitexact(gpu_daw, c2):
    P, w = c2, c2["w"]
    assert P["src"].shape[1] == 64 and P["rec"].shape[1] == 401 and (w.nz, w.nx, w.nt) == (201, 801, 3000)
    print("C2 modeling_batch 64 shots, one launch", flush=True)
    all64 = _model(gpu_daw, P, P["src"])
    print("C2 modeling_batch 64 shots, 16 per launch", flush=True)
    by16 = _model(gpu_daw, P, P["src"], max_batch=16)
    assert np.array_equal(all64, by16)
    print("C2 modeling_batch 64 shots, reversed order", flush=True)
    rev = _model(gpu_daw, P, np.ascontiguousarray(P["src"][:, ::-1]))
    assert np.array_equal(all64, rev[::-1])
    for k in (0, 63):
        print("C2 oracle forward, shot", k, flush=True)
        ref = O.forward(P["vp"], w.npml, P["h"], w.ofs, w.nt, P["dt"], P["s"], P["src"][:, k].copy(), P["rec"])
        assert np.abs(ref).max() > 0
        assert np.array_equal(all64[k], ref), (k, rel_l2(all64[k], ref))


def test_c2_transpose_dot_product(gpu_daw, c2):
    P, w = c2, c2["w"]
    rng = np.random.default_rng(12)
    s = rng.standard_normal(w.nt).astype(np.float32)
    d = rng.standard_normal((P["rec"].shape[1], w.nt)).astype(np.float32)
    src = P["src"][:, 31].copy()
    print("C2 forward and transpose, shot 31", flush=True)
    fs = gpu_daw.submodeling(P["vp"], w.npml, P["h"], w.ofs, w.nt, P["dt"], s, src, P["rec"]).astype(np.float64)
    ftd = gpu_daw.submodeling_transpose(P["vp"], w.npml, P["h"], w.ofs, w.nt, P["dt"], d, src,
                                        P["rec"]).astype(np.float64)
    lhs, rhs = float(np.sum(fs * d)), float(np.sum(s.astype(np.float64) * ftd))
    scale = float(np.linalg.norm(fs) * np.linalg.norm(d))
    print("C2 dot test", lhs, rhs, abs(lhs - rhs) / scale, flush=True)
    assert abs(lhs - rhs) <= 1e-6 * scale


def test_c2_redatum_batch_each_shot_as_alone(gpu_daw, c2):
    import torch
    P, w = c2, c2["w"]
    its = 2
    print("C2 observed data", flush=True)
    dobs = torch.from_numpy(_model(gpu_daw, P, P["src"])).cuda()
    nd = P["dat"].shape[1]

    def run(cols):
        datum = torch.zeros((len(cols), nd, w.nt), dtype=torch.float32, device="cuda")
        hist = gpu_daw.redatum_batch(P["c0"], w.npml, P["h"], w.ofs, w.nt, P["dt"], P["s_dev"],
                                     np.ascontiguousarray(P["src"][:, cols]), P["rec"], P["dat"],
                                     dobs[cols].contiguous(), w.mu, its, datum_dev=datum, band=w.band)
        return datum.cpu().numpy(), hist

    print("C2 redatum_batch 64 shots", flush=True)
    all_d, all_h = run(list(range(64)))
    for k in (0, 40, 63):
        print("C2 redatum_batch shot", k, "alone", flush=True)
        one_d, one_h = run([k])
        assert np.abs(one_d).max() > 0
        assert rel_l2(all_d[k], one_d[0]) <= 1e-6, (k, rel_l2(all_d[k], one_d[0]))
        assert np.allclose(all_h[k], one_h[0], rtol=1e-9, atol=0), k
