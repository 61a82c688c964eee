"""Interior tiles of the one-step kernel take a branch-free body (no sponge
factors, clamps, free surface or injection work), which must give the
reference's fields bit for bit like the general body.

The grids here are large enough that most tiles of every tile shape are
interior; the cases put the receivers (record band, residual injection) near
the surface and deep inside the grid, so that interior tiles meet record-band
rows and injection rows cut tiles out of the interior set.  `_interior_tiles`
counts the tiles that qualify, so each case is known to exercise the path.
"""
import numpy as np
import pytest

import oracle as O
from cases import Case, rel_l2

pytestmark = pytest.mark.gpu

TILES = [(8, 128), (16, 64), (16, 128), (20, 64), (24, 64), (32, 64), (32, 128), (64, 64)]


def _interior_tiles(c, tz, tx, rows_blocked):
    """Tiles whose pressure rectangle (tile plus a one/two-cell ring) lies in
    the sponge-free interior, below the free surface and off the blocked
    (injection) rows."""
    n1e, n2e = c.n1 + 2 * c.npml, c.n2 + 2 * c.npml
    n = 0
    for z0 in range(0, n1e, tz):
        for x0 in range(0, n2e, tx):
            pzl, pzh, pxl, pxh = z0 - 1, z0 + tz + 1, x0 - 1, x0 + tx + 1
            ok = pzl >= c.npml and pzh < c.npml + c.n1 and pxl >= c.npml and pxh < c.npml + c.n2
            ok = ok and not (c.ofs == 1 and pzl <= c.npml)
            ok = ok and not any(pzl <= z <= pzh for z in rows_blocked)
            n += ok
    return n


def _window_rows(tab, npml, n1e):
    # rows a 21-point Hicks window around each point touches (0-based extended)
    z = tab[0] - 1 if tab.ndim == 1 else tab[0, :] - 1
    return sorted({int(r) + k for r in np.atleast_1d(z) for k in range(-10, 11) if 0 <= int(r) + k < n1e})


@pytest.fixture
def tiles(gpu_daw):
    yield gpu_daw
    gpu_daw.set_tile(0)
    gpu_daw.set_snapshot_format("fp32")


def _case(**kw):
    # 224 x 484 extended grid; vp up to ~8300 m/s, so dt = 0.5 ms keeps the
    # scheme stable; a 30 Hz wavelet starts early enough to reach the tiles
    return Case(nz=200, nx=460, npml=12, nt=200, dt=5e-4, f0=30.0, **kw)


@pytest.mark.parametrize("src_z,rec_z", [(10.0, 15.0), (650.0, 700.0)])
def test_forward_interior_tiles(tiles, src_z, rec_z):
    c = _case(src_z=src_z, rec_z=rec_z)
    n1e = c.n1 + 2 * c.npml
    src_rows = _window_rows(c.src, c.npml, n1e)
    for type_src in (0, 1):
        ref = O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec, type_src=type_src)
        assert np.abs(ref).max() > 0
        for tile in TILES:
            assert _interior_tiles(c, tile[0], tile[1], src_rows) >= 4
            tiles.set_tile(*tile)
            print("forward interior tiles", src_z, rec_z, type_src, tile, flush=True)
            sis = tiles.submodeling(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec, type_src=type_src)
            assert np.array_equal(sis, ref), (tile, type_src, rel_l2(sis, ref))


@pytest.mark.parametrize("rec_z", [15.0, 700.0])
def test_hessian_interior_tiles(tiles, rec_z):
    c = _case(rec_z=rec_z)
    n1e = c.n1 + 2 * c.npml
    rec_rows = _window_rows(c.rec, c.npml, n1e)
    res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
    ref = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    ref_e = O.esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
    ref16 = {s: O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec, snap=s) for s in ("bf16", "fp16")}
    for tile in TILES:
        assert _interior_tiles(c, tile[0], tile[1], rec_rows) >= 4
        tiles.set_tile(*tile)
        print("hessian interior tiles", rec_z, tile, flush=True)
        a = tiles.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
        assert np.array_equal(a, ref), (tile, rel_l2(a, ref))
        e = tiles.submodeling_adjoint_esource(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
        assert np.array_equal(e, ref_e), (tile, rel_l2(e, ref_e))
        for snap in ("bf16", "fp16"):
            tiles.set_snapshot_format(snap)
            q = tiles.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
            tiles.set_snapshot_format("fp32")
            assert np.array_equal(q, ref16[snap]), (tile, snap, rel_l2(q, ref16[snap]))


def test_modeling_batch_deep_datum_interior_tiles(tiles):
    """Datum receivers deep in the grid: their record-band rows fall inside
    interior tiles of the forward sweep, which store p into the band."""
    import torch
    c = _case(dat_z=900.0)
    ns = 3
    xmax = (c.n2 - 1) * float(c.h)
    src = O.acqui_table(np.full(ns, 10.0), np.linspace(0.3, 0.7, ns) * xmax, c.h, c.npml, 1)
    s_dev = torch.from_numpy(c.s).cuda()
    nr, nd = c.rec.shape[1], c.dat.shape[1]
    for tile in [(16, 64), (32, 128)]:
        tiles.set_tile(*tile)
        sis = torch.zeros((ns, nr, c.nt), dtype=torch.float32, device="cuda")
        dat = torch.zeros((ns, nd, c.nt), dtype=torch.float32, device="cuda")
        print("modeling_batch deep datum", tile, flush=True)
        tiles.modeling_batch(c.vp, c.npml, c.h, 1, c.nt, c.dt, s_dev, src, c.rec, sis, dat_tab=c.dat, datum_dev=dat)
        sis, dat = sis.cpu().numpy(), dat.cpu().numpy()
        for s in range(ns):
            osis, owf = O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, src[:, s].copy(), c.rec, want_wavefield=True)
            assert np.array_equal(sis[s], osis)
            assert np.array_equal(dat[s], O.resample(owf, c.npml, c.dat, 1))


def _ring_case(tz, edge, **kw):
    """A case whose 21-row Hicks window (of the source and of every receiver)
    ends exactly on a ring row of one tz-row tile: edge "top" puts the
    window's last row on the top ring row z0-1 of the tile at z0, "bottom" its
    first row on the bottom ring row z0+tz+1.  That tile's own rows hold no
    other window row, so it must still apply the terms of its ring row.  The
    offset of +0.37 / -0.38 cell from the nearest node gives that edge row a
    nonzero coefficient (it is exactly zero on the other side)."""
    npml, h = 12, 10.0
    if edge == "top":
        z0 = -(-71 // tz) * tz
        zrec, frac = z0 - 11, 0.37
    else:
        z0 = max(0, -(-(49 - tz) // tz) * tz)
        zrec, frac = z0 + tz + 11, -0.38
    depth = (zrec - npml + frac) * h
    c = _case(src_z=depth, rec_z=depth, **kw)
    assert c.npml == npml and float(c.h) == h
    assert int(c.rec[0, 0]) - 1 == zrec and int(c.src[0]) - 1 == zrec
    t = np.asarray(O.hicks2d_tile(int(c.rec[2, 0]), int(c.rec[3, 0]), int(c.rec[4, 0]), c.ofs))
    assert np.abs(t[:, 20 if edge == "top" else 0]).max() > 0      # axis 1 is the z offset -10..10
    ring = z0 - 1 if edge == "top" else z0 + tz + 1
    assert ring == (zrec + 10 if edge == "top" else zrec - 10)
    return c


@pytest.mark.parametrize("edge", ["top", "bottom"])
def test_injection_window_edge_on_tile_ring(tiles, edge):
    """Tiles whose pressure rows (z0-1 .. z0+tz+1) miss every Hicks window
    row skip the injection lookup; a window that ends on a tile's ring row
    must still reach that tile, in the forward sweep (source) and in the
    back-propagation of the residuals (receivers)."""
    for tile in TILES:
        c = _ring_case(tile[0], edge)
        ref = O.forward(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
        res = c.rng.standard_normal((c.rec.shape[1], c.nt)).astype(np.float32)
        ref_h = O.esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
        tiles.set_tile(*tile)
        print("injection window edge on tile ring", edge, tile, flush=True)
        sis = tiles.submodeling(c.vp, c.npml, c.h, 1, c.nt, c.dt, c.s, c.src, c.rec)
        assert np.array_equal(sis, ref), (tile, rel_l2(sis, ref))
        a = tiles.submodeling_adjoint_esource_data(c.vp0, c.npml, c.h, 1, c.nt, c.dt, res, c.rec)
        assert np.array_equal(a, ref_h), (tile, rel_l2(a, ref_h))
