"""Synthetic workload generators: deterministic, the stated shapes and
distributions."""
import numpy as np

from dmap_amd import workload as W


def test_deterministic():
    a, b = W.c1_small(n=64, width=64, height=64), W.c1_small(n=64, width=64, height=64)
    assert (a.texels == b.texels).all() and (a.rays == b.rays).all()


def test_fbm_range_and_tiling():
    f = W.fbm(128, 4, 4, 1)
    assert f.min() >= -1 and f.max() <= 1 and f.std() > 0.05
    r = W.ridged(128, 2, 8, 2)
    assert r.min() >= 0 and r.max() <= 1


def test_quad_plane():
    t = W.quad_plane(8, 8.0)
    assert t.shape == (128, 24)
    # every triangle wound counter-clockwise seen from +z, normals +z, uv = xy / 8
    p = t[:, :9].reshape(-1, 3, 3)
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert (n[:, 2] > 0).all()
    assert np.allclose(t[:, 18:24].reshape(-1, 3, 2), p[:, :, :2] / 8)


def test_pinhole_unit_directions_tile_order():
    o, d = W.pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0), 60, 64, 32)
    assert o.shape == (2048, 3)
    assert np.allclose(np.linalg.norm(d, axis=1), 1, atol=1e-6)


def test_bounce_rays_leave_the_surface():
    tex = W.c2_map(256)
    nodes = W.close_tile(tex)
    o, d = W.bounce_rays(nodes, 0.1, 0.0, True, 8.0, 5000, 3)
    h, hs, ht = W.field(nodes, 0.1, 0.0, True, o[:, 0] / 8.0, o[:, 1] / 8.0)
    assert (o[:, 2] > h - 1e-6).all()
    n = np.stack([-hs * 8, -ht * 8, np.full(len(h), 64.0)], 1)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert (np.sum(n * d, axis=1) > -1e-6).all()


def test_c2_shape():
    wl = W.c2_incoherent(n=256, width=64, height=32, n_bounce=100)
    assert wl.rays.shape == (8, 2048 + 100) and wl.n_primary == 2048
    assert np.isinf(wl.rays[7]).all() and (wl.rays[6] == 0).all()
