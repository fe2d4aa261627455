"""Object set-up on the device against the oracle: roots, height ranges,
AABBs (footprint_roots, triangle_height_range, triangle_aabb), and the
error paths of the C ABI."""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["c2_256", "curved", "rotated"])
def test_setup_matches_oracle(torch_gpu, kind):
    import dmap_amd
    if kind == "c2_256":
        tex, tris, strength = W.c2_map(256), W.quad_plane(8, 8.0), 0.1
    else:
        tex = np.random.default_rng(3).random((64, 64)).astype(np.float32)
        tris = scenes.curved_quad() if kind == "curved" else scenes.rotated_chart_quad()
        strength = 0.05
    tile = dmap_amd.DisplacementTile(tex, strength)
    obj = dmap_amd.ObjectIntersector(tile, tris)
    sc = oracle.Scene(W.close_tile(tex).astype(np.float64), strength, 0.0, True, tris)
    for k in range(len(tris)):
        g, r = obj.triangle_info(k), sc.triangle_info(k)
        assert g["root_level"] == r["root_level"]
        assert np.array_equal(np.sort(g["roots"], axis=0), np.sort(r["roots"], axis=0))
        assert g["h_range"][0] <= r["h_range"][0] and g["h_range"][1] >= r["h_range"][1]
        assert g["h_range"][0] >= np.nextafter(np.float32(r["h_range"][0]), -np.inf)
        assert g["h_range"][1] <= np.nextafter(np.float32(r["h_range"][1]), np.inf)
        assert (g["aabb"][0] <= r["aabb"][0]).all() and (g["aabb"][1] >= r["aabb"][1]).all()
        # padded for fp32 rays by 8 ulps of the coordinates, as rtc_bounds does
        pad = 1e-5 * (1 + np.abs(r["aabb"]).max())
        assert (g["aabb"][1] - g["aabb"][0] <= (r["aabb"][1] - r["aabb"][0]) + 2 * pad).all()


def test_error_paths(torch_gpu):
    import dmap_amd
    with pytest.raises(dmap_amd.DmapError, match="2\\^L"):
        dmap_amd.DisplacementTile(np.zeros((6, 6), np.float32))
    tile = dmap_amd.DisplacementTile(np.zeros((8, 8), np.float32))
    bad = W.quad_plane(1, 1.0)[:1].copy()
    bad[0, 20:24] = bad[0, 18:20].tolist() * 2  # all three texture coordinates equal
    with pytest.raises(dmap_amd.DmapError, match="degenerate"):
        dmap_amd.ObjectIntersector(tile, bad)
    obj = dmap_amd.ObjectIntersector(tile, W.quad_plane(1, 1.0))
    import torch
    rays = torch.zeros((8, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(dmap_amd.DmapError, match="prim_mask"):
        obj.intersect(rays, prim_mask=5)
    with pytest.raises(dmap_amd.DmapError):
        obj.intersect(torch.zeros((7, 4), dtype=torch.float32, device="cuda"))


def test_degenerate_rays_miss(torch_gpu):
    import dmap_amd
    import torch
    tile = dmap_amd.DisplacementTile(np.random.default_rng(1).random((32, 32)).astype(np.float32), 0.05)
    obj = dmap_amd.ObjectIntersector(tile, W.quad_plane(1, 1.0))
    r = np.zeros((8, 6), np.float32)
    r[0:3] = 0.5
    r[2] = 1.0
    r[5] = -1.0  # straight down: a hit
    r[7] = np.inf
    r[3:6, 1] = 0.0  # zero direction
    r[0, 2] = np.nan  # NaN origin
    r[6, 3], r[7, 3] = 2.0, 1.0  # empty interval
    r[7, 4] = 0.5  # t_max before the surface
    r[3, 5] = np.inf  # infinite direction component
    res = obj.intersect(torch.from_numpy(r).cuda())
    t = res["t"].cpu().numpy()
    assert np.isfinite(t[0]) and abs(t[0] - 1.0) < 0.06
    assert np.isinf(t[1:]).all()
    assert (res["prim"].cpu().numpy()[1:] == -1).all()
    empty = obj.intersect(torch.zeros((8, 0), dtype=torch.float32, device="cuda"))
    assert empty["t"].numel() == 0


def test_root_level_15_matches_oracle(torch_gpu):
    """The highest root level the walk supports: a 1024-texel tile repeated 35
    times across the quad puts the roots at level 15, whose trail bits are the
    top four of the walk's 64-bit trail."""
    import dmap_amd
    from tests.gpu_util import gpu_hits, oracle_hits
    from tests.parity import compare
    tex = np.random.default_rng(3).random((1024, 1024)).astype(np.float32)
    tris = scenes.span_quad(10, 45)
    tile = dmap_amd.DisplacementTile(tex, 2e-5)
    obj = dmap_amd.ObjectIntersector(tile, tris)
    assert [obj.triangle_info(k)["root_level"] for k in range(2)] == [15, 15]
    rays = scenes.rays(20000, seed=41)
    g = gpu_hits(obj, rays)
    ref, d, _ = oracle_hits(W.close_tile(tex).astype(np.float64), 2e-5, 0.0, True, tris, rays)
    cmp = compare(g, ref, d, tris, 1024)
    print({k: v for k, v in cmp.items() if k != "bad_index"})
    assert cmp["hits_ref"] > 8000
    assert cmp["unexplained"] == 0


def test_root_level_16_is_refused(torch_gpu):
    """A root level above 15 would shift the walk's trail past its 64 bits:
    object creation refuses it with the level in the message."""
    import dmap_amd
    tex = np.random.default_rng(3).random((1024, 1024)).astype(np.float32)
    tile = dmap_amd.DisplacementTile(tex, 2e-5)
    with pytest.raises(dmap_amd.DmapError, match="root level 16"):
        dmap_amd.ObjectIntersector(tile, scenes.span_quad(10, 80))
