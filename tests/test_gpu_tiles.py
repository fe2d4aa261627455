"""A displacement asset loaded from disk (tiles.load_asset) is the same
resident tile as one built from the same texels: same nodes, same pyramid,
same hits; PNG maps are box-downsampled and offset by -strength * midlevel
as DisplacementAsset does (displacement_asset.cpp:16-35)."""
import numpy as np
import pytest

from dmap_amd import tiles as T
from dmap_amd import workload as W

pytestmark = pytest.mark.gpu


def test_load_asset_png_and_tile_file(torch_gpu, tmp_path):
    import dmap_amd
    torch = torch_gpu
    rng = np.random.default_rng(6)
    img = rng.integers(0, 65536, size=(128, 128))
    T.write_png(tmp_path / "h.png", img, depth=16, filters=[4])
    tile = T.load_asset(tmp_path / "h.png", strength=0.08, midlevel=0.5, resolution=64)
    ref_tex = T.downsample_box(img / 65535.0, 64).astype(np.float32)
    ref = dmap_amd.DisplacementTile(ref_tex, 0.08, -0.04, True)
    assert np.array_equal(tile.nodes(), ref.nodes())
    for k in range(tile.n_levels):
        a, b = tile.level(k), ref.level(k)
        assert all(np.array_equal(a[c], b[c]) for c in a)
    T.save_tile(tmp_path / "h.dmt", ref_tex, 0.08, -0.04, True)
    tile2 = T.load_asset(tmp_path / "h.dmt")
    o, d = W.pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 40.0, 64, 64)
    rays = torch.from_numpy(W.pack_rays(o, d)).cuda()
    res = [dmap_amd.ObjectIntersector(t, W.quad_plane(1, 1.0)).intersect(rays) for t in (tile, ref, tile2)]
    for k in ("t", "st", "ng", "prim", "texel"):
        assert torch.equal(res[0][k], res[1][k]) and torch.equal(res[2][k], res[1][k])
