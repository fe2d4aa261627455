"""The top-level BVH (heap-ordered, complete, padded with empty leaves) over
objects of 1 to 200 triangles: the closest hit through it equals the oracle's
flat loop over every triangle (ObjectIntersector::intersect_flat,
displaced_intersector.cpp:578-599), and any-hit agrees with closest-hit. The
counts cover a single leaf, padded trees (3, 5, 127, 129, 200 triangles), a
full tree (64, 128), and both sides of the 128-leaf limit of the workgroup
copy in LDS (single-object launches read the tree from LDS up to 128 leaves
and from global memory beyond), each with the host builder's tree (median
splits) and the device builder's (Morton order)."""
import numpy as np
import pytest

from dmap_amd import workload as W
from tests.gpu_util import gpu_hits, oracle_hits
from tests.parity import compare

pytestmark = pytest.mark.gpu


def _rays(n, lo, hi, seed):
    """Half from above towards points of [lo, hi] (xy), half near the
    surface over it in every direction."""
    rng = np.random.default_rng(seed)
    r = np.zeros((8, n), dtype=np.float32)
    half = n // 2
    o = np.stack([rng.uniform(lo[0] - 0.5, hi[0] + 0.5, half), rng.uniform(lo[1] - 0.5, hi[1] + 0.5, half),
                  rng.uniform(0.5, 2.0, half)])
    tgt = np.stack([rng.uniform(lo[0], hi[0], half), rng.uniform(lo[1], hi[1], half), np.zeros(half)])
    r[0:3, :half] = o
    r[3:6, :half] = tgt - o
    m = n - half
    r[0:3, half:] = np.stack([rng.uniform(lo[0], hi[0], m), rng.uniform(lo[1], hi[1], m),
                              rng.uniform(-0.05, 0.15, m)])
    d = rng.normal(size=(3, m))
    r[3:6, half:] = d / np.linalg.norm(d, axis=0)
    r[6] = 0.0
    r[7] = np.inf
    return r


@pytest.mark.parametrize("bvh", ["host", "device"])
@pytest.mark.parametrize("n_tri", [1, 3, 5, 64, 127, 128, 129, 200])
def test_bvh_closest_and_any_hit(torch_gpu, n_tri, bvh):
    import dmap_amd
    torch = torch_gpu
    tris = W.quad_plane(10, 10.0)[:n_tri]
    tex = W.c2_map(64)
    tile = dmap_amd.DisplacementTile(tex, 0.1, 0.0, True)
    obj = dmap_amd.ObjectIntersector(tile, tris, bvh=bvh)
    xy = tris[:, 0:9].reshape(-1, 3, 3)[:, :, 0:2].reshape(-1, 2)
    rays = _rays(4096, xy.min(axis=0), xy.max(axis=0), seed=n_tri)
    g = gpu_hits(obj, rays)
    ref, d, _ = oracle_hits(W.close_tile(tex).astype(np.float64), 0.1, 0.0, True, tris, rays)
    c = compare(g, ref, d, tris, 64)
    print(n_tri, {k: v for k, v in c.items() if k != "bad_index"})
    assert c["hits_ref"] > 1000
    assert c["unexplained"] == 0
    occ = obj.occluded(torch.from_numpy(rays).cuda()).cpu().numpy().astype(bool)
    hit = np.isfinite(g["t"])
    # any-hit finds a crossing wherever closest-hit does, and none where it does not
    assert (occ == hit).mean() > 0.999, (occ != hit).sum()
