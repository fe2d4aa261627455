"""Rays that reach the special cases of the box and slab tests: direction
components that are exactly zero (the clip's parallel-ray branch, in world
space for the BVH and in tangent space for the node tests), rays in the
coordinate planes, grazing rays along the surface, and directions scaled by
2^-40 and 2^40 (t scaled by 2^40 and 2^-40, so products and quotients sit
far from 1 in magnitude). The closest hit through the GPU equals the
oracle's on the same rays (Aabb3::clip node_bounds.cpp:25-44 for the
parallel case; ObjectIntersector::intersect_flat displaced_intersector.cpp
:578-599 for the object), to the parity bar of tests/parity.py."""
import numpy as np
import pytest

from dmap_amd import workload as W
from tests.gpu_util import gpu_hits, oracle_hits
from tests.parity import compare

pytestmark = pytest.mark.gpu

N_LEAF = 64
STRENGTH = 0.1


def _axis_rays(n, seed):
    """Origins over and around the surface of quad_plane(3, 3.0); directions
    along +-x, +-y, +-z, in the xz / yz / xy planes, and general."""
    rng = np.random.default_rng(seed)
    r = np.zeros((8, n), dtype=np.float32)
    r[0] = rng.uniform(-0.2, 3.2, n)
    r[1] = rng.uniform(-0.2, 3.2, n)
    r[2] = rng.uniform(-0.05, 0.2, n)
    kind = rng.integers(0, 5, n)
    d = rng.normal(size=(3, n))
    axis = rng.integers(0, 3, n)
    sign = np.where(rng.random(n) < 0.5, -1.0, 1.0)
    for k in range(n):
        if kind[k] == 0:  # along one axis
            d[:, k] = 0.0
            d[axis[k], k] = sign[k]
        elif kind[k] == 1:  # in a coordinate plane
            d[axis[k], k] = 0.0
        elif kind[k] == 2:  # straight down from above
            r[2, k] = 1.0
            d[:, k] = (0.0, 0.0, -1.0)
        elif kind[k] == 3:  # grazing along the surface, in the xy plane
            r[2, k] = rng.uniform(0.0, 0.1)
            d[2, k] = 0.0
    d /= np.linalg.norm(d, axis=0)
    r[3:6] = d
    r[6] = 0.0
    r[7] = np.inf
    return r


@pytest.fixture(scope="module")
def scene():
    import dmap_amd
    tris = W.quad_plane(3, 3.0)
    tex = W.c2_map(N_LEAF)
    tile = dmap_amd.DisplacementTile(tex, STRENGTH, 0.0, True)
    obj = dmap_amd.ObjectIntersector(tile, tris)
    return obj, tris, W.close_tile(tex).astype(np.float64)


@pytest.mark.parametrize("scale_exp", [0, -40, 40])
def test_axis_and_scaled_rays_match_oracle(torch_gpu, scene, scale_exp):
    obj, tris, nodes = scene
    rays = _axis_rays(6000, seed=11)
    rays[3:6] *= np.float32(2.0 ** scale_exp)  # exact: a power of two
    assert np.isfinite(rays[3:6]).all()
    g = gpu_hits(obj, rays)
    ref, d, _ = oracle_hits(nodes, STRENGTH, 0.0, True, tris, rays)
    c = compare(g, ref, d.astype(np.float64), tris, N_LEAF)
    print(scale_exp, {k: v for k, v in c.items() if k != "bad_index"})
    assert c["hits_ref"] > 1000
    assert c["unexplained"] == 0


def test_scaled_directions_scale_t(torch_gpu, scene):
    """Scaling every direction by 2^k scales each hit's t by 2^-k: the hits
    of the three runs are the same points."""
    obj, tris, _ = scene
    rays = _axis_rays(6000, seed=12)
    base = gpu_hits(obj, rays)
    hit = np.isfinite(base["t"])
    for e in (-40, 40):
        r = rays.copy()
        r[3:6] *= np.float32(2.0 ** e)
        g = gpu_hits(obj, r)
        same = np.isfinite(g["t"]) == hit
        assert same.mean() > 0.999, (~same).sum()
        both = hit & np.isfinite(g["t"])
        rel = np.abs(g["t"][both].astype(np.float64) * 2.0 ** e - base["t"][both]) / base["t"][both]
        assert np.quantile(rel, 0.999) < 1e-4, np.quantile(rel, 0.999)
