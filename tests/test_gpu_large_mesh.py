"""A base mesh of 51,200 triangles (a torus, interpolated normals, one tile
wrapped around it eight times): object set-up time, and closest hits against
the oracle's flat loop over every triangle (displaced_intersector.cpp:578-599)
on a ray sample, under tests/parity.py's bar (no unexplained difference in
hit/miss, t, st or n); the GPU's BVH stands in for that loop."""
import time

import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests.parity import compare

pytestmark = pytest.mark.gpu


def torus(nu=160, nv=160, R=1.0, r=0.35, tiles=8):
    """Rows p0 p1 p2, m0 m1 m2, t0 t1 t2 of a (nu x nv)-quad torus; the
    texture coordinates run over `tiles` repeats of the tile around the ring."""
    def pt(i, j):
        u, v = 2 * np.pi * i / nu, 2 * np.pi * j / nv
        c = np.array([np.cos(u), np.sin(u), 0.0])
        p = (R + r * np.cos(v)) * c + r * np.sin(v) * np.array([0, 0, 1.0])
        n = np.cos(v) * c + np.sin(v) * np.array([0, 0, 1.0])
        return p, n, np.array([tiles * i / nu, j / nv])
    rows = []
    for j in range(nv):
        for i in range(nu):
            (p00, n00, t00), (p10, n10, t10) = pt(i, j), pt(i + 1, j)
            (p01, n01, t01), (p11, n11, t11) = pt(i, j + 1), pt(i + 1, j + 1)
            rows.append(np.concatenate([p00, p10, p11, n00, n10, n11, t00, t10, t11]))
            rows.append(np.concatenate([p00, p11, p01, n00, n11, n01, t00, t11, t01]))
    return np.asarray(rows)


def test_large_mesh_setup_and_parity(torch_gpu):
    import dmap_amd
    torch = torch_gpu
    tris = torus()
    tex = W.c2_map(256)
    tile = dmap_amd.DisplacementTile(tex, 0.02, 0.0, True)
    t0 = time.perf_counter()
    obj = dmap_amd.ObjectIntersector(tile, tris)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    print(f"[large mesh] {len(tris)} triangles: object set-up {setup_s * 1e3:.1f} ms", flush=True)
    assert setup_s < 5.0
    rng = np.random.default_rng(3)
    n = 4096
    o = rng.normal(size=(n, 3)) * np.array([0.4, 0.4, 0.2]) + np.array([0.0, 0.0, 2.5])
    tgt = rng.normal(size=(n, 3)) * np.array([0.9, 0.9, 0.2])
    d = tgt - o
    rays = W.pack_rays(o.astype(np.float32), d.astype(np.float32))
    res = obj.intersect(torch.from_numpy(rays).cuda())
    g = {"t": res["t"].cpu().numpy(), "prim": res["prim"].cpu().numpy(), "st": res["st"].cpu().numpy().T,
         "ng": res["ng"].cpu().numpy().T}
    sc = oracle.Scene(W.close_tile(tex).astype(np.float64), 0.02, 0.0, True, tris)
    dd = rays[3:6].T.astype(np.float64)
    ref = sc.intersect(rays[0:3].T.astype(np.float64), dd)
    cmp = compare(g, ref, dd, tris, tex.shape[0])
    print({k: v for k, v in cmp.items() if k not in ("bad_index", "ray_flag")}, flush=True)
    assert cmp["hits_ref"] > 0.3 * n
    assert cmp["unexplained"] == 0
