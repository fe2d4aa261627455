"""Kernels #2/#3 parity: GPU closest hits against the oracle's on the same
seeded rays (tests/parity.py states the bar and the bands)."""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests.parity import compare

pytestmark = pytest.mark.gpu


def run_gpu(wl, rays, options=None):
    import dmap_amd
    import torch
    tile = dmap_amd.DisplacementTile(wl.texels, wl.strength, wl.offset, wl.repeat)
    obj = dmap_amd.ObjectIntersector(tile, wl.triangles)
    res = obj.intersect(torch.from_numpy(rays).cuda(), options=options, stats=True)
    torch.cuda.synchronize()
    g = {k: res[k].cpu().numpy() for k in ("t", "prim")}
    g["st"] = res["st"].cpu().numpy().T
    g["ng"] = res["ng"].cpu().numpy().T
    g["texel"] = res["texel"].cpu().numpy().T
    g["stats"] = {k: v.cpu().numpy() for k, v in res["stats"].items()}
    return g


def run_both(wl, idx=None, options=None, oracle_kw=None):
    rays = wl.rays if idx is None else np.ascontiguousarray(wl.rays[:, idx])
    g = run_gpu(wl, rays, options)
    sc = oracle.Scene(wl.nodes, wl.strength, wl.offset, wl.repeat, wl.triangles)
    o = rays[0:3].T.astype(np.float64)
    d = rays[3:6].T.astype(np.float64)
    ref = sc.intersect(o, d, rays[6].astype(np.float64), rays[7].astype(np.float64), stats=True,
                       **(oracle_kw or {}))
    return g, ref, d


def report(cmp):
    return {k: v for k, v in cmp.items() if k not in ("bad_index", "ray_flag")}


def test_c1_closest_hit_parity(torch_gpu):
    wl = W.c1_small()
    g, ref, d = run_both(wl)
    cmp = compare(g, ref, d, wl.triangles, wl.texels.shape[0])
    print(report(cmp))
    assert cmp["hits_ref"] > 40000
    assert cmp["unexplained"] == 0, report(cmp)
    check_counters(g, ref, "C1", wl)


def check_counters(g, ref, label, wl, idx=None):
    """The per-ray counters of TraversalStats (displaced_intersector.h:89-105)
    on the GPU: node tests per level sum to the node tests; tests after the
    final hit are 0 on a miss, within the totals, and exclude the hit's leaf.
    The walk from the roots (walk_start "roots", the reference's walk)
    finds the same hits, bit for bit, as the default walk from the block
    that holds the ray's clip to the triangle's height shell, with more node
    tests; against the oracle, level by level over the rays, its padded fp32
    tests keep every node the double tests keep and test a passing cell's
    four children together, so it tests at least as many nodes per level."""
    import dmap_amd
    rays = wl.rays if idx is None else np.ascontiguousarray(wl.rays[:, idx])
    r = run_gpu(wl, rays, dmap_amd.TraversalOptions(walk_start="roots"))
    for key in ("t", "prim", "st", "ng", "texel"):
        assert np.array_equal(r[key], g[key]), (label, key)
    O = ref["stats"]
    for run, G in (("shell", g["stats"]), ("roots", r["stats"])):
        per = G["node_tests_per_level"]
        assert per.shape[0] == 16
        assert (per.sum(axis=0) == G["node_tests"]).all()
        hit = np.isfinite(g["t"])
        for name, total in (("node_tests_after_hit", "node_tests"), ("leaf_tests_after_hit", "leaf_tests")):
            assert (G[name][~hit] == 0).all() and (G[name] <= G[total]).all()
        assert (G["leaf_tests_after_hit"][hit] <= G["leaf_tests"][hit] - 1).all()
        print(f"[{label}] {run}: node tests per level {per.sum(axis=1).astype(np.int64).tolist()}; after hit: "
              f"nodes {int(G['node_tests_after_hit'].sum())}, leaves {int(G['leaf_tests_after_hit'].sum())}",
              flush=True)
    # The leaves met can differ in order (a block need not align with its
    # parent's), so per-ray leaf counts are not compared; node tests are fewer.
    assert g["stats"]["node_tests"].sum() < r["stats"]["node_tests"].sum()
    gl = r["stats"]["node_tests_per_level"].sum(axis=1).astype(np.int64)
    ol = O["node_tests_per_level"].sum(axis=0)
    assert ol[16:].sum() == 0
    print(f"[{label}] oracle: node tests per level {ol[:16].tolist()}; after hit: nodes "
          f"{int(O['node_tests_after_hit'].sum())}, leaves {int(O['leaf_tests_after_hit'].sum())}", flush=True)
    assert (gl >= ol[:16]).all()
    assert r["stats"]["node_tests"].sum() >= O["node_tests"].sum()


def test_c2_small_closest_hit_parity(torch_gpu):
    """The incoherent workload's structure at 256^2 and 2 x 64k rays."""
    wl = W.c2_incoherent(n=256, width=256, height=256, n_bounce=65536)
    g, ref, d = run_both(wl)
    cmp = compare(g, ref, d, wl.triangles, wl.texels.shape[0])
    print(report(cmp))
    assert cmp["unexplained"] == 0, report(cmp)
    check_counters(g, ref, "C2 at 256^2", wl)
