"""Kernels #2/#3 parity: GPU closest hits against the oracle's on the same
seeded rays (tests/parity.py states the bar and the bands)."""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests.parity import compare

pytestmark = pytest.mark.gpu


def run_both(wl, idx=None, options=None, oracle_kw=None):
    import dmap_amd
    import torch
    rays = wl.rays if idx is None else np.ascontiguousarray(wl.rays[:, idx])
    tile = dmap_amd.DisplacementTile(wl.texels, wl.strength, wl.offset, wl.repeat)
    obj = dmap_amd.ObjectIntersector(tile, wl.triangles)
    res = obj.intersect(torch.from_numpy(rays).cuda(), options=options, stats=True)
    torch.cuda.synchronize()
    g = {k: res[k].cpu().numpy() for k in ("t", "prim")}
    g["st"] = res["st"].cpu().numpy().T
    g["ng"] = res["ng"].cpu().numpy().T
    g["texel"] = res["texel"].cpu().numpy().T
    g["stats"] = {k: v.cpu().numpy() for k, v in res["stats"].items()}
    sc = oracle.Scene(wl.nodes, wl.strength, wl.offset, wl.repeat, wl.triangles)
    o = rays[0:3].T.astype(np.float64)
    d = rays[3:6].T.astype(np.float64)
    ref = sc.intersect(o, d, rays[6].astype(np.float64), rays[7].astype(np.float64), stats=True,
                       **(oracle_kw or {}))
    return g, ref, d


def report(cmp):
    return {k: v for k, v in cmp.items() if k != "bad_index"}


def test_c1_closest_hit_parity(torch_gpu):
    wl = W.c1_small()
    g, ref, d = run_both(wl)
    cmp = compare(g, ref, d, wl.triangles, wl.texels.shape[0])
    print(report(cmp))
    assert cmp["hits_ref"] > 40000
    assert cmp["unexplained"] == 0, report(cmp)
    # Work counters: the padded fp32 tests keep every node the double tests keep.
    assert g["stats"]["node_tests"].sum() >= ref["stats"]["node_tests"].sum()


def test_c2_small_closest_hit_parity(torch_gpu):
    """The incoherent workload's structure at 256^2 and 2 x 64k rays."""
    wl = W.c2_incoherent(n=256, width=256, height=256, n_bounce=65536)
    g, ref, d = run_both(wl)
    cmp = compare(g, ref, d, wl.triangles, wl.texels.shape[0])
    print(report(cmp))
    assert cmp["unexplained"] == 0, report(cmp)
