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


def _level_table(rec, levels):
    """test_traversal_cost.cpp:387-418: per level the tests, the prunes of
    each kind, the median interval lengths (over tests where all three pass
    with t_max = inf) and the share where the slab's is shorter; then check
    (2)'s crossover levels."""
    rows = []
    for lv in range(levels):
        r = rec[rec["level"] == lv]
        full = (r["box_length"] >= 0) & (r["slab_length"] >= 0) & (r["both_length"] >= 0)
        med = [float(np.median(r[k][full])) if full.any() else 0.0 for k in ("box_length", "slab_length", "both_length")]
        rows.append({"tests": len(r), "slab_prunes": int(((r["box_pass"] == 1) & (r["slab_pass"] == 0)).sum()),
                     "box_prunes": int(((r["box_pass"] == 0) & (r["slab_pass"] == 1)).sum()),
                     "both_prune": int(((r["box_pass"] == 0) & (r["slab_pass"] == 0)).sum()),
                     "median": med, "slab_shorter": float((r["slab_length"][full] < r["box_length"][full]).mean())
                     if full.any() else 0.0})
    crossover = pruning = -1
    for lv, row in enumerate(rows):
        if row["tests"] and row["median"][1] < row["median"][0] and crossover == lv - 1:
            crossover = lv
        if row["slab_prunes"] > row["box_prunes"] and pruning == lv - 1:
            pruning = lv
    return rows, crossover, pruning


def test_node_records_against_the_oracle(torch_gpu):
    """NodeTestRecord (displaced_intersector.h:77-87) on the GPU against the
    oracle's, along the box walk from the roots (the walk the reference
    records, test_traversal_cost.cpp:377-386), on 3,000 rays of the C1 map
    that reach the pyramid. Every oracle record has the GPU's record of the
    same (ray, triangle, cell); a test the double oracle passes the padded
    fp32 test passes; interval lengths agree to 1e-4 (absolute, plus 1e-4
    relative), or the oracle's fails and the GPU's is at most 1e-4 long;
    the per-level table of the reference's check (2) agrees with the same
    crossover levels: prune counts within 3% of the level's tests (the
    padded tests keep some nodes the double tests prune: at the texel level
    the slab's median interval is 3e-5 long and its padding about 4e-6) and
    medians within the per-record bar. The GPU's own records match its counters, ray by ray in seq
    order, also on the default walk from the shell's block."""
    import dmap_amd
    torch = torch_gpu
    wl = W.c1_small()
    sc = oracle.Scene(wl.nodes, wl.strength, wl.offset, wl.repeat, wl.triangles)
    allr = sc.intersect(wl.rays[0:3].T, wl.rays[3:6].T, wl.rays[6], wl.rays[7], bound="box", stats=True)
    idx = np.nonzero(allr["stats"]["node_tests"] > 0)[0]
    idx = np.sort(np.random.default_rng(7).choice(idx, size=3000, replace=False))
    rays = np.ascontiguousarray(wl.rays[:, idx])
    ref = sc.records(rays[0:3].T, rays[3:6].T, rays[6], rays[7], bound="box")
    tile = dmap_amd.DisplacementTile(wl.texels, wl.strength, wl.offset, wl.repeat)
    obj = dmap_amd.ObjectIntersector(tile, wl.triangles)
    rt = torch.from_numpy(rays).cuda()
    for start in ("roots", "shell"):
        opt = dmap_amd.TraversalOptions(bound="box", walk_start=start)
        res = obj.intersect(rt, options=opt, stats=True, records=4 * len(ref) + 1024)
        torch.cuda.synchronize()
        rec = res["records"]
        nt = res["stats"]["node_tests"].cpu().numpy()
        assert res["record_count"] == len(rec) == nt.sum()
        assert (np.bincount(rec["ray"], minlength=len(idx)) == nt).all()
        order = np.lexsort((rec["seq"], rec["ray"]))
        starts = np.concatenate([[0], np.cumsum(nt)[:-1]])
        assert (rec["seq"][order] == np.arange(len(rec)) - np.repeat(starts, nt)).all()
        if start == "roots":
            gpu = rec
    print(f"[records] oracle {len(ref)}, GPU from the roots {len(gpu)}, from the shell's block {len(rec)}", flush=True)
    assert len(ref) > 20000 and len(gpu) >= len(ref)
    # Match by (ray, triangle, level, i, j).
    def key(r):
        return ((r["ray"].astype(np.int64) * 4 + r["prim"]) * 16 + r["level"]) * (1 << 40) + \
               (r["i"].astype(np.int64) + 2048) * (1 << 20) + (r["j"].astype(np.int64) + 2048)
    kr, kg = key(ref), key(gpu)
    assert len(np.unique(kr)) == len(kr) and len(np.unique(kg)) == len(kg)
    pos = np.searchsorted(kg[np.argsort(kg)], kr)
    g = gpu[np.argsort(kg)][np.minimum(pos, len(gpu) - 1)]
    found = key(g) == kr
    print(f"[records] oracle node tests the GPU walk did not make: {int((~found).sum())}", flush=True)
    assert (~found).sum() == 0
    assert (g["box_pass"][ref["box_pass"] == 1] == 1).all() and (g["slab_pass"][ref["slab_pass"] == 1] == 1).all()
    flips = int((g["box_pass"] != ref["box_pass"]).sum() + (g["slab_pass"] != ref["slab_pass"]).sum())
    for name in ("box_length", "slab_length", "both_length"):
        o_len, g_len = ref[name], g[name].astype(np.float64)
        ok = o_len >= 0
        err = np.abs(g_len[ok] - o_len[ok])
        assert (g_len[ok] >= 0).all() and (err <= 1e-4 + 1e-4 * o_len[ok]).all(), (name, err.max())
        assert (g_len[~ok] <= 1e-4).all(), name
        print(f"[records] {name}: {ok.sum()} both measured, max |GPU - oracle| {err.max():.2e}; "
              f"{int((g_len[~ok] >= 0).sum())} kept only by the GPU's padding", flush=True)
    print(f"[records] pass flags that differ (GPU passes, oracle fails): {flips}", flush=True)
    levels = int(ref["level"].max()) + 1
    rows_o, cross_o, prune_o = _level_table(ref, levels)
    rows_g, cross_g, prune_g = _level_table(g, levels)
    for lv in range(levels):
        ro, rg = rows_o[lv], rows_g[lv]
        print(f"[records] level {lv}: oracle {ro}, GPU {rg}", flush=True)
        assert rg["tests"] == ro["tests"]  # the matched records: one per oracle test
        for a in ("slab_prunes", "box_prunes", "both_prune"):
            assert abs(rg[a] - ro[a]) <= max(3, 0.03 * ro["tests"]), (lv, a)
        for mo, mg in zip(ro["median"], rg["median"]):
            assert abs(mg - mo) <= 1e-4 + 1e-4 * mo, (lv, mo, mg)
    print(f"[records] crossover levels: oracle {cross_o}, {prune_o}; GPU {cross_g}, {prune_g}", flush=True)
    assert (cross_g, prune_g) == (cross_o, prune_o)
