"""GPU against the oracle across the traversal options and object kinds:
node-test modes, the two-triangle leaf, any-hit, prim_mask, the fp64
device run, interpolated normals, general charts, clamp and offset."""
import numpy as np
import pytest

from dmap_amd import workload as W
from tests import scenes
from tests.gpu_util import gpu_hits, oracle_hits
from tests.parity import compare

pytestmark = pytest.mark.gpu


def build(tex, strength, offset, repeat, tris, build_mode="fold"):
    import dmap_amd
    tile = dmap_amd.DisplacementTile(tex, strength, offset, repeat, build_mode)
    return dmap_amd.ObjectIntersector(tile, tris)


def summary(c):
    return {k: v for k, v in c.items() if k != "bad_index"}


@pytest.mark.parametrize("bound,k", [("box", 1), ("slab", 1), ("both", 0), ("both", 3), ("slab", 8)])
def test_bound_modes(torch_gpu, bound, k):
    import dmap_amd
    tex = np.random.default_rng(4).random((128, 128)).astype(np.float32)
    tris = W.quad_plane(1, 1.0)
    obj = build(tex, 0.06, 0.0, True, tris)
    rays = scenes.rays(40000, 5)
    opt = dmap_amd.TraversalOptions(bound=bound, slab_max_level=k)
    g = gpu_hits(obj, rays, options=opt)
    ref, d, _ = oracle_hits(W.close_tile(tex), 0.06, 0.0, True, tris, rays, bound=bound, slab_max_level=k)
    c = compare(g, ref, d, tris, 128)
    print(summary(c))
    assert c["unexplained"] == 0, summary(c)


@pytest.mark.parametrize("kind", ["curved", "rotated"])
def test_general_normals_and_charts(torch_gpu, kind):
    tex = np.random.default_rng(6).random((64, 64)).astype(np.float32)
    tris = scenes.curved_quad() if kind == "curved" else scenes.rotated_chart_quad()
    obj = build(tex, 0.05, 0.0, True, tris)
    rays = scenes.rays(40000, 7)
    g = gpu_hits(obj, rays)
    ref, d, _ = oracle_hits(W.close_tile(tex), 0.05, 0.0, True, tris, rays)
    c = compare(g, ref, d, tris, 64)
    print(summary(c))
    assert c["hits_ref"] > 10000
    assert c["unexplained"] == 0, summary(c)
    both = np.isfinite(g["t"]) & np.isfinite(ref["t"])
    assert np.abs(g["bary"][both] - ref["bary"][both]).max() < 1e-4


def test_clamp_and_offset_and_direct_build(torch_gpu):
    tex = np.random.default_rng(8).random((64, 64)).astype(np.float32)
    tris = W.quad_plane(1, 1.0)
    obj = build(tex, 0.07, -0.035, False, tris, "direct")
    rays = scenes.rays(30000, 9)
    g = gpu_hits(obj, rays)
    ref, d, _ = oracle_hits(W.close_tile(tex, repeat=False), 0.07, -0.035, False, tris, rays, build="direct")
    c = compare(g, ref, d, tris, 64)
    print(summary(c))
    assert c["unexplained"] == 0, summary(c)


def test_two_triangle_leaf(torch_gpu):
    import dmap_amd
    tex = np.random.default_rng(10).random((64, 64)).astype(np.float32)
    tris = W.quad_plane(1, 1.0)
    obj = build(tex, 0.05, 0.0, True, tris)
    rays = scenes.rays(40000, 11)
    g = gpu_hits(obj, rays, options=dmap_amd.TraversalOptions(leaf="two_triangle"))
    ref, d, _ = oracle_hits(W.close_tile(tex), 0.05, 0.0, True, tris, rays, leaf="two_triangle")
    # The texel mesh's own edges (texel boundaries and the (i+1, j)-(i, j+1)
    # diagonals) are where fp32 and fp64 crossings may fall on either side.
    st = np.where(np.isfinite(ref["t"])[:, None], ref["st"], g["st"])
    f = st * 64 - np.floor(st * 64)
    near = (np.minimum(f, 1 - f) < 1e-4).any(axis=1) | (np.abs(f.sum(axis=1) - 1) < 1e-4)
    c = compare(g, ref, d, tris, 64)
    bad = c["bad_index"]
    print(summary(c), "on mesh edges:", int(near[bad].sum()))
    assert (~near[bad]).sum() == 0


def test_any_hit_and_prim_mask(torch_gpu):
    import dmap_amd
    import torch
    tex = W.c2_map(256)
    tris = W.quad_plane(8, 8.0)
    obj = build(tex, 0.1, 0.0, True, tris)
    wl = W.c2_incoherent(n=256, width=128, height=64, n_bounce=8192)
    rays = wl.rays
    g = gpu_hits(obj, rays)
    occ = obj.occluded(torch.from_numpy(rays).cuda()).cpu().numpy().astype(bool)
    assert (occ == np.isfinite(g["t"])).all()
    short = rays.copy()
    short[7] = np.where(np.isfinite(g["t"]), g["t"] * 0.999, np.inf)
    occ2 = obj.occluded(torch.from_numpy(short).cuda()).cpu().numpy().astype(bool)
    assert not occ2[np.isfinite(g["t"])].any()
    # prim_mask = k: the walk of triangle k alone (the emitter-pdf door).
    for k in (0, 37, 127):
        gk = gpu_hits(obj, rays, prim_mask=k)
        ref, d, _ = oracle_hits(W.close_tile(tex), 0.1, 0.0, True, tris[k:k + 1], rays)
        ref["prim"] = np.where(ref["prim"] >= 0, k, -1)
        c = compare(gk, ref, d, tris, 256)
        assert c["unexplained"] == 0, summary(c)
        assert ((gk["prim"] == k) | (gk["prim"] == -1)).all()


def test_fp64_device_run(torch_gpu):
    """The same solve in fp64 on the device: agrees with the double oracle
    to the Newton tolerance, and gives the fp32 error report."""
    import dmap_amd
    wl = W.c2_incoherent(n=256, width=128, height=128, n_bounce=16384)
    obj = build(wl.texels, 0.1, 0.0, True, wl.triangles)
    g64 = gpu_hits(obj, wl.rays, options=dmap_amd.TraversalOptions(precision="fp64"))
    g32 = gpu_hits(obj, wl.rays)
    ref, d, _ = oracle_hits(wl.nodes, 0.1, 0.0, True, wl.triangles, wl.rays)
    c = compare(g64, ref, d, wl.triangles, 256, rel=1e-6)
    print("fp64", summary(c))
    assert c["unexplained"] == 0 and c["hitmiss_flips"] == 0
    both = np.isfinite(g32["t"]) & np.isfinite(g64["t"])
    err = np.abs(g32["t"][both] - g64["t"][both]) / g64["t"][both]
    print(f"fp32 vs fp64 t: max {err.max():.2e}, p99 {np.quantile(err, 0.99):.2e}")
    assert (np.isfinite(g32["t"]) == np.isfinite(g64["t"])).mean() > 0.9999


def test_reorder_does_not_change_results(torch_gpu):
    """The three ray orders (sorted, as given, near-first) give every ray the
    same hit, bit for bit; the near-first order puts the bounce rays (their
    origins on the surface, inside the object's box) ahead of the camera's."""
    import dmap_amd
    wl = W.c2_incoherent(n=256, width=128, height=128, n_bounce=16384)
    obj = build(wl.texels, 0.1, 0.0, True, wl.triangles)
    out = {order: gpu_hits(obj, wl.rays, options=dmap_amd.TraversalOptions(reorder=order))
           for order in ("sort", "given", "near_first")}
    for order in ("sort", "near_first"):
        for key in ("t", "prim", "st", "ng", "texel"):
            assert np.array_equal(out[order][key], out["given"][key]), (order, key)


def _near_classes(rays, lo, hi):
    """The near-first class of each ray, restated in double: for a ray whose
    fp32 origin lies in the box [lo, hi], floor(-2 log2(L / diagonal))
    clamped to 0 .. 15, L its path length in the box from tmin to where it
    leaves the box (or tmax); 15 when L / diagonal is not a positive finite
    number; 16 for the other rays. Also whether -2 log2(L / diagonal) lies
    within 1e-4 of a class boundary, where the device's fp32 may round to
    the other side."""
    o, d = rays[0:3], rays[3:6].astype(np.float64)
    tmin, tmax = rays[6].astype(np.float64), rays[7].astype(np.float64)
    inside = np.all((o >= lo[:, None]) & (o <= hi[:, None]), axis=0)
    lo64, hi64, o64 = lo.astype(np.float64), hi.astype(np.float64), o.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        wall = np.where(d > 0, hi64[:, None], lo64[:, None])
        t_axis = np.where(d != 0, (wall - o64) / d, np.inf)
        t_exit = np.minimum(tmax, t_axis.min(axis=0))
        x = (t_exit - tmin) * np.linalg.norm(d, axis=0) / np.linalg.norm(hi64 - lo64)
        v = -2.0 * np.log2(x)
    ok = (x > 0) & np.isfinite(v)
    cls = np.where(ok, np.clip(np.floor(np.where(ok, v, 0.0)), 0, 15), 15).astype(np.int64)
    cls = np.where(inside, cls, 16)
    near_edge = inside & ok & (np.abs(v - np.round(v)) < 1e-4)
    return cls, near_edge


def test_ray_orders_are_the_stated_permutations(torch_gpu):
    """dmap_ray_order. Near-first: a permutation that takes the rays whose
    origin is in the object's box first, in classes of their path length in
    the box (half an octave of the diagonal each, longest first), every class
    in its given order, then the other rays in their given order; checked
    against the classes restated on the host. Sort: a permutation. Given,
    and any order under 4,096 rays: the identity."""
    import torch

    import dmap_amd
    wl = W.c2_incoherent(n=256, width=128, height=128, n_bounce=16384)
    obj = build(wl.texels, 0.1, 0.0, True, wl.triangles)
    n = wl.rays.shape[1]
    rays = torch.from_numpy(np.ascontiguousarray(wl.rays)).cuda()
    boxes = np.stack([obj.triangle_info(k)["aabb"] for k in range(wl.triangles.shape[0])])
    lo, hi = boxes[:, 0].min(axis=0).astype(np.float32), boxes[:, 1].max(axis=0).astype(np.float32)
    cls, edge = _near_classes(wl.rays, lo, hi)
    n_in = int((cls < 16).sum())
    assert 0 < n_in < n and len(np.unique(cls[cls < 16])) >= 4, np.bincount(cls)

    perm = obj.ray_order(rays, "near_first").cpu().numpy().astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert (cls[perm[:n_in]] < 16).all() and (cls[perm[n_in:]] == 16).all()
    assert (np.diff(perm[n_in:]) > 0).all()  # the outside rays as given
    inner = perm[:n_in][~edge[perm[:n_in]]]  # the rays whose class the fp32 rounding cannot move
    c = cls[inner]
    assert (np.diff(c) >= 0).all()  # longest first
    same = np.diff(c) == 0
    assert (np.diff(inner)[same] > 0).all()  # stable within a class

    srt = obj.ray_order(rays, "sort").cpu().numpy().astype(np.int64)
    assert np.array_equal(np.sort(srt), np.arange(n)) and not np.array_equal(srt, np.arange(n))
    assert np.array_equal(obj.ray_order(rays, "given").cpu().numpy(), np.arange(n))
    small = rays[:, :4095].contiguous()
    for order in ("near_first", "sort"):
        assert np.array_equal(obj.ray_order(small, order).cpu().numpy(), np.arange(4095))
    with pytest.raises(dmap_amd.DmapError):
        obj.ray_order(rays, "shuffled")


@pytest.mark.parametrize("kind", ["flat", "curved"])
def test_unpolished_fp32_root_meets_the_bar(torch_gpu, kind):
    """The fp32 Newton root itself, without the two double polishing steps
    (polish=False): the fp32 search decides hit/miss, the triangle and the
    texel, so those equal the polished run's; its t, st and n sit within
    1e-3 (relative) of the double oracle's, ten times the parity bar. The
    fp32 solve stops at 16 ulps of its residual, which on rays that leave
    the surface at a short distance is up to a few 1e-4 of t; the polish
    brings the reported values within 1e-4 (the parity tests)."""
    import dmap_amd
    if kind == "flat":
        wl = W.c2_incoherent(n=256, width=128, height=128, n_bounce=16384)
        tris, tex, strength, nodes, rays = wl.triangles, wl.texels, 0.1, wl.nodes, wl.rays
    else:
        tex = np.random.default_rng(3).random((64, 64)).astype(np.float32)
        tris, strength = scenes.curved_quad(), 0.05
        nodes = W.close_tile(tex).astype(np.float64)
        rays = scenes.rays(20000, seed=12)
    obj = build(tex, strength, 0.0, True, tris)
    raw = gpu_hits(obj, rays, options=dmap_amd.TraversalOptions(polish=False))
    pol = gpu_hits(obj, rays)
    ref, d, _ = oracle_hits(nodes, strength, 0.0, True, tris, rays)
    c = compare(raw, ref, d, tris, tex.shape[0], rel=1e-3)
    print("unpolished", summary(c))
    assert c["unexplained"] == 0 and c["hitmiss_flips"] == 0
    for key in ("prim", "texel"):
        assert np.array_equal(raw[key], pol[key]), key
    both = np.isfinite(raw["t"])
    assert (both == np.isfinite(pol["t"])).all()
    assert not np.array_equal(raw["t"][both], pol["t"][both])  # the polish does move the reported t


def _tilted_flat_quad():
    """rotated_chart_quad (constant normal, rotated chart) with its positions
    and normals turned about a skew axis and shifted: a tangent frame that is
    no axis permutation."""
    tris = scenes.rotated_chart_quad()
    a = np.array([1.0, 2.0, 0.5]) / np.linalg.norm([1.0, 2.0, 0.5])
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    R = np.eye(3) + np.sin(0.7) * K + (1 - np.cos(0.7)) * K @ K
    for row in tris:
        for v in range(3):
            row[3 * v:3 * v + 3] = R @ row[3 * v:3 * v + 3] + np.array([0.3, -0.2, 0.5])
            row[9 + 3 * v:12 + 3 * v] = R @ row[9 + 3 * v:12 + 3 * v]
    return tris


@pytest.mark.parametrize("case", ["c2", "c2_fp64", "tilted", "span", "offset_clamp", "box"])
def test_walk_start_does_not_change_results(torch_gpu, case):
    """A constant-normal triangle's walk from the finest block that holds
    the ray's clip to its height shell (the default) finds the same hits,
    bit for bit, as the walk from its roots (walk_start "roots", the
    reference's walk), with fewer node tests: on the incoherent workload's
    structure (fp32 and fp64), a tilted frame with a rotated chart, a chart
    over 6 x 6 repeats of the tile from negative coordinates, negative
    heights with clamping, and the box-only bound."""
    import dmap_amd
    opt = {}
    if case in ("c2", "c2_fp64"):
        wl = W.c2_incoherent(n=256, width=128, height=128, n_bounce=16384)
        tex, strength, offset, repeat, tris, rays = wl.texels, 0.1, 0.0, True, wl.triangles, wl.rays
        if case == "c2_fp64":
            opt["precision"] = "fp64"
    else:
        tex = np.random.default_rng(21).random((64, 64)).astype(np.float32)
        strength, offset, repeat = 0.06, 0.0, True
        tris = {"tilted": _tilted_flat_quad, "span": lambda: scenes.span_quad(-2.5, 3.5)}.get(
            case, lambda: W.quad_plane(1, 1.0))()
        if case == "offset_clamp":
            offset, repeat = -0.05, False
        if case == "box":
            opt["bound"] = "box"
        rays = scenes.rays(40000, 23)
        if case == "tilted":
            # The scene's rays, carried into the tilted quad's frame.
            rays = rays.copy()
            a = np.array([1.0, 2.0, 0.5]) / np.linalg.norm([1.0, 2.0, 0.5])
            K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
            R = np.eye(3) + np.sin(0.7) * K + (1 - np.cos(0.7)) * K @ K
            rays[0:3] = (R @ rays[0:3].astype(np.float64) + np.array([[0.3], [-0.2], [0.5]])).astype(np.float32)
            rays[3:6] = (R @ rays[3:6].astype(np.float64)).astype(np.float32)
    obj = build(tex, strength, offset, repeat, tris)
    shell = gpu_hits(obj, rays, options=dmap_amd.TraversalOptions(**opt), stats=True)
    roots = gpu_hits(obj, rays, options=dmap_amd.TraversalOptions(walk_start="roots", **opt), stats=True)
    hits = int(np.isfinite(shell["t"]).sum())
    ns, nr = int(shell["stats"]["node_tests"].sum()), int(roots["stats"]["node_tests"].sum())
    print(f"[walk start] {case}: {hits} hits of {rays.shape[1]}; node tests {ns} from the shell's block, "
          f"{nr} from the roots", flush=True)
    assert hits > rays.shape[1] // 10
    for key in ("t", "prim", "st", "bary", "ng", "texel"):
        assert np.array_equal(shell[key], roots[key]), key
    assert ns < nr
