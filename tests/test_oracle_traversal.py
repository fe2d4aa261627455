"""The oracle's walk and leaves (displaced_intersector.cpp, node_bounds.cpp
restated) pinned by closed forms, brute force and the reference's own
validation invariants (validate_traversal.cpp:148-461,
validate_node_bounds.cpp:124-170, test_traversal_cost.cpp:242-297)."""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W

N_TEX = 16


def small_scene(n=N_TEX, amp=0.08, seed=5, repeat=True, flat=None):
    rng = np.random.default_rng(seed)
    tex = rng.random((n, n)).astype(np.float32) if flat is None else np.full((n, n), flat, np.float32)
    nodes = W.close_tile(tex, repeat).astype(np.float64)
    tris = W.quad_plane(1, 1.0)
    return oracle.Scene(nodes, amp, 0.0, repeat, tris), nodes * amp, tris


def ray_set(n, seed=9):
    """Half from above the unit quad towards it, half near the surface in random directions
    (the two distributions of traversal_test_util.cpp:83-111)."""
    rng = np.random.default_rng(seed)
    m = n // 2
    o1 = np.column_stack([rng.uniform(-0.2, 1.2, m), rng.uniform(-0.2, 1.2, m), rng.uniform(0.5, 1.5, m)])
    tgt = np.column_stack([rng.uniform(0, 1, m), rng.uniform(0, 1, m), np.zeros(m)])
    d1 = tgt - o1
    o2 = np.column_stack([rng.uniform(0, 1, n - m), rng.uniform(0, 1, n - m), rng.uniform(-0.1, 0.1, n - m)])
    d2 = rng.normal(size=(n - m, 3))
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    return np.vstack([o1, o2]), np.vstack([d1, d2])


def closed_form_hits(V, o, d):
    """Nearest root of the ray against every texel's bilinear patch
    S = (s, t, h(s, t)) (flat unit quad, normal +z): per texel the height
    along the ray is quadratic in t. Returns t (inf on a miss)."""
    n = V.shape[0] - 1
    w = 1.0 / n
    jj, ii = np.mgrid[0:n, 0:n]
    ii, jj = ii.ravel(), jj.ravel()
    h00, h10 = V[jj, ii], V[jj, ii + 1]
    h01, h11 = V[jj + 1, ii], V[jj + 1, ii + 1]
    ha, hb, T = h10 - h00, h01 - h00, h11 - h10 - h01 + h00
    best = np.full(len(o), np.inf)
    for r in range(len(o)):
        a0, a1 = (o[r, 0] - ii * w) / w, d[r, 0] / w
        b0, b1 = (o[r, 1] - jj * w) / w, d[r, 1] / w
        c0 = h00 + ha * a0 + hb * b0 + T * a0 * b0
        c1 = ha * a1 + hb * b1 + T * (a0 * b1 + a1 * b0)
        c2 = T * a1 * b1
        # (o_z - c0) + t (d_z - c1) - t^2 c2 = 0
        A, B, C = -c2, d[r, 2] - c1, o[r, 2] - c0
        ts = []
        lin = np.abs(A) < 1e-300
        with np.errstate(divide="ignore", invalid="ignore"):
            t_lin = np.where(lin & (B != 0), -C / B, np.inf)
            disc = B * B - 4 * A * C
            sq = np.sqrt(np.maximum(disc, 0))
            q = -0.5 * (B + np.sign(B + (B == 0)) * sq)
            r1 = np.where(~lin & (disc >= 0), q / A, np.inf)
            r2 = np.where(~lin & (disc >= 0) & (q != 0), C / q, np.inf)
        for tt in (t_lin, r1, r2):
            a, b = a0 + a1 * tt, b0 + b1 * tt
            ok = (tt > 0) & np.isfinite(tt) & (a >= 0) & (a <= 1) & (b >= 0) & (b <= 1)
            if ok.any():
                best[r] = min(best[r], tt[ok].min())
    return best


def test_newton_leaf_matches_bilinear_closed_form():
    sc, V, _ = small_scene()
    o, d = ray_set(1200)
    res = sc.intersect(o, d, threads=1)
    ref = closed_form_hits(V, o, d)
    hit_o, hit_r = np.isfinite(res["t"]), np.isfinite(ref)
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    cos = np.abs(np.sum(dn * res["ng"], axis=1))
    grazing = hit_o & (cos < 0.1)  # validate_traversal.cpp:116
    agree = (hit_o == hit_r)
    both = hit_o & hit_r
    assert hit_r.sum() > 500
    assert ((~agree) & ~grazing).sum() == 0
    # Newton stops within residual_tolerance = 1e-10 x mean edge of the ray
    # (displaced_intersector.cpp:150), a shift of tolerance / cos along it.
    tol = 4 * 1.14e-10 / np.maximum(cos[both], 0.1) + 1e-12
    dist = np.abs(res["t"][both] - ref[both]) * np.linalg.norm(d[both], axis=1)
    assert (dist[~grazing[both]] <= tol[~grazing[both]]).all()


def mt_mesh_hits(V, o, d):
    """Moller-Trumbore over the texel mesh split along (i+1, j)-(i, j+1)
    (displaced_tessellation.cpp:11-31): nearest t (inf on a miss)."""
    n = V.shape[0] - 1
    w = 1.0 / n
    jj, ii = np.mgrid[0:n, 0:n]
    ii, jj = ii.ravel(), jj.ravel()

    def P(i, j):
        return np.column_stack([i * w, j * w, V[j, i]])

    S00, S10, S01, S11 = P(ii, jj), P(ii + 1, jj), P(ii, jj + 1), P(ii + 1, jj + 1)
    A = np.vstack([S00, S10])
    B = np.vstack([S10, S11])
    C = np.vstack([S01, S01])
    e1, e2 = B - A, C - A
    best = np.full(len(o), np.inf)
    for r in range(len(o)):
        p = np.cross(d[r], e2)
        det = np.sum(e1 * p, axis=1)
        with np.errstate(divide="ignore", invalid="ignore"):
            inv = 1.0 / det
            s = o[r] - A
            b1 = np.sum(s * p, axis=1) * inv
            q = np.cross(s, e1)
            b2 = (q @ d[r]) * inv
            t = np.sum(e2 * q, axis=1) * inv
        ok = (det != 0) & (b1 >= 0) & (b1 <= 1) & (b2 >= 0) & (b1 + b2 <= 1) & (t > 0)
        if ok.any():
            best[r] = t[ok].min()
    return best


def test_two_triangle_leaf_matches_mesh():
    """validate_traversal.cpp:148-208: the two-triangle leaf is the texel mesh."""
    sc, V, _ = small_scene()
    o, d = ray_set(1200, seed=21)
    res = sc.intersect(o, d, leaf="two_triangle", threads=1)
    ref = mt_mesh_hits(V, o, d)
    hit_o, hit_r = np.isfinite(res["t"]), np.isfinite(ref)
    assert hit_r.sum() > 500
    assert (hit_o == hit_r).sum() >= len(o) - 2  # edge-on cases only
    both = hit_o & hit_r
    assert np.allclose(res["t"][both], ref[both], rtol=1e-10, atol=1e-12)


def test_never_miss_aimed_rays():
    """validate_traversal.cpp:392-461: rays aimed at surface points hit at or before them."""
    sc, _, tris = small_scene(seed=7)
    rng = np.random.default_rng(3)
    n = 3000
    k = rng.integers(0, 2, n)
    u = np.sqrt(rng.random(n))
    b = np.column_stack([1 - u, u * rng.random(n)])
    t = tris[k, 18:24].reshape(-1, 3, 2)
    st = t[:, 0] + b[:, :1] * (t[:, 1] - t[:, 0]) + b[:, 1:2] * (t[:, 2] - t[:, 0])
    S, _ = sc.surface_points(k.astype(np.int32), st)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ts = 1.14 * 10 ** (-1.5 + 2 * rng.random(n))
    res = sc.intersect(S - ts[:, None] * dirs, dirs, threads=1)
    cos = np.abs(np.sum(dirs * np.nan_to_num(res["ng"]), axis=1))
    tol = 1e-9 * (ts + 1.5) + 4 * 1.14e-10 / np.maximum(cos, 1e-3)
    assert np.isfinite(res["t"]).all()
    assert (res["t"] <= ts + tol).all()


@pytest.mark.parametrize("bound,k", [("box", 7), ("slab", 7), ("both", 0), ("both", 2), ("both", 7)])
def test_bound_modes_give_identical_hits(bound, k):
    """validate_traversal.cpp:320-390; and `both` never tests more nodes than `box`
    (test_traversal_cost.cpp:268-297)."""
    sc, _, _ = small_scene(n=64, seed=8)
    o, d = ray_set(4000, seed=4)
    ref = sc.intersect(o, d, bound="both", slab_max_level=1, stats=True)
    got = sc.intersect(o, d, bound=bound, slab_max_level=k, stats=True)
    assert (np.isfinite(ref["t"]) == np.isfinite(got["t"])).all()
    both = np.isfinite(ref["t"])
    assert np.allclose(got["t"][both], ref["t"][both], rtol=0, atol=1e-8)
    if bound == "box":
        assert (ref["stats"]["node_tests"] <= got["stats"]["node_tests"]).all()


def test_flat_map_is_a_plane():
    sc, _, _ = small_scene(flat=0.5, amp=0.1)
    o, d = ray_set(2000, seed=13)
    res = sc.intersect(o, d, threads=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        tp = (0.05 - o[:, 2]) / d[:, 2]
    p = o + tp[:, None] * d
    inside = (tp > 0) & (p[:, 0] >= 0) & (p[:, 0] <= 1) & (p[:, 1] >= 0) & (p[:, 1] <= 1)
    edge = np.minimum.reduce([p[:, 0], 1 - p[:, 0], p[:, 1], 1 - p[:, 1], np.abs(p[:, 0] - p[:, 1])])
    clear = np.abs(edge) > 1e-9
    assert (np.isfinite(res["t"])[clear] == inside[clear]).all()
    m = inside & clear
    assert np.allclose(res["t"][m], tp[m], rtol=1e-9)
    assert np.allclose(res["ng"][m], [0, 0, 1], atol=1e-12)


def test_occluded_equals_closest_hit_exists():
    sc, _, _ = small_scene(n=32)
    o, d = ray_set(3000, seed=17)
    res = sc.intersect(o, d)
    occ = sc.occluded(o, d)
    assert (occ == np.isfinite(res["t"])).all()
    tmax = np.where(np.isfinite(res["t"]), res["t"] * 0.999, 1e9)
    occ2 = sc.occluded(o, d, tmax=tmax)
    # With t_max short of the first hit, only a hit nearer than that occludes: none.
    assert not occ2[np.isfinite(res["t"])].any()


def test_node_bounds_are_conservative():
    """validate_node_bounds.cpp:124-170: dense surface points lie in every
    node's box; rays through them have t inside box, slab and both."""
    sc, V, tris = small_scene(n=16, seed=2)
    rng = np.random.default_rng(1)
    k = 0
    checks = 0
    for level in range(5):
        m = 16 >> level
        for (i, j) in rng.integers(0, m, size=(6, 2)):
            nb = sc.node_bound(k, level, int(i), int(j))
            w = 2.0 ** level / 16
            g = (np.arange(9) + 0.5) / 9
            ss, tt = np.meshgrid(i * w + g * w, j * w + g * w)
            st = np.column_stack([ss.ravel(), tt.ravel()])
            inside = st[:, 0] >= st[:, 1]  # triangle 0 of the quad: s >= t
            st = st[inside]
            if len(st) == 0:
                continue
            S, _ = sc.surface_points(np.zeros(len(st), np.int32), st)
            # tangent coordinates of the flat unit quad are (x, y, z) itself
            lo, hi = nb["box"]
            assert (S >= lo - 1e-12).all() and (S <= hi + 1e-12).all()
            for p in S[::7]:
                dvec = rng.normal(size=3)
                dvec /= np.linalg.norm(dvec)
                origin = p - 0.7 * dvec
                for mode in ("box", "slab", "both"):
                    ok, t01 = sc.node_interval(k, level, int(i), int(j), origin, dvec, 0.0, np.inf, mode)
                    assert ok and t01[0] - 1e-9 <= 0.7 <= t01[1] + 1e-9, (level, i, j, mode, t01)
                    checks += 1
    assert checks > 100


def test_triangle_aabb_contains_surface():
    sc, _, tris = small_scene(n=16, seed=4)
    rng = np.random.default_rng(5)
    for k in range(2):
        info = sc.triangle_info(k)
        u = np.sqrt(rng.random(2000))
        b = np.column_stack([1 - u, u * rng.random(2000)])
        t = tris[k, 18:24].reshape(3, 2)
        st = t[0] + b[:, :1] * (t[1] - t[0]) + b[:, 1:2] * (t[2] - t[0])
        S, _ = sc.surface_points(np.full(len(st), k, np.int32), st)
        assert (S >= info["aabb"][0] - 1e-12).all() and (S <= info["aabb"][1] + 1e-12).all()
        assert info["root_level"] == 3 and len(info["roots"]) == 4


def test_classify_square_against_sampling():
    rng = np.random.default_rng(6)
    tri = np.array([0.1, 0.2, 0.8, 0.3, 0.3, 0.9])
    for _ in range(300):
        c = rng.random(2)
        s = rng.uniform(0.01, 0.2)
        cls = oracle.classify_square(tri, c[0], c[1], s)
        g = np.linspace(-s, s, 41)
        X, Y = np.meshgrid(c[0] + g, c[1] + g)
        P = np.column_stack([X.ravel(), Y.ravel()])
        a, b, cc = tri[0:2], tri[2:4], tri[4:6]

        def side(p, q, r):
            return (q[0] - p[0]) * (r[:, 1] - p[1]) - (q[1] - p[1]) * (r[:, 0] - p[0])
        ins = (side(a, b, P) >= 0) & (side(b, cc, P) >= 0) & (side(cc, a, P) >= 0)
        if cls == 0:
            assert not ins.any()
        elif cls == 1:
            assert ins.all()


def test_root_level_grows_with_the_uv_span():
    """footprint_roots (descent_sampler.cpp:66-83) over a repeated tile: the
    inputs of the GPU root-level tests sit at levels 15 and 16, and the
    oracle (whose walk has no level limit, int64 cells) still answers at 16
    with hits on the surface."""
    tex = np.random.default_rng(3).random((1024, 1024)).astype(np.float32)
    nodes = W.close_tile(tex).astype(np.float64)
    from tests import scenes
    for (lo, hi), level in (((10, 45), 15), ((10, 80), 16)):
        tris = scenes.span_quad(lo, hi)
        sc = oracle.Scene(nodes, 2e-5, 0.0, True, tris)
        assert [sc.triangle_info(k)["root_level"] for k in range(2)] == [level, level]
        o = np.column_stack([np.linspace(0.1, 0.9, 50), np.linspace(0.2, 0.8, 50), np.full(50, 0.5)])
        d = np.tile([0.1, -0.05, -1.0], (50, 1))
        res = sc.intersect(o, d, threads=1)
        assert np.isfinite(res["t"]).all()
        S, _ = sc.surface_points(res["prim"], res["st"])
        assert np.linalg.norm(S - (o + res["t"][:, None] * d), axis=1).max() <= 2 * sc.triangle_info(0)["residual_tolerance"]


def test_work_counters_per_level_and_after_hit():
    """TraversalStats (displaced_intersector.h:89-105): node tests per level
    sum to the node tests; `both` at every slab level k tests no more nodes
    than `box` on every ray at every level (test_traversal_cost.cpp:268-297,
    check (1)); tests after the final hit are counted in walks that hit only,
    never exceed the totals, and exclude the leaf that made the hit."""
    sc, _, _ = small_scene(n=64, seed=8)
    o, d = ray_set(4000, seed=4)
    box = sc.intersect(o, d, bound="box", stats=True)
    bs = box["stats"]
    assert (bs["node_tests_per_level"].sum(axis=1) == bs["node_tests"]).all()
    assert bs["node_tests_per_level"][:, 7:].sum() == 0  # a 64^2 map has levels 0..6
    for k in range(0, 7):
        both = sc.intersect(o, d, bound="both", slab_max_level=k, stats=True)
        assert (both["stats"]["node_tests_per_level"] <= bs["node_tests_per_level"]).all(), k
    hit = np.isfinite(box["t"])
    assert hit.sum() > 1000 and (~hit).sum() > 100
    for name, total in (("node_tests_after_hit", "node_tests"), ("leaf_tests_after_hit", "leaf_tests")):
        after = bs[name]
        assert (after[~hit] == 0).all()
        assert (after <= bs[total]).all()
    assert (bs["leaf_tests_after_hit"][hit] <= bs["leaf_tests"][hit] - 1).all()
    # Some walks go on after their final hit (the child order is not front to back).
    assert bs["node_tests_after_hit"][hit].sum() > 0


def _box_clip_length(lo, hi, o, d, t_min):
    """The length of a ray's stretch inside the box [lo, hi] from t_min on
    (-1 when it misses), by the slab method in numpy."""
    t0, t1 = t_min, np.inf
    for a in range(3):
        if d[a] != 0.0:
            ta, tb = sorted(((lo[a] - o[a]) / d[a], (hi[a] - o[a]) / d[a]))
            t0, t1 = max(t0, ta), min(t1, tb)
        elif not lo[a] <= o[a] <= hi[a]:
            return -1.0
    if t0 > t1 + 1e-9 * max(1.0, abs(t0), abs(t1)):
        return -1.0
    return abs(t1 - t0)


@pytest.mark.parametrize("kind", ["flat", "curved"])
def test_node_records_follow_the_walk(kind):
    """NodeTestRecord (displaced_intersector.h:77-87, record_node_test at
    displaced_intersector.cpp:159-172) along the box walk, the walk the
    reference records (test_traversal_cost.cpp:377-386): one record per node
    test, per ray and per level as TraversalStats counts them, each ray's in
    seq order; the box's decision is the walk's (a passing cell with
    children is followed by one of its children, a failing one by a cell
    no finer); a test that passes with the walk's t_max passes with
    t_max = inf; both's interval lies inside the box's and the slab's; and
    the box length equals a numpy slab clip of the node's box with the ray
    in the triangle's tangent frame."""
    from tests import scenes
    if kind == "flat":
        sc, _, tris = small_scene(n=32, seed=3)
    else:
        tex = np.random.default_rng(4).random((32, 32)).astype(np.float32)
        tris = scenes.curved_quad()
        sc = oracle.Scene(W.close_tile(tex).astype(np.float64), 0.06, 0.0, True, tris)
    o, d = ray_set(1500, seed=13)
    res = sc.intersect(o, d, bound="box", stats=True, threads=1)
    rec = sc.records(o, d, bound="box")
    st = res["stats"]
    assert len(rec) == st["node_tests"].sum() > 10000
    assert (np.bincount(rec["ray"], minlength=len(o)) == st["node_tests"]).all()
    per_level = np.zeros_like(st["node_tests_per_level"])
    np.add.at(per_level, (rec["ray"], rec["level"]), 1)
    assert (per_level == st["node_tests_per_level"]).all()
    # Ray by ray in walk order, seq counting each ray's tests from 0.
    assert (np.diff(rec["ray"]) >= 0).all()
    starts = np.concatenate([[0], np.cumsum(st["node_tests"])[:-1]])
    assert (rec["seq"] == np.arange(len(rec)) - np.repeat(starts, st["node_tests"])).all()
    # The box's decision is the walk's.
    same = (rec["ray"][1:] == rec["ray"][:-1]) & (rec["prim"][1:] == rec["prim"][:-1])
    down = (rec["box_pass"][:-1] == 1) & (rec["level"][:-1] > 0)
    assert same.sum() > 5000 and (down & same).sum() > 1000
    assert (rec["level"][1:][down & same] == rec["level"][:-1][down & same] - 1).all()
    assert (rec["level"][1:][~down & same] >= rec["level"][:-1][~down & same]).all()
    # With t_max = inf a test passes whenever it passed with the walk's t_max.
    assert (rec["box_length"][rec["box_pass"] == 1] >= 0).all()
    assert (rec["slab_length"][rec["slab_pass"] == 1] >= 0).all()
    both = rec["both_length"] >= 0
    assert both.sum() > 1000
    assert (rec["box_length"][both] >= 0).all() and (rec["slab_length"][both] >= 0).all()
    assert (rec["both_length"][both] <= np.minimum(rec["box_length"], rec["slab_length"])[both] * (1 + 1e-9) + 1e-9).all()
    # The slab shortens some intervals and prunes some nodes the box keeps.
    assert ((rec["box_pass"] == 1) & (rec["slab_pass"] == 0)).sum() > 0
    assert (rec["slab_length"][both] < rec["box_length"][both]).mean() > 0.1
    # The box length against an independent clip of the node's box.
    rng = np.random.default_rng(2)
    for q in rng.choice(len(rec), size=400, replace=False):
        r = rec[q]
        q0, M = sc.tangent_frame(int(r["prim"]))
        ot, dt = M @ (o[r["ray"]] - q0), M @ d[r["ray"]]
        lo, hi = sc.node_bound(int(r["prim"]), int(r["level"]), int(r["i"]), int(r["j"]))["box"]
        want = _box_clip_length(lo, hi, ot, dt, 0.0)
        if want < 0:
            assert r["box_length"] < 0
        else:
            assert abs(r["box_length"] - want) <= 1e-9 * (1 + want), (r, want)
