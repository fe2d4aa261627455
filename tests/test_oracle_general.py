"""The oracle's general-normal and general-chart branch, pinned on the CPU.

The flat +z quads of the benchmark workloads give the normal's affine form no
deviation, so the Min-Range normalize remainder, the curved node bounds and
the curved Newton solve are idle there. These tests run the reference's own
validation checks on objects where all of them work:

  * curved:  the unit quad with outward-leaning interpolated vertex normals;
  * rotated: the unit quad whose corners map to a rotated, scaled square of
             the tile (a general chart J);
  * tilted:  both of the above, with the quad's positions and normals rotated
             and moved in 3-D (a general tangent frame).

Independent of the oracle's chart algebra (displaced_surface.cpp:54-91), the
tests restate the surface from its definition: barycentric weights of the
texture triangle, the interpolated position and normal, the bilinear height
(displaced_surface.cpp:29-47, 99-108, 162-166). Checks, after the reference:

  * Min-Range normalize encloses every normalized point of its input
    (affine_form.cpp:64-95), by dense sampling;
  * dense surface points lie in every node's box, satisfy the slab identity,
    and rays through them have t inside box, slab and both
    (validate_node_bounds.cpp:124-170);
  * rays aimed at surface points never miss and never hit late
    (validate_traversal.cpp:392-461);
  * at every Newton hit |S(st) - (o + t d)| <= 2 residual tolerances, st in
    the leaf texel and the domain, and ng the unit Su x Sv
    (validate_traversal.cpp:210-240);
  * the Newton hits agree with a texel-aligned fine mesh within twice its
    chord error (validate_traversal.cpp:242-318, chord errors
    traversal_test_util.cpp:170-192);
  * every bound mode gives the same hits (validate_traversal.cpp:320-390).
"""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests import scenes

N_TEX = 16
AMP = 0.08
KINDS = ("curved", "rotated", "tilted")


def _rotation(axis, angle):
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(angle) * K + (1 - np.cos(angle)) * K @ K


def tilted_quad():
    tris = scenes.rotated_chart_quad()
    tris[:, 9:18] = scenes.curved_quad()[:, 9:18]
    R = _rotation((1.0, 2.0, 0.5), 0.7)
    shift = np.array([0.3, -0.2, 0.5])
    for row in tris:
        for v in range(3):
            row[3 * v:3 * v + 3] = R @ row[3 * v:3 * v + 3] + shift
            row[9 + 3 * v:12 + 3 * v] = R @ row[9 + 3 * v:12 + 3 * v]
    return tris


def triangles(kind):
    return {"curved": scenes.curved_quad, "rotated": scenes.rotated_chart_quad, "tilted": tilted_quad}[kind]()


def make(kind, n=N_TEX, seed=5, amp=AMP, smooth=False):
    rng = np.random.default_rng(seed)
    if smooth:
        x = (np.arange(n) + 0.5) / n
        X, Y = np.meshgrid(x, x)
        tex = 0.5 + 0.3 * np.sin(2 * np.pi * X + 1.0) * np.cos(2 * np.pi * Y) + 0.05 * rng.random((n, n))
    else:
        tex = rng.random((n, n))
    nodes = W.close_tile(tex.astype(np.float32), True).astype(np.float64)
    tris = triangles(kind)
    return oracle.Scene(nodes, amp, 0.0, True, tris), nodes, tris


# ------------------------------------------------ the surface from its definition

def _bary(tri, st):
    """Barycentric weights (b1, b2) of tile points st (n, 2) in the texture triangle."""
    t = tri[18:24].reshape(3, 2)
    T = np.column_stack([t[1] - t[0], t[2] - t[0]])
    return np.linalg.solve(T, (st - t[0]).T).T


def height(nodes, amp, st):
    """HeightGrid::h under repeat: bilinear in the closed node grid, values[j, i] at (i/(W-1), j/(W-1))."""
    Wn = nodes.shape[0]
    x = (st[:, 0] - np.floor(st[:, 0])) * (Wn - 1)
    y = (st[:, 1] - np.floor(st[:, 1])) * (Wn - 1)
    i = np.minimum(np.floor(x).astype(np.int64), Wn - 2)
    j = np.minimum(np.floor(y).astype(np.int64), Wn - 2)
    a, b = x - i, y - j
    v = (1 - a) * (1 - b) * nodes[j, i] + a * (1 - b) * nodes[j, i + 1] + (1 - a) * b * nodes[j + 1, i] \
        + a * b * nodes[j + 1, i + 1]
    return amp * v


def surface(tri, nodes, amp, st):
    """S(s, t) = P + h N_hat with P and N = M / |M| interpolated barycentrically."""
    b = _bary(tri, st)
    p = tri[0:9].reshape(3, 3)
    m = tri[9:18].reshape(3, 3)
    P = p[0] + b[:, :1] * (p[1] - p[0]) + b[:, 1:2] * (p[2] - p[0])
    M = m[0] + b[:, :1] * (m[1] - m[0]) + b[:, 1:2] * (m[2] - m[0])
    Nh = M / np.linalg.norm(M, axis=1, keepdims=True)
    return P + height(nodes, amp, st)[:, None] * Nh, Nh


def chart_frame(tri):
    """The chart's q0 = P(0, 0), e1 = dP/ds, e2 = dP/dt and the tangent frame [e1 e2 ng]^-1."""
    t = tri[18:24].reshape(3, 2)
    p = tri[0:9].reshape(3, 3)
    T = np.column_stack([t[1] - t[0], t[2] - t[0]])
    E = np.column_stack([p[1] - p[0], p[2] - p[0]]) @ np.linalg.inv(T)
    e1, e2 = E[:, 0], E[:, 1]
    q0 = p[0] - E @ t[0]
    ng = np.cross(e1, e2)
    ng /= np.linalg.norm(ng)
    return q0, np.linalg.inv(np.column_stack([e1, e2, ng])), ng


def inside_domain(tri, st, eps=0.0):
    b = _bary(tri, st)
    return (b[:, 0] >= eps) & (b[:, 1] >= eps) & (1 - b[:, 0] - b[:, 1] >= eps)


def sample_points(tri, n, rng):
    u = np.sqrt(rng.random(n))
    b = np.column_stack([1 - u, u * rng.random(n)])
    t = tri[18:24].reshape(3, 2)
    return t[0] + b[:, :1] * (t[1] - t[0]) + b[:, 1:2] * (t[2] - t[0])


def mean_edge(tri):
    p = tri[0:9].reshape(3, 3)
    return (np.linalg.norm(p[1] - p[0]) + np.linalg.norm(p[2] - p[1]) + np.linalg.norm(p[0] - p[2])) / 3


def object_rays(tris, nodes, amp, n, seed):
    """Half camera-like rays from above the object towards random surface
    points, half from near the surface in random directions (the two
    distributions of traversal_test_util.cpp:83-111), in the object's frame."""
    rng = np.random.default_rng(seed)
    m = n // 2
    k = rng.integers(0, len(tris), n)
    S = np.empty((n, 3))
    G = np.empty((n, 3))
    for q in range(len(tris)):
        sel = k == q
        st = sample_points(tris[q], sel.sum(), rng)
        S[sel], _ = surface(tris[q], nodes, amp, st)
        G[sel] = chart_frame(tris[q])[2]
    lat = rng.normal(size=(n, 3)) * 0.5
    o1 = S[:m] + rng.uniform(0.4, 1.5, (m, 1)) * G[:m] + lat[:m]
    d1 = S[:m] - o1
    o2 = S[m:] + rng.uniform(-0.08, 0.08, (n - m, 1)) * G[m:]
    d2 = rng.normal(size=(n - m, 3))
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    return np.vstack([o1, o2]), np.vstack([d1, d2])


@pytest.mark.parametrize("kind", KINDS)
def test_surface_and_frame_match_their_definition(kind):
    """displaced_position, the unit Su x Sv normal and tangent_frame against
    the barycentric restatement; the normal by central differences of S."""
    sc, nodes, tris = make(kind)
    rng = np.random.default_rng(11)
    for k in range(len(tris)):
        st = sample_points(tris[k], 400, rng)
        # Keep away from texel edges, where S is only C0.
        f = (st * N_TEX) % 1.0
        st = st[(np.minimum(f, 1 - f) > 0.02).all(axis=1)]
        S_or, n_or = sc.surface_points(np.full(len(st), k, np.int32), st)
        S_def, _ = surface(tris[k], nodes, AMP, st)
        assert np.abs(S_or - S_def).max() < 1e-13
        hs = 1e-6
        dS_ds = (surface(tris[k], nodes, AMP, st + [hs, 0])[0] - surface(tris[k], nodes, AMP, st - [hs, 0])[0]) / (2 * hs)
        dS_dt = (surface(tris[k], nodes, AMP, st + [0, hs])[0] - surface(tris[k], nodes, AMP, st - [0, hs])[0]) / (2 * hs)
        nf = np.cross(dS_ds, dS_dt)
        nf /= np.linalg.norm(nf, axis=1, keepdims=True)
        assert np.abs(n_or - nf).max() < 1e-7
        q0, Tt, _ = chart_frame(tris[k])
        o_or, T_or = sc.tangent_frame(k)
        assert np.allclose(o_or, q0, atol=1e-13) and np.allclose(T_or, Tt, atol=1e-12)


def _true_unit(form, eu, ev, ek):
    """Points c + a eu + b ev + k ek of the vector form (3, 4), normalized."""
    c, a, b, k = form[:, 0], form[:, 1], form[:, 2], form[:, 3]
    X = c[None] + a[None] * eu[:, None] + b[None] * ev[:, None] + k[None] * ek
    return X / np.linalg.norm(X, axis=1, keepdims=True)


def _check_enclosure(form, rng, g=np.linspace(-1, 1, 9)):
    """Every normalized point of `form` lies in normalize(form) at the same
    noise symbols; returns the number of points checked."""
    out = oracle.affine_normalize(form)
    assert np.isfinite(out).all() and (out[:, 3] >= 0).all()
    EU, EV = np.meshgrid(g, g)
    eu, ev = EU.ravel(), EV.ravel()
    corners = np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).reshape(3, -1).T
    E = np.vstack([corners, rng.uniform(-1, 1, (20, 3))]) if form[:, 3].any() else np.zeros((1, 3))
    lo = out[:, 0][None] + out[:, 1][None] * eu[:, None] + out[:, 2][None] * ev[:, None] - out[:, 3][None]
    hi = lo + 2 * out[:, 3][None]
    n = 0
    for ek in E:
        X = _true_unit(form, eu, ev, np.broadcast_to(ek, (len(eu), 3)))
        ok = np.isfinite(X).all(axis=1)
        assert ((X[ok] >= lo[ok] - 1e-12) & (X[ok] <= hi[ok] + 1e-12)).all(), (form, out)
        n += ok.sum()
    return n


def _random_form(rng, spread_max=0.35, remainder=False, positive=False):
    c = rng.normal(size=3)
    if positive:
        c = np.abs(c)
    c *= rng.uniform(0.5, 2.0) / np.linalg.norm(c)
    spread = rng.uniform(0.01, spread_max) * np.linalg.norm(c)
    a, b = rng.normal(size=3) * spread, rng.normal(size=3) * spread
    k = rng.uniform(0, 1, 3) * spread if remainder else np.zeros(3)
    return np.column_stack([c, a, b, k])


def test_affine_normalize_encloses_dense_samples():
    """Min-Range normalize (affine_form.cpp:64-95): for every value of the
    noise symbols (and every point of the remainder boxes), the normalized
    vector lies in the output form at the same symbols.

    Every call site of the reference (node_bounds.cpp:102, :265,
    displaced_intersector.cpp:251) normalizes a form with zero remainders,
    and those are checked densely. With remainders, sq_length bounds the
    cross term 2 sum_q c_q k_q e_q by 2 |c . k| (affine_form.cpp:83), which
    is its maximum only when every c_q k_q has one sign: those forms are
    checked too, and one mixed-sign form is kept below as a recorded
    counterexample, so that a change to the restated bound is noticed."""
    rng = np.random.default_rng(2027)
    n_checked = 0
    for _ in range(300):
        n_checked += _check_enclosure(_random_form(rng), rng)
    for _ in range(100):
        n_checked += _check_enclosure(_random_form(rng, remainder=True, positive=True), rng)
    assert n_checked > 300000
    # The counterexample: c = (1, -1, 0.2), remainders on x and y only.
    form = np.array([[1.0, 0.05, 0.0, 0.3], [-1.0, 0.0, 0.05, 0.3], [0.2, 0.0, 0.0, 0.0]])
    out = oracle.affine_normalize(form)
    X = _true_unit(form, np.zeros(1), np.zeros(1), np.array([[1.0, -1.0, 0.0]]))[0]
    assert (X > out[:, 0] + out[:, 3]).any() or (X < out[:, 0] - out[:, 3]).any()
    # A form whose squared length reaches 0 falls back to [-1, 1]^3.
    out = oracle.affine_normalize(np.array([[0.1, 0.2, 0, 0], [0, 0, 0.1, 0], [0, 0, 0, 0.05]]))
    assert np.array_equal(out, np.array([[0.0, 0, 0, 1]] * 3))


@pytest.mark.parametrize("kind", KINDS)
def test_node_bounds_contain_the_surface(kind):
    """validate_node_bounds.cpp:124-170 on every node of every level over each
    triangle's footprint: surface points in the box (tangent frame), the slab
    identity F = m_z rho with |rho| <= r, rays through the points inside the
    box, slab and both intervals."""
    sc, nodes, tris = make(kind)
    rng = np.random.default_rng(2027)
    tol = 1e-9
    n_points = n_rays = n_cells = 0
    for k in range(len(tris)):
        info = sc.triangle_info(k)
        q0, Tt, _ = chart_frame(tris[k])
        scale = mean_edge(tris[k])
        t = tris[k][18:24].reshape(3, 2)
        dlo, dhi = t.min(axis=0), t.max(axis=0)
        for level in range(info["root_level"], -1, -1):
            w = 2.0 ** level / N_TEX
            i_lo, i_hi = int(np.floor(dlo[0] / w)), int(np.ceil(dhi[0] / w)) - 1
            j_lo, j_hi = int(np.floor(dlo[1] / w)), int(np.ceil(dhi[1] / w)) - 1
            for j in range(j_lo, j_hi + 1):
                for i in range(i_lo, i_hi + 1):
                    if oracle.classify_square(tris[k][18:24], (i + 0.5) * w, (j + 0.5) * w, 0.5 * w) == 0:
                        continue
                    st = np.column_stack([(i + rng.random(128)) * w, (j + rng.random(128)) * w])
                    st = st[inside_domain(tris[k], st)][:16]
                    if len(st) == 0:
                        continue
                    n_cells += 1
                    nb = sc.node_bound(k, level, i, j)
                    lo, hi = nb["box"]
                    box_scale = np.linalg.norm(hi - lo) + scale
                    S, Nh = surface(tris[k], nodes, AMP, st)
                    p = (S - q0) @ Tt.T
                    assert (p >= lo - tol * box_scale).all() and (p <= hi + tol * box_scale).all(), \
                        (kind, k, level, i, j, p.min(0) - lo, hi - p.max(0))
                    node = nb["node"]
                    m = Nh @ Tt.T
                    F = p[:, 2] * (1 + node["gu"] * m[:, 0] + node["gv"] * m[:, 1]) - m[:, 2] * (
                        node["h0"] + node["gu"] * (p[:, 0] - nb["centre"][0]) + node["gv"] * (p[:, 1] - nb["centre"][1]))
                    assert (np.abs(F) <= np.abs(m[:, 2]) * node["r"] + tol * (node["r"] + np.abs(F) + box_scale)).all()
                    n_points += len(st)
                    for q in range(len(st)):
                        d = rng.normal(size=3)
                        d /= np.linalg.norm(d)
                        ts = scale * 10.0 ** (-1.5 + 2.0 * rng.random())
                        o = S[q] - ts * d
                        for mode in ("box", "slab", "both"):
                            ok, t01 = sc.node_interval(k, level, i, j, o, d, 0.0, 1e3 * ts, mode)
                            assert ok and t01[0] - tol * ts <= ts <= t01[1] + tol * ts, (kind, level, i, j, mode)
                        n_rays += 1
    assert n_cells > 100 and n_points > 2000 and n_rays == n_points


@pytest.mark.parametrize("kind", KINDS)
def test_never_miss_aimed_rays(kind):
    """validate_traversal.cpp:392-461: rays aimed at surface points hit at or before them."""
    sc, nodes, tris = make(kind, seed=7, smooth=True)
    rng = np.random.default_rng(3)
    n = 3000
    k = rng.integers(0, len(tris), n)
    S = np.empty((n, 3))
    for q in range(len(tris)):
        sel = k == q
        S[sel], _ = surface(tris[q], nodes, AMP, sample_points(tris[q], sel.sum(), rng))
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    ts = 1.14 * 10 ** (-1.5 + 2 * rng.random(n))
    res = sc.intersect(S - ts[:, None] * dirs, dirs, threads=1)
    cos = np.abs(np.sum(dirs * np.nan_to_num(res["ng"]), axis=1))
    tol = 1e-9 * (ts + 1.5) + 4 * 1.14e-10 / np.maximum(cos, 1e-3)
    assert np.isfinite(res["t"]).all()
    assert (res["t"] <= ts + tol).all()


@pytest.mark.parametrize("kind", KINDS)
def test_newton_hits_lie_on_the_surface(kind):
    """validate_traversal.cpp:210-240: residual |S(st) - (o + t d)| at most twice
    the residual tolerance, st inside the leaf texel and the triangle's
    domain, and ng the unit Su x Sv at st (texel interiors)."""
    sc, nodes, tris = make(kind)
    o, d = object_rays(tris, nodes, AMP, 3000, seed=19)
    res = sc.intersect(o, d, threads=1)
    hit = np.isfinite(res["t"])
    assert hit.sum() > 1500
    w = 1.0 / N_TEX
    for k in range(len(tris)):
        sel = hit & (res["prim"] == k)
        assert sel.sum() > 300
        st = res["st"][sel]
        S, _ = surface(tris[k], nodes, AMP, st)
        resid = np.linalg.norm(S - (o[sel] + res["t"][sel, None] * d[sel]), axis=1)
        assert resid.max() <= 2.0 * sc.triangle_info(k)["residual_tolerance"]
        tx = res["texel"][sel]
        assert ((st[:, 0] >= tx[:, 0] * w) & (st[:, 0] <= (tx[:, 0] + 1) * w)).all()
        assert ((st[:, 1] >= tx[:, 1] * w) & (st[:, 1] <= (tx[:, 1] + 1) * w)).all()
        assert inside_domain(tris[k], st, -1e-12).all()
        f = (st * N_TEX) % 1.0
        inner = (np.minimum(f, 1 - f) > 1e-3).all(axis=1)
        hs = 1e-7
        sti = st[inner]
        dS_ds = (surface(tris[k], nodes, AMP, sti + [hs, 0])[0] - surface(tris[k], nodes, AMP, sti - [hs, 0])[0])
        dS_dt = (surface(tris[k], nodes, AMP, sti + [0, hs])[0] - surface(tris[k], nodes, AMP, sti - [0, hs])[0])
        nf = np.cross(dS_ds, dS_dt)
        nf /= np.linalg.norm(nf, axis=1, keepdims=True)
        assert np.abs(res["ng"][sel][inner] - nf).max() < 1e-6


def _mt_nearest(o, d, A, B, C):
    """Moller-Trumbore over the triangles (A, B, C): nearest t > 0, the face
    and its barycentrics (inf, -1 on a miss)."""
    e1, e2 = B - A, C - A
    p = np.cross(d, e2)
    det = np.sum(e1 * p, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / det
        s = o - A
        b1 = np.sum(s * p, axis=1) * inv
        q = np.cross(s, e1)
        b2 = (q @ d) * inv
        t = np.sum(e2 * q, axis=1) * inv
    ok = (det != 0) & (b1 >= 0) & (b2 >= 0) & (b1 + b2 <= 1) & (t > 0)
    if not ok.any():
        return np.inf, -1, 0.0, 0.0
    f = np.flatnonzero(ok)[np.argmin(t[ok])]
    return t[f], f, b1[f], b2[f]


def _outline(tris):
    """The texture-space corners of the quad (the union of its two triangles' domains), in order."""
    pts = np.unique(np.vstack([t[18:24].reshape(3, 2) for t in tris]).round(12), axis=0)
    c = pts.mean(axis=0)
    return pts[np.argsort(np.arctan2(pts[:, 1] - c[1], pts[:, 0] - c[0]))]


def _boundary_distance(corners, st):
    dist = np.full(len(st), np.inf)
    for e in range(len(corners)):
        a, b = corners[e], corners[(e + 1) % len(corners)]
        ab = b - a
        u = np.clip(((st - a) @ ab) / (ab @ ab), 0, 1)
        dist = np.minimum(dist, np.linalg.norm(st - (a + u[:, None] * ab), axis=1))
    return dist


@pytest.mark.parametrize("kind", KINDS)
def test_newton_hits_match_the_fine_texel_mesh(kind):
    """validate_traversal.cpp:242-318: the Newton leaf against a texel-aligned
    mesh with 4 x 4 sub-cells per texel, the sub-cell corners on the surface.
    Where both hit, |dt| <= 2 eps / max(cos, 0.1) + 1e-5 (t + scale), eps twice
    the mesh's largest chord error at edge midpoints and face centroids
    (traversal_test_util.cpp:170-192). Rays whose nearer hit is grazing
    (cos < 0.1) or lies within the chord band of the origin, and rays with a
    hit near the boundary of the object's domain, where the mesh is cut to
    whole sub-cells, are set aside as the reference sets them aside; none may
    be left unexplained."""
    sub = 4
    sc, nodes, tris = make(kind, smooth=True)
    # The quad's two triangles share one smooth surface S(s, t) (positions and
    # normals are affine over the whole quad): one mesh over their union.
    f = N_TEX * sub
    sub_w = 1.0 / f
    g = np.arange(f + 1) * sub_w
    GS, GT = np.meshgrid(g, g)
    V = np.column_stack([GS.ravel(), GT.ravel()])

    def owner(st):
        own = np.full(len(st), -1)
        for k in range(len(tris)):
            own[(own < 0) & inside_domain(tris[k], st, -1e-12)] = k
        return own

    def surf(st):
        k = owner(st)
        out = np.full((len(st), 3), np.nan)
        for q in range(len(tris)):
            out[k == q] = surface(tris[q], nodes, AMP, st[k == q])[0]
        return out

    Sv = surf(V)
    faces = []
    for b in range(f):
        for a in range(f):
            q = [b * (f + 1) + a, b * (f + 1) + a + 1, (b + 1) * (f + 1) + a, (b + 1) * (f + 1) + a + 1]
            if np.isnan(Sv[q]).any():
                continue
            faces.append((q[0], q[1], q[2]))
            faces.append((q[1], q[3], q[2]))
    faces = np.array(faces)
    assert len(faces) > 1000
    A, B, C = Sv[faces[:, 0]], Sv[faces[:, 1]], Sv[faces[:, 2]]
    eps = 0.0
    for (x, y) in ((0, 1), (1, 2), (2, 0)):
        st = 0.5 * (V[faces[:, x]] + V[faces[:, y]])
        eps = max(eps, np.linalg.norm(surf(st) - 0.5 * (Sv[faces[:, x]] + Sv[faces[:, y]]), axis=1).max())
    st = (V[faces[:, 0]] + V[faces[:, 1]] + V[faces[:, 2]]) / 3
    eps = 2.0 * max(eps, np.linalg.norm(surf(st) - (A + B + C) / 3, axis=1).max())
    assert 0 < eps < 0.01

    corners = _outline(tris)
    o, d = object_rays(tris, nodes, AMP, 1500, seed=23)
    res = sc.intersect(o, d, threads=1)
    scale = 2.0
    agree = skipped = unexplained = 0
    for r in range(len(o)):
        ta = res["t"][r]
        tb, fb, b1, b2 = _mt_nearest(o[r], d[r], A, B, C)
        ha, hb = np.isfinite(ta), np.isfinite(tb)
        if not ha and not hb:
            agree += 1
            continue
        dl = np.linalg.norm(d[r])
        dn = d[r] / dl
        cos_a = abs(dn @ res["ng"][r]) if ha else 0.0
        if hb:
            nb = np.cross(B[fb] - A[fb], C[fb] - A[fb])
            cos_b = abs(dn @ nb) / np.linalg.norm(nb)
        else:
            cos_b = 0.0
        a_nearer = ha and (not hb or ta <= tb)
        cos_n, t_n = (cos_a, ta) if a_nearer else (cos_b, tb)
        if ha and hb and abs(ta - tb) * dl <= 2 * eps / max(cos_a, 0.1) + 1e-5 * (ta * dl + scale):
            agree += 1
            continue
        near_boundary = False
        if ha:
            near_boundary |= _boundary_distance(corners, res["st"][r][None])[0] < 2 * sub_w + 2 * eps / max(cos_a, 0.1)
        if hb:
            st_b = (1 - b1 - b2) * V[faces[fb, 0]] + b1 * V[faces[fb, 1]] + b2 * V[faces[fb, 2]]
            near_boundary |= _boundary_distance(corners, st_b[None])[0] < 2 * sub_w + 2 * eps / max(cos_b, 0.1)
        near_origin = t_n * dl < 2 * eps / max(cos_n, 0.1)
        if near_boundary or near_origin or cos_n < 0.1:
            skipped += 1
        else:
            unexplained += 1
    assert agree > 0.85 * len(o), (agree, skipped, unexplained)
    assert unexplained == 0, (agree, skipped, unexplained, eps)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("bound,k", [("box", 7), ("slab", 7), ("both", 0), ("both", 7)])
def test_bound_modes_give_identical_hits(kind, bound, k):
    """validate_traversal.cpp:320-390 on the general objects."""
    sc, nodes, tris = make(kind, n=32, seed=8)
    o, d = object_rays(tris, nodes, AMP, 2000, seed=4)
    ref = sc.intersect(o, d, bound="both", slab_max_level=1, stats=True)
    got = sc.intersect(o, d, bound=bound, slab_max_level=k, stats=True)
    assert (np.isfinite(ref["t"]) == np.isfinite(got["t"])).all()
    both = np.isfinite(ref["t"])
    assert both.sum() > 800
    assert np.allclose(got["t"][both], ref["t"][both], rtol=0, atol=1e-8)
    if bound == "box":
        assert (ref["stats"]["node_tests"] <= got["stats"]["node_tests"]).all()
