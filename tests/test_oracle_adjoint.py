"""The oracle's adjoint (implicit-function gradient of t and ng with respect to
the node values) pinned by central finite differences of the oracle's own
double forward, on a flat quad (constant normal) and on a quad with
interpolated normals. A node whose perturbation moves any hit to another
texel or flips a hit/miss is a non-differentiable event and is skipped."""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests import scenes as S

N_TEX = 16
AMP = 0.05
EPS = 1e-5


def _scene(nodes, tris):
    return oracle.Scene(nodes, AMP, 0.0, True, tris)


def _rays(n=576, seed=21):
    o, d = W.pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 30.0, 24, 24)
    return o.astype(np.float64)[:n], d.astype(np.float64)[:n]


def _loss(nodes, tris, o, d, lt, ln):
    sc = _scene(nodes, tris)
    _, t, ng = sc.adjoint(o, d, lt, ln)
    ref = sc.intersect(o, d)
    return t, ng, ref


@pytest.mark.parametrize("kind", ["flat", "curved"])
def test_oracle_adjoint_matches_central_differences(kind):
    rng = np.random.default_rng(7)
    tex = rng.random((N_TEX, N_TEX))
    nodes = W.close_tile(tex, True).astype(np.float64)
    tris = W.quad_plane(1, 1.0) if kind == "flat" else S.curved_quad()
    o, d = _rays()
    n = o.shape[0]
    lt = rng.normal(size=n)
    ln = rng.normal(size=(n, 3))
    grad, t0, ng0 = _scene(nodes, tris).adjoint(o, d, lt, ln)
    hit = np.isfinite(t0)
    assert hit.sum() > 300
    base = _scene(nodes, tris).intersect(o, d)
    # Nodes with gradient, checked one by one.
    cand = np.argwhere(np.abs(grad) > 1e-8)
    pick = cand[rng.permutation(len(cand))[:24]]
    checked = 0
    for j, i in pick:
        vals = []
        moved = False
        for sgn in (1.0, -1.0):
            pert = nodes.copy()
            pert[j, i] += sgn * EPS
            t, ng, ref = _loss(pert, tris, o, d, lt, ln)
            moved |= bool(np.any(np.isfinite(t) != hit)) or bool(np.any(ref["texel"][hit] != base["texel"][hit]))
            both = hit & np.isfinite(t)
            vals.append(np.sum(lt[both] * t[both]) + np.sum(ln[both] * ng[both]))
        if moved:
            continue
        fd = (vals[0] - vals[1]) / (2 * EPS)
        g = grad[j, i]
        assert abs(fd - g) <= 1e-7 * max(1.0, abs(g)), (kind, j, i, fd, g)
        checked += 1
    assert checked >= 12


def test_oracle_adjoint_polished_root_is_the_forward_hit():
    """The adjoint's double root agrees with the forward's accepted root."""
    rng = np.random.default_rng(3)
    nodes = W.close_tile(rng.random((N_TEX, N_TEX)), True)
    tris = W.quad_plane(1, 1.0)
    o, d = _rays()
    sc = _scene(nodes, tris)
    _, t, ng = sc.adjoint(o, d, np.ones(o.shape[0]))
    ref = sc.intersect(o, d)
    h = np.isfinite(ref["t"])
    assert np.array_equal(np.isfinite(t), h)
    assert np.max(np.abs(t[h] - ref["t"][h]) / ref["t"][h]) < 1e-9
    assert np.max(np.abs(ng[h] - ref["ng"][h])) < 1e-7


def test_oracle_adjoint_linear_in_the_upstream_gradient():
    """grad(a lt + b lt') = a grad(lt) + b grad(lt') (the adjoint is a linear map)."""
    rng = np.random.default_rng(4)
    nodes = W.close_tile(rng.random((N_TEX, N_TEX)), True)
    sc = _scene(nodes, S.curved_quad())
    o, d = _rays()
    n = o.shape[0]
    l1, l2 = rng.normal(size=n), rng.normal(size=n)
    m1, m2 = rng.normal(size=(n, 3)), rng.normal(size=(n, 3))
    g1 = sc.adjoint(o, d, l1, m1)[0]
    g2 = sc.adjoint(o, d, l2, m2)[0]
    g12 = sc.adjoint(o, d, 2 * l1 - 3 * l2, 2 * m1 - 3 * m2)[0]
    assert np.allclose(g12, 2 * g1 - 3 * g2, rtol=1e-10, atol=1e-12)
