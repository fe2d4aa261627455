"""The oracle's renderer hit record (DisplacedGeometry::surface_point and the
geometric frame of compute_intersection, displaced_geometry.cpp:126-183)
pinned by what fixes it: p is the displaced surface point, dpdu and dpdv are
central differences of p in the object's uv (through the uv map, flip_v
included), ng is the unit normal of dpdu x dpdv oriented along the base
face's winding, and (t, b, ng) is an orthonormal frame with b along
ng x dpdu; twosided faces ng towards the viewer."""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests import scenes as S


def _scene(kind):
    rng = np.random.default_rng(8)
    tex = rng.random((16, 16))
    tris = {"flat": W.quad_plane(1, 1.0), "curved": S.curved_quad(), "rotated": S.rotated_chart_quad()}[kind]
    return oracle.Scene(W.close_tile(tex, True), 0.05, 0.0, True, tris), tris


def _p_at(sc, tris, k, uo, vo, su, sv, ou, ov, flip):
    """p as a function of the object's uv: tile (s, t) = (uo su + ou, v' sv + ov), v' = 1 - vo under flip."""
    s = uo * su + ou
    t = ((1.0 - vo) if flip else vo) * sv + ov
    return sc.displaced_position(k, s, t)


@pytest.mark.parametrize("kind", ["flat", "curved", "rotated"])
@pytest.mark.parametrize("uvmap", [((1.0, 1.0), (0.0, 0.0), False), ((2.0, 0.5), (0.1, -0.2), True)])
def test_hit_record_pins(kind, uvmap):
    (su, sv), (ou, ov), flip = uvmap
    sc, tris = _scene(kind)
    rng = np.random.default_rng(5)
    for _ in range(40):
        k = int(rng.integers(len(tris)))
        u, v = rng.random(2)
        if u + v > 1:
            u, v = 1 - u, 1 - v
        d = rng.normal(size=3)
        rec = sc.hit_record(k, u, v, d, (su, sv), (ou, ov), flip, twosided=False)
        row = tris[k]
        t0, t1, t2 = row[18:20], row[20:22], row[22:24]
        st = t0 + u * (t1 - t0) + v * (t2 - t0)
        assert np.allclose(rec["p"], sc.displaced_position(k, st[0], st[1]), atol=1e-12)
        # uv: the object's coordinates of st
        uo = (st[0] - ou) / su
        vo = (st[1] - ov) / sv
        if flip:
            vo = 1.0 - vo
        assert np.allclose(rec["uv"], [uo, vo], atol=1e-12)
        # dpdu, dpdv: central differences of p in (uo, vo), away from texel edges
        fx, fy = (st * 16) % 1.0
        if min(fx, fy, 1 - fx, 1 - fy) > 1e-3:
            hh = 1e-7
            for key, du, dv in (("dpdu", hh, 0.0), ("dpdv", 0.0, hh)):
                fd = (_p_at(sc, tris, k, uo + du, vo + dv, su, sv, ou, ov, flip) -
                      _p_at(sc, tris, k, uo - du, vo - dv, su, sv, ou, ov, flip)) / (2 * hh)
                assert np.allclose(rec[key], fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max()), (key, rec[key], fd)
        # ng: unit, along dpdu x dpdv up to sign, with the winding
        wind = np.cross(row[3:6] - row[0:3], row[6:9] - row[0:3])
        c = np.cross(rec["dpdu"], rec["dpdv"])
        assert abs(np.linalg.norm(rec["ng"]) - 1) < 1e-6
        assert np.linalg.norm(np.cross(rec["ng"], c / np.linalg.norm(c))) < 1e-5
        assert np.dot(rec["ng"], wind) > 0
        # the frame
        F = np.stack([rec["t"], rec["b"], rec["ng"]])
        assert np.allclose(F @ F.T, np.eye(3), atol=1e-5)
        b = np.cross(rec["ng"], rec["dpdu"])
        assert np.allclose(rec["b"], b / np.linalg.norm(b), atol=1e-5)
        # twosided: ng towards -d, with dpdu and t turned as well
        two = sc.hit_record(k, u, v, d, (su, sv), (ou, ov), flip, twosided=True)
        assert np.dot(two["ng"], -d) >= 0
        sign = 1.0 if np.dot(rec["ng"], -d) >= 0 else -1.0
        assert np.allclose(two["ng"], sign * rec["ng"]) and np.allclose(two["dpdu"], sign * rec["dpdu"])
        assert np.allclose(two["t"], sign * rec["t"]) and np.allclose(two["b"], rec["b"])
