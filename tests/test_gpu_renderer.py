"""Renderer-facing calls: the hit record (dmap_hit_record) against the
oracle's restatement of DisplacedGeometry::surface_point + compute_intersection
at the same hits, and the batched filter-resume loop (intersect_filtered)
against its independent equivalents: accept-all = the plain closest hit,
reject-triangle-k = the object without triangle k, and a texel-parity filter
= the same loop run on the oracle."""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests import scenes as S

pytestmark = pytest.mark.gpu


def _setup(kind, torch, n_tex=64, res=96):
    import dmap_amd
    rng = np.random.default_rng(12)
    tex = rng.random((n_tex, n_tex)).astype(np.float32)
    tris = {"flat": W.quad_plane(2, 1.0), "curved": S.curved_quad(), "rotated": S.rotated_chart_quad()}[kind]
    tile = dmap_amd.DisplacementTile(tex, 0.05)
    obj = dmap_amd.ObjectIntersector(tile, tris)
    o, d = W.pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 40.0, res, res)
    rays = torch.from_numpy(W.pack_rays(o, d)).cuda()
    sc = oracle.Scene(W.close_tile(tex).astype(np.float64), 0.05, 0.0, True, tris)
    return dmap_amd, tile, obj, rays, sc, tris, o, d


@pytest.mark.parametrize("kind", ["flat", "curved", "rotated"])
@pytest.mark.parametrize("twosided", [False, True])
def test_hit_record_matches_oracle(torch_gpu, kind, twosided):
    torch = torch_gpu
    dmap_amd, tile, obj, rays, sc, tris, o, d = _setup(kind, torch)
    hits = obj.intersect(rays)
    uvmap = ((2.0, 0.5), (0.1, -0.2), True)
    rec = obj.hit_record(rays, hits, uv_scale=uvmap[0], uv_offset=uvmap[1], flip_v=uvmap[2], twosided=twosided)
    R = {k: v.cpu().numpy() for k, v in rec.items()}
    prim = hits["prim"].cpu().numpy()
    bary = hits["bary"].cpu().numpy().astype(np.float64)
    hit = np.flatnonzero(prim >= 0)
    assert len(hit) > 1000
    miss = np.flatnonzero(prim < 0)
    assert all(np.all(R[k][:, miss] == 0) for k in R)
    for r in hit[:: max(1, len(hit) // 600)]:
        ref = sc.hit_record(int(prim[r]), bary[0, r], bary[1, r], d[r].astype(np.float64), *uvmap, twosided=twosided)
        scale = 1 + np.abs(ref["p"]).max()
        assert np.allclose(R["p"][:, r], ref["p"], atol=2e-7 * scale)
        assert np.allclose(R["ng"][:, r], ref["ng"], atol=2e-6)
        for k in ("dpdu", "dpdv"):
            assert np.allclose(R[k][:, r], ref[k], rtol=2e-6, atol=2e-6 * np.abs(ref[k]).max())
        assert np.allclose(R["uv"][:, r], ref["uv"], atol=1e-6)
        assert np.allclose(R["t"][:, r], ref["t"], atol=2e-5) and np.allclose(R["b"][:, r], ref["b"], atol=2e-5)
    # The record's p is the hit point o + t d, to fp32 resolution.
    t = hits["t"].cpu().numpy()
    p_ray = o[hit].astype(np.float64) + t[hit, None] * d[hit].astype(np.float64)
    assert np.max(np.abs(p_ray - R["p"][:, hit].T)) < 5e-6


def test_filtered_accept_all_and_without_triangle(torch_gpu):
    torch = torch_gpu
    dmap_amd, tile, obj, rays, sc, tris, o, d = _setup("flat", torch, res=128)
    plain = obj.intersect(rays)
    allf = dmap_amd.intersect_filtered(obj, rays, lambda h: torch.ones_like(h["t"], dtype=torch.bool))
    for k in plain:
        assert torch.equal(plain[k], allf[k]), k
    k = 3
    keep = [q for q in range(len(tris)) if q != k]
    obj_wo = dmap_amd.ObjectIntersector(tile, tris[keep])
    ref = obj_wo.intersect(rays)
    filt = dmap_amd.intersect_filtered(obj, rays, lambda h: h["prim"] != k)
    assert bool(((plain["prim"] == k).sum()) > 100)
    # the object without k numbers its triangles 0..n-2
    remap = torch.tensor(keep + [-1], dtype=torch.int32, device="cuda")
    ref_prim = remap[torch.where(ref["prim"] >= 0, ref["prim"], torch.full_like(ref["prim"], len(keep)))]
    assert torch.equal(filt["prim"], ref_prim)
    for key in ("t", "st", "bary", "ng", "texel"):
        assert torch.equal(filt[key], ref[key]), key
    occ = dmap_amd.occluded_filtered(obj, rays, lambda h: h["prim"] != k)
    assert torch.equal(occ, torch.isfinite(ref["t"]))


def _resume_np(t):
    """dmap_amd.filter_resume in numpy fp32."""
    t = t.astype(np.float32)
    return np.maximum(t + np.abs(t) * np.float32(1e-5), np.nextafter(t, np.float32(np.inf)))


def test_filtered_texel_parity_matches_oracle_loop(torch_gpu):
    """The filter-resume loop on the GPU against the same loop on the oracle."""
    torch = torch_gpu
    dmap_amd, tile, obj, rays, sc, tris, o, d = _setup("flat", torch, n_tex=16, res=64)
    assert np.array_equal(_resume_np(np.array([2.0, -3.0, 0.0], np.float32)),
                          dmap_amd.filter_resume(torch.tensor([2.0, -3.0, 0.0], device="cuda")).cpu().numpy())
    # Grazing rays cross many texels: reject hits on texels with i + j even.
    rng = np.random.default_rng(4)
    n = 4096
    og = np.column_stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), np.full(n, 0.2)])
    tg = np.column_stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), np.zeros(n)])
    dg = tg - og
    dg[:, 2] *= 0.3
    t0 = 0.0
    rg = torch.from_numpy(W.pack_rays(og.astype(np.float32), dg.astype(np.float32))).cuda()
    accept = lambda h: (h["texel"][0] + h["texel"][1]) % 2 == 1  # noqa: E731
    filt = dmap_amd.intersect_filtered(obj, rg, accept)
    # The same loop on the oracle.
    of, df = og.astype(np.float32).astype(np.float64), dg.astype(np.float32).astype(np.float64)
    tmin = np.full(n, t0)
    t_ref = np.full(n, np.inf)
    texel_ref = np.zeros((n, 2), dtype=np.int64)
    todo = np.arange(n)
    for _ in range(64):
        if len(todo) == 0:
            break
        r = sc.intersect(of[todo], df[todo], tmin[todo])
        hit = np.isfinite(r["t"])
        ok = hit & ((r["texel"][:, 0] + r["texel"][:, 1]) % 2 == 1)
        t_ref[todo[ok]] = r["t"][ok]
        texel_ref[todo[ok]] = r["texel"][ok]
        again = hit & ~ok
        tmin[todo[again]] = _resume_np(r["t"][again])
        todo = todo[again]
    tg_ = filt["t"].cpu().numpy()
    hit_g, hit_r = np.isfinite(tg_), np.isfinite(t_ref)
    flips = int((hit_g != hit_r).sum())
    assert flips <= n * 1e-3, flips
    both = hit_g & hit_r
    assert both.sum() > 150
    same_texel = np.all(filt["texel"].cpu().numpy().T[both] == texel_ref[both], axis=1)
    assert same_texel.mean() > 0.999
    assert np.max(np.abs(tg_[both][same_texel] - t_ref[both][same_texel]) / np.abs(t_ref[both][same_texel])) < 1e-4

    # every accepted GPU hit satisfies the filter
    tex_g = filt["texel"].cpu().numpy()
    assert np.all(((tex_g[0] + tex_g[1]) % 2 == 1)[hit_g])


def test_no_hit_behind_the_origin(torch_gpu):
    """A negative t_min does not make crossings behind the origin count:
    newton_from rejects a root whose point lies behind the origin
    (displaced_intersector.cpp:356-357), so the reported t is >= 0 and a
    filter loop never resumes from a negative t. Rays that leave the surface
    downwards with t_min = -4 cross it only at t in (-4, 0): no hit, on the
    GPU and on the oracle; reversed, the same rays hit it."""
    torch = torch_gpu
    dmap_amd, tile, obj, rays, sc, tris, o, d = _setup("flat", torch, n_tex=16, res=64)
    rng = np.random.default_rng(4)
    n = 4096
    phi = rng.uniform(0, 2 * np.pi, n)
    og = np.column_stack([rng.uniform(0.1, 0.9, n), rng.uniform(0.1, 0.9, n), np.full(n, -0.01)]).astype(np.float32)
    dg = np.column_stack([0.2 * np.cos(phi), 0.2 * np.sin(phi), np.full(n, -0.06)]).astype(np.float32)
    for sign, expect_hits in ((1.0, False), (-1.0, True)):
        r = W.pack_rays(og, sign * dg)
        r[6] = -4.0
        g = obj.intersect(torch.from_numpy(r).cuda())
        ref = sc.intersect(og.astype(np.float64), sign * dg.astype(np.float64), np.full(n, -4.0))
        hit_g, hit_r = torch.isfinite(g["t"]).cpu().numpy(), np.isfinite(ref["t"])
        if expect_hits:
            assert hit_r.sum() > 0.9 * n
            assert (g["t"].cpu().numpy()[hit_g] >= 0).all()
            assert (hit_g == hit_r).sum() >= n - 2
        else:
            assert not hit_r.any() and not hit_g.any()


def _rotation(seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def _move(tris, R, T):
    """The triangle rows with positions R p + T and vertex normals R m."""
    out = tris.copy()
    for v in range(3):
        out[:, 3 * v:3 * v + 3] = tris[:, 3 * v:3 * v + 3] @ R.T + T
        out[:, 9 + 3 * v:12 + 3 * v] = tris[:, 9 + 3 * v:12 + 3 * v] @ R.T
    return out


def _host(rec):
    return {k: v.cpu().numpy() for k, v in rec.items()}


UVMAP = dict(uv_scale=(2.0, 0.5), uv_offset=(0.1, -0.2), flip_v=True)


@pytest.mark.parametrize("kind", ["flat", "curved", "rotated"])
@pytest.mark.parametrize("twosided", [False, True])
def test_world_record_of_a_rigid_instance(torch_gpu, kind, twosided):
    """dmap_hit_record_world. With the identity it is the object-space
    record (p, dpdu, dpdv, uv bit for bit; ng, t, b to 1e-7: ng is
    renormalized in double). With a rotation and translation (R, T) it
    equals the plain record of the same mesh moved to world space (R p + T,
    R m) at the same hits (a record depends on the hit's triangle and
    barycentrics only): p to 1e-6 of the scene's size, ng, t, b to 1e-5,
    dpdu, dpdv to 1e-5 relative, uv to 1e-6. Tracing the moved mesh with the
    world rays finds the same hits: hit/miss on 99.9% of the rays, the same
    triangle on the common hits and p to 2e-6 of the scene's size (the two
    traces see the rays rounded to fp32 in different frames)."""
    torch = torch_gpu
    dmap_amd, tile, obj, rays, sc, tris, o, d = _setup(kind, torch)
    hits = obj.intersect(rays)
    a = _host(obj.hit_record(rays, hits, twosided=twosided, **UVMAP))
    b = _host(obj.hit_record(rays, hits, twosided=twosided, object_to_world=np.hstack([np.eye(3), np.zeros((3, 1))]),
                             **UVMAP))
    for k in ("p", "dpdu", "dpdv", "uv"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("ng", "t", "b"):
        assert np.abs(a[k] - b[k]).max() <= 1e-7, k
    R, T = _rotation(5), np.array([0.3, -1.2, 2.0])
    moved = dmap_amd.ObjectIntersector(tile, _move(tris, R, T))
    o_w = (o.astype(np.float64) @ R.T + T).astype(np.float32)
    d_w = (d.astype(np.float64) @ R.T).astype(np.float32)
    rays_w = torch.from_numpy(W.pack_rays(o_w, d_w)).cuda()
    o_o = ((o_w.astype(np.float64) - T) @ R).astype(np.float32)
    d_o = (d_w.astype(np.float64) @ R).astype(np.float32)
    rays_o = torch.from_numpy(W.pack_rays(o_o, d_o)).cuda()
    h_i = obj.intersect(rays_o)
    xf = np.hstack([R, T[:, None]])
    ri = _host(obj.hit_record(rays_o, h_i, twosided=twosided, object_to_world=xf, **UVMAP))
    # The moved mesh's record at the instance's hits.
    rw = _host(moved.hit_record(rays_w, h_i, twosided=twosided, **UVMAP))
    pi = h_i["prim"].cpu().numpy()
    ok = np.flatnonzero(pi >= 0)
    assert len(ok) > 1000
    scale = 1 + np.abs(rw["p"][:, ok]).max()
    assert np.abs(ri["p"][:, ok] - rw["p"][:, ok]).max() <= 1e-6 * scale
    for k in ("ng", "t", "b"):
        assert np.abs(ri[k][:, ok] - rw[k][:, ok]).max() <= 1e-5, k
    for k in ("dpdu", "dpdv"):
        err = np.abs(ri[k][:, ok] - rw[k][:, ok]).max(axis=0) / np.abs(rw[k][:, ok]).max(axis=0)
        assert err.max() <= 1e-5, k
    assert np.abs(ri["uv"][:, ok] - rw["uv"][:, ok]).max() <= 1e-6
    # The moved mesh traced with the world rays.
    h_w = moved.intersect(rays_w)
    pw = h_w["prim"].cpu().numpy()
    assert ((pi >= 0) == (pw >= 0)).mean() > 0.999
    both = np.flatnonzero((pi >= 0) & (pw >= 0))
    assert (pi[both] == pw[both]).mean() > 0.999
    rt = _host(moved.hit_record(rays_w, h_w, twosided=twosided, **UVMAP))
    same = both[pi[both] == pw[both]]
    assert np.abs(ri["p"][:, same] - rt["p"][:, same]).max() <= 2e-6 * scale


def _camera_diffs(o, d, res, fov_deg, eps):
    """Offset rays of a pinhole camera's rays: the same origins, the
    directions moved by eps pixel spacings along the image's x and y."""
    fwd = d / np.linalg.norm(d, axis=1, keepdims=True)
    right = np.cross(fwd, [0.0, 0.0, 1.0])
    right /= np.linalg.norm(right, axis=1, keepdims=True)
    up = np.cross(right, fwd)
    step = 2 * np.tan(np.radians(fov_deg) / 2) / res * np.linalg.norm(d, axis=1, keepdims=True)
    rx_d = (d + eps * step * right).astype(np.float32)
    ry_d = (d + eps * step * up).astype(np.float32)
    return np.concatenate([o.T, rx_d.T, o.T, ry_d.T]).astype(np.float32), rx_d, ry_d


@pytest.mark.parametrize("instanced", [False, True])
def test_uv_partials_against_offset_rays(torch_gpu, instanced):
    """The uv partials on a displaced plane that is exactly flat (one
    quad, constant texels, so the tangent plane is the surface): the offset
    rays' own hits give (u, v) and p at px, py exactly, so du/dx, dv/dx,
    du/dy, dv/dy equal the hit uv differences and dp/dx, dp/dy the hit
    position differences, to fp32 rounding (1e-4 relative). Instanced: the
    same with a rotation and translation, the differentials in world space,
    against the world-space records."""
    import dmap_amd
    torch = torch_gpu
    tex = np.full((64, 64), 0.3, dtype=np.float32)
    tile = dmap_amd.DisplacementTile(tex, 0.05)
    tris = W.quad_plane(1, 1.0)
    obj = dmap_amd.ObjectIntersector(tile, tris)
    res = 64
    o, d = W.pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 40.0, res, res)
    o, d = o.astype(np.float64), d.astype(np.float64)
    diffs, rx_d, ry_d = _camera_diffs(o, d, res, 40.0, 0.5)
    R, T = (_rotation(9), np.array([1.0, 2.0, -0.5])) if instanced else (np.eye(3), np.zeros(3))
    xf = np.hstack([R, T[:, None]]) if instanced else None
    # Object-space rays of the camera; the differentials go in world space.
    rays = torch.from_numpy(W.pack_rays(o.astype(np.float32), d.astype(np.float32))).cuda()
    wd = diffs.astype(np.float64).reshape(4, 3, -1)
    wd[0] = (R @ wd[0]) + T[:, None]
    wd[2] = (R @ wd[2]) + T[:, None]
    wd[1] = R @ wd[1]
    wd[3] = R @ wd[3]
    D = torch.from_numpy(np.ascontiguousarray(wd.reshape(12, -1).astype(np.float32))).cuda()
    hits = obj.intersect(rays)
    rec = _host(obj.hit_record(rays, hits, diffs=D, object_to_world=xf, **UVMAP))
    out, offset_hit = {}, hits["prim"].cpu().numpy() >= 0
    for name, dd in (("x", rx_d), ("y", ry_d)):
        rr = torch.from_numpy(W.pack_rays(o.astype(np.float32), dd)).cuda()
        hh = obj.intersect(rr)
        offset_hit &= hh["prim"].cpu().numpy() >= 0
        out[name] = _host(obj.hit_record(rr, hh, object_to_world=xf, **UVMAP))
    # Rays whose offset rays hit the quad too (the plane's uv is one affine
    # map over both triangles).
    hit = np.flatnonzero(offset_hit)
    assert len(hit) > 1500
    for name, du, dv, dp in (("x", "dudx", "dvdx", "dpdx"), ("y", "dudy", "dvdy", "dpdy")):
        duv = out[name]["uv"][:, hit] - rec["uv"][:, hit]
        got = np.stack([rec[du][hit], rec[dv][hit]])
        assert np.abs(got - duv).max() <= 1e-4 * np.abs(duv).max() + 2e-6, name
        dpp = out[name]["p"][:, hit] - rec["p"][:, hit]
        assert np.abs(rec[dp][:, hit] - dpp).max() <= 1e-4 * np.abs(dpp).max() + 2e-6, name
    miss = np.flatnonzero(hits["prim"].cpu().numpy() < 0)
    assert all(np.all(rec[k][..., miss] == 0) for k in ("dudx", "dvdx", "dpdx"))
