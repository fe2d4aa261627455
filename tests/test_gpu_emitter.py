"""The emitter-pdf door (DisplacedAreaLight::pdf, displaced_area_light.cpp:
174-186) on the GPU: dmap_amd.emitter.light_pdf walks one light triangle
alone (prim_mask) and converts an area density to solid angle."""
import numpy as np
import pytest

from dmap_amd import workload as W

pytestmark = pytest.mark.gpu


def _obj(tex, strength, tris):
    import dmap_amd
    tile = dmap_amd.DisplacementTile(tex, strength, 0.0, True)
    return dmap_amd.ObjectIntersector(tile, tris)


def _shading(n, seed, targets_lo, targets_hi):
    """Shading points above the unit square and unit directions towards
    random targets on the plane z = 0 (some off the light triangle)."""
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(-0.2, 1.2, n), rng.uniform(-0.2, 1.2, n), rng.uniform(0.3, 1.5, n)])
    y = np.stack([rng.uniform(targets_lo, targets_hi, n), rng.uniform(targets_lo, targets_hi, n), np.zeros(n)])
    w = y - p
    w /= np.linalg.norm(w, axis=0)
    return p.astype(np.float32), w.astype(np.float32)


def _flat_triangle_pdf(tri, p, w):
    """Closed form on a flat triangle with a uniform area density 1 / area:
    the plane hit y = p + t w inside the triangle, density t^2 |w|^2 / area
    / |cos|; 0 when the direction misses. Also the distance of the hit to
    the triangle's nearest edge, in the plane."""
    a, b, c = tri[0:3], tri[3:6], tri[6:9]
    nrm = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(nrm)
    nu = nrm / np.linalg.norm(nrm)
    p, w = p.astype(np.float64), w.astype(np.float64)
    denom = nu @ w
    t = (nu @ (a[:, None] - p)) / denom
    y = p + t * w
    edge_dist = np.full(p.shape[1], np.inf)
    inside = t > 0
    for u, v in ((a, b), (b, c), (c, a)):
        e = v - u
        side = np.cross(e, y.T - u) @ nu  # > 0 on the inner side (counter-clockwise winding)
        inside &= side >= 0
        edge_dist = np.minimum(edge_dist, np.abs(side) / np.linalg.norm(e))
    cos = np.abs(denom) / np.linalg.norm(w, axis=0)
    pdf = np.where(inside, t * t * (w * w).sum(axis=0) / area / cos, 0.0)
    return pdf, inside, edge_dist, area


def test_flat_light_pdf_is_the_closed_form(torch_gpu):
    """On a flat light (a zero map) the door's hit is the plane hit: hit or
    miss as the closed form says (away from the triangle's edges), and the
    density t^2 / (area |cos|) within 1e-5 relative (fp32 rays)."""
    import torch

    from dmap_amd import emitter
    tris = W.quad_plane(1, 1.0)
    obj = _obj(np.zeros((64, 64), np.float32), 0.05, tris)
    p, w = _shading(20000, 11, -0.1, 1.1)
    for k in range(2):
        ref, inside, edge, area = _flat_triangle_pdf(tris[k, 0:9], p, w)
        pdf, st, hit = emitter.light_pdf(obj, k, torch.from_numpy(p), torch.from_numpy(w), 1.0 / area)
        pdf, hit = pdf.cpu().numpy(), hit.cpu().numpy()
        clear = edge > 1e-5
        assert (hit[clear] == inside[clear]).all(), k
        both = hit & inside & clear
        assert both.sum() > 5000
        rel = np.abs(pdf[both] - ref[both]) / ref[both]
        print(f"[emitter] flat triangle {k}: {int(both.sum())} hits, max rel err {rel.max():.2e}", flush=True)
        assert rel.max() < 1e-5
        assert (pdf[~hit] == 0).all()
        # a density that depends on the hit's tile coordinates is read there
        pdf2, st2, _ = emitter.light_pdf(obj, k, torch.from_numpy(p), torch.from_numpy(w),
                                          lambda s: (1.0 + s[:, 0]) / area)
        pdf2 = pdf2.cpu().numpy()
        s0 = st2.cpu().numpy()[:, 0]
        assert np.allclose(pdf2[both], ref[both] * (1.0 + s0[both]), rtol=1e-5)


def test_displaced_light_door_finds_the_triangle_walks_hit(torch_gpu):
    """On a displaced light the door (the light triangle's walk alone) meets
    the surface where the full closest hit does when that hit is on the
    light triangle (1e-6 relative: the full walk can reach the triangle with
    a t_max already narrowed by the other one, and the two fp32 searches
    then meet the same root, polished in double, from different starts),
    and never before the full hit otherwise; its density is the area
    density times dist^2 / |cos| at that hit."""
    import torch

    from dmap_amd import emitter
    tris = W.quad_plane(1, 1.0)
    tex = W.c1_small(n=128, width=8, height=8).texels
    obj = _obj(tex, 0.05, tris)
    p, w = _shading(20000, 12, 0.0, 1.0)
    rays = np.zeros((8, p.shape[1]), np.float32)
    rays[0:3], rays[3:6], rays[7] = p, w, np.inf
    full = obj.intersect(torch.from_numpy(rays).cuda())
    t_full = full["t"].cpu().numpy()
    prim_full = full["prim"].cpu().numpy()
    for k in range(2):
        pdf, st, hit = emitter.light_pdf(obj, k, torch.from_numpy(p), torch.from_numpy(w), 2.0)
        door = obj.intersect(torch.from_numpy(rays).cuda(), prim_mask=k)
        t_door = door["t"].cpu().numpy()
        mine = prim_full == k
        assert mine.sum() > 5000
        assert np.allclose(t_door[mine], t_full[mine], rtol=1e-6, atol=0), k
        fin = np.isfinite(t_full)
        assert (t_door[fin] >= t_full[fin] * (1 - 1e-6)).all()
        ng = door["ng"].cpu().numpy().astype(np.float64)
        w64 = w.astype(np.float64)
        wn2 = (w64 * w64).sum(axis=0)
        cos = np.abs((ng * w64).sum(axis=0)) / np.sqrt(wn2)
        ref = 2.0 * t_door.astype(np.float64) ** 2 * wn2 / np.where(cos > 0, cos, 1.0)
        got = pdf.cpu().numpy()
        ok = np.isfinite(t_door) & (cos >= emitter.MIN_COS_Y)
        assert ok.sum() > 5000 and np.allclose(got[ok], ref[ok], rtol=1e-10)
        assert (got[~np.isfinite(t_door)] == 0).all()
        print(f"[emitter] displaced triangle {k}: {int(mine.sum())} rays whose closest hit is on it, "
              f"{int(np.isfinite(t_door).sum())} door hits", flush=True)
