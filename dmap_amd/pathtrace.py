"""Wavefront path tracer over the batched intersector (config 3: path
tracing with GPU intersection).

The reference renders with ks's SmallPT through embree callbacks, one ray at
a time (path_trace.cpp:124-139, direct_lighting.cpp:185-232); ks is not part
of the reference repository. This tracer is the batched equivalent for the
configuration's scene: a Lambertian displaced surface, one rectangular area
light and a constant environment, paths of at most `depth` bounces.

Per bounce, the live paths go through one closest-hit batch and their
next-event rays through one any-hit batch (dmap_intersect_closest /
dmap_intersect_any); paths that end are dropped from the next batch. The
estimator:

* next-event estimation on the light: a uniform point on the rectangle, a
  shadow ray to it (t_max shortened by 1e-4, as ks's mesh lights do,
  direct_lighting.cpp:57), weighted by the power heuristic against the
  cosine-sampled BSDF;
* BSDF sampling (cosine-weighted, exact for the Lambertian lobe) continues
  the path; a sampled ray that meets the light before the surface adds the
  light with the complementary weight, one that escapes adds the
  environment (not light-sampled, weight 1);
* new rays start 1e-4 off the surface along the geometric normal, on the
  side they leave from.

Random numbers come from a counter-based hash of (pixel, sample, bounce,
dimension), so the same image comes out of any object with the intersector
calls below (closest, occluded), whatever computes them. The light does not
occlude (its back face is open).
"""
import math

import numpy as np

LIGHT_EPS = 1e-4
SPAWN_EPS = 1e-4


class Scene:
    """One displaced object, a one-sided rectangular light (corner, edges,
    radiance; emits along edge_u x edge_v) and a constant environment."""

    def __init__(self, intersector, albedo=0.7, env=0.2, light_corner=(3.0, 3.0, 3.0), light_u=(2.0, 0.0, 0.0),
                 light_v=(0.0, 2.0, 0.0), radiance=10.0, light_down=True):
        self.its = intersector
        self.albedo = float(albedo)
        self.env = float(env)
        c = np.asarray(light_corner, dtype=np.float64)
        u = np.asarray(light_u, dtype=np.float64)
        v = np.asarray(light_v, dtype=np.float64)
        n = np.cross(u, v)
        self.area = float(np.linalg.norm(n))
        n = n / self.area
        if light_down and n[2] > 0:
            u, v, n = v, u, -n
        self.light = (c, u, v, n)
        self.radiance = float(radiance)


# ------------------------------------------------------------------ random

def _hash(torch, x):
    """A 32-bit integer mix (lowbias32 constants) on int64 tensors."""
    x = x & 0xFFFFFFFF
    x = x ^ (x >> 16)
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x = x ^ (x >> 15)
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x = x ^ (x >> 16)
    return x


def uniform(torch, path_id, bounce, dim, seed):
    """Uniform [0, 1) per path for (bounce, dimension): float64."""
    key = _hash(torch, path_id * 0x9E3779B1 + seed)
    key = _hash(torch, key ^ (bounce * 0x85EBCA77 + dim * 0xC2B2AE3D))
    return (key.to(torch.float64) + 0.5) / 4294967296.0


# ------------------------------------------------------------------- tracer

def _length(x):
    """Euclidean length of the columns of a (3, n) tensor as (x0 x0 + x1 x1) + x2 x2,
    each product rounded on its own: the device kernels' arithmetic (a fused
    norm reduction may contract the products into FMAs and round
    differently)."""
    return (x * x).sum(0, keepdim=True).sqrt()


def _cross(torch, a, b):
    return torch.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def _light_hit(torch, scene, o, d):
    """Distance along (o, d) to the light rectangle's front face, inf if none."""
    c, u, v, n = (torch.tensor(x, dtype=torch.float64, device=o.device)[:, None] for x in scene.light)
    dn = (d * n).sum(0)
    t = ((c - o) * n).sum(0) / dn
    p = o + t * d - c
    uu = (p * u).sum(0) / (u * u).sum(0)
    vv = (p * v).sum(0) / (v * v).sum(0)
    ok = (dn < 0) & (t > 0) & (uu >= 0) & (uu <= 1) & (vv >= 0) & (vv <= 1)
    return torch.where(ok, t, torch.full_like(t, math.inf))


def render(torch, scene, cam_o, cam_d, spp=4, depth=4, seed=0, strategy="mis", stats=None, trace_log=None):
    """Radiance per pixel, (n_pixels,) float64. cam_o, cam_d: (3, n_pixels)
    float64 camera rays on the intersector's device. strategy 'mis'
    (NEE + BSDF with the power heuristic) or 'bsdf' (BSDF sampling only).
    trace_log: a list that receives, for every intersector call in order,
    ("closest" or "occluded", the pixel index of each ray of the batch)."""
    dev = cam_o.device
    n_pix = cam_o.shape[1]
    c, lu, lv, ln = (torch.tensor(x, dtype=torch.float64, device=dev)[:, None] for x in scene.light)
    image = torch.zeros(n_pix, dtype=torch.float64, device=dev)
    rays_closest = rays_shadow = 0
    for s in range(spp):
        pix = torch.arange(n_pix, device=dev, dtype=torch.int64)
        pid = pix * spp + s
        o, d = cam_o.clone(), cam_d.clone()
        d = d / _length(d)
        beta = torch.ones(n_pix, dtype=torch.float64, device=dev)
        pdf_prev = torch.zeros(n_pix, dtype=torch.float64, device=dev)  # 0: camera ray (no MIS)
        for b in range(depth + 1):
            if pix.numel() == 0:
                break
            hit = scene.its.closest(o, d)
            rays_closest += pix.numel()
            if trace_log is not None:
                trace_log.append(("closest", pix.cpu()))
            t_surf = hit["t"]
            t_light = _light_hit(torch, scene, o, d)
            on_light = t_light < t_surf
            # light reached by this ray (camera or BSDF sampled)
            if bool(on_light.any()):
                cos_l = -(d * ln).sum(0)
                if strategy == "mis":
                    pdf_l = (t_light * t_light) / (scene.area * cos_l.clamp_min(1e-12))
                    w = torch.where(pdf_prev > 0, pdf_prev ** 2 / (pdf_prev ** 2 + pdf_l ** 2), torch.ones_like(pdf_l))
                else:
                    w = torch.ones_like(t_light)
                image.index_add_(0, pix, torch.where(on_light, beta * w * scene.radiance, torch.zeros_like(beta)))
            escaped = ~torch.isfinite(t_surf) & ~on_light
            if bool(escaped.any()):
                image.index_add_(0, pix, torch.where(escaped, beta * scene.env, torch.zeros_like(beta)))
            alive = torch.isfinite(t_surf) & ~on_light
            if b == depth or not bool(alive.any()):
                break
            sel = alive.nonzero().squeeze(1)
            pix, pid, beta = pix[sel], pid[sel], beta[sel]
            o, d, t = o[:, sel], d[:, sel], t_surf[sel]
            ng = hit["ng"][:, sel]
            p = o + t * d
            # shade on the side the ray came from
            ng = torch.where((ng * d).sum(0) > 0, -ng, ng)
            # frame about ng
            a = torch.where(ng[0].abs() < 0.9, torch.tensor([1.0, 0.0, 0.0], device=dev, dtype=torch.float64)[:, None],
                            torch.tensor([0.0, 1.0, 0.0], device=dev, dtype=torch.float64)[:, None]).expand_as(ng)
            t1 = _cross(torch, a, ng)
            t1 = t1 / _length(t1)
            t2 = _cross(torch, ng, t1)
            f = scene.albedo / math.pi
            origin = p + SPAWN_EPS * ng
            if strategy == "mis":
                # next-event estimation on the light
                ua = uniform(torch, pid, b, 0, seed)
                va = uniform(torch, pid, b, 1, seed)
                q = c + ua * lu + va * lv
                to_l = q - origin
                dist2 = (to_l * to_l).sum(0)
                dist = dist2.sqrt()
                wi = to_l / dist
                cos_s = (wi * ng).sum(0)
                cos_l = -(wi * ln).sum(0)
                ok = (cos_s > 0) & (cos_l > 0)
                if bool(ok.any()):
                    k = ok.nonzero().squeeze(1)
                    occ = scene.its.occluded(origin[:, k], wi[:, k], dist[k] - LIGHT_EPS)
                    rays_shadow += k.numel()
                    if trace_log is not None:
                        trace_log.append(("occluded", pix[k].cpu()))
                    vis = torch.zeros_like(ok)
                    vis[k] = ~occ
                    pdf_l = dist2 / (scene.area * cos_l.clamp_min(1e-12))
                    pdf_b = cos_s.clamp_min(0) / math.pi
                    w = pdf_l ** 2 / (pdf_l ** 2 + pdf_b ** 2)
                    contrib = beta * f * cos_s * scene.radiance * w / pdf_l
                    image.index_add_(0, pix, torch.where(ok & vis, contrib, torch.zeros_like(contrib)))
            # cosine-weighted continuation: beta *= f cos / pdf = albedo
            u1 = uniform(torch, pid, b, 2, seed)
            u2 = uniform(torch, pid, b, 3, seed)
            r, phi = u1.sqrt(), 2 * math.pi * u2
            lx, ly, lz = r * phi.cos(), r * phi.sin(), (1 - u1).clamp_min(0).sqrt()
            d = lx * t1 + ly * t2 + lz * ng
            beta = beta * scene.albedo
            pdf_prev = lz / math.pi
            o = origin
    if stats is not None:
        stats["closest_rays"] = rays_closest
        stats["shadow_rays"] = rays_shadow
    return image / spp


def render_device(scene, cam_o, cam_d, spp=4, depth=4, seed=0, strategy="mis", stats=None):
    """render() in one library call (dmap_pathtrace): the same estimator,
    random numbers and arithmetic, with the shading, next-event and
    continuation steps as device kernels between the batched closest-hit and
    any-hit calls, all samples of a pixel in one wavefront. scene.its must
    be a GpuIntersector."""
    its = scene.its
    if not isinstance(its, GpuIntersector):
        raise TypeError("render_device traces through the library: the scene needs a GpuIntersector")
    image, n_closest, n_shadow = its.obj.pathtrace(cam_o, cam_d, spp, depth, scene.light, scene.radiance, scene.env,
                                                   scene.albedo, seed=seed, strategy=strategy, options=its.options)
    if stats is not None:
        stats["closest_rays"] = n_closest
        stats["shadow_rays"] = n_shadow
    return image


# ------------------------------------------------------------ intersectors

class GpuIntersector:
    """The batched GPU closest hit / any hit on float64 device tensors (rays
    go to the library in fp32, as its ABI takes them)."""

    def __init__(self, obj, options=None):
        self.obj = obj
        self.options = options

    def _rays(self, torch, o, d, tmax=None):
        n = o.shape[1]
        r = torch.empty((8, n), dtype=torch.float32, device=o.device)
        r[0:3] = o.to(torch.float32)
        r[3:6] = d.to(torch.float32)
        r[6] = 0.0
        r[7] = float("inf") if tmax is None else tmax.to(torch.float32)
        return r

    def closest(self, o, d):
        import torch
        res = self.obj.intersect(self._rays(torch, o, d), options=self.options)
        return {"t": res["t"].to(torch.float64), "ng": res["ng"].to(torch.float64)}

    def occluded(self, o, d, tmax):
        import torch
        return self.obj.occluded(self._rays(torch, o, d, tmax), options=self.options).to(torch.bool)


def c3_scene(torch, n=4096, res=1024, device="cuda"):
    """Config 3: the incoherent workload's surface family (64 quads on
    [0, 8]^2, fBm + ridged at amplitude 0.1 / 0.03) at n^2 texels, the same
    pinhole camera at res x res, a 2 x 2 light at height 3 facing down over
    the middle (radiance 10), environment 0.2, albedo 0.7."""
    import dmap_amd
    from dmap_amd import serving as SV
    from dmap_amd import workload as W
    tex = SV.scene_map(n, W.NOISE_SEED + 3, device)
    tile = dmap_amd.DisplacementTile(tex, 0.1, 0.0, True, "fold")
    obj = dmap_amd.ObjectIntersector(tile, W.quad_plane(8, 8.0))
    o, d = W.pinhole_rays((4.0, -0.5, 3.5), (4.0, 4.0, 0.0), 60.0, res, res)
    cam_o = torch.from_numpy(o.T.astype(np.float64)).to(device)
    cam_d = torch.from_numpy(d.T.astype(np.float64)).to(device)
    scene = Scene(GpuIntersector(obj), albedo=0.7, env=0.2, light_corner=(3.0, 3.0, 3.0), light_u=(2.0, 0.0, 0.0),
                  light_v=(0.0, 2.0, 0.0), radiance=10.0)
    scene.tile, scene.obj = tile, obj
    return scene, cam_o, cam_d
