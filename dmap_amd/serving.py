"""The serving workload and its driver.

Eight resident scenes, each 64 quads tiling [0, 8]^2 over one seeded
fBm(6) + ridged displacement (amplitude 0.1 / 0.03, the incoherent
workload's family) of 512^2, 1024^2, 2048^2 (x3), 4096^2 (x2) and 8192^2
fp32 texels. Each scene has a pool of 1,048,576 pinhole primaries (a
1024 x 1024 camera in 8 x 8 pixel blocks) and 1,048,576 cosine-weighted
bounce rays. A request picks a scene uniformly and a ray count log-uniform
in [1k, 1M]; 70% of its rays are a contiguous run of the scene's primaries
(coherent), 30% a contiguous run of its bounces (incoherent). Requests
arrive as a Poisson process at a fixed rate; the driver submits each at its
arrival time, polls completions and releases their hits at once.

The maps are generated on the device in float64 (value noise with the same
lattice seeds as the host generator) so the 8192^2 scene takes no host
time; nothing here is compared with the oracle: serving results are checked
against the single-request entry point instead.
"""
import math
import time

import numpy as np

from . import workload as W

SCENE_SIZES = (512, 1024, 2048, 2048, 2048, 4096, 4096, 8192)
POOL = 1 << 20
PRIMARY_FRACTION = 0.7
MIN_RAYS, MAX_RAYS = 1_000, 1_000_000
SMALL = 65_536


def _value_noise_t(torch, n, period, rng, device):
    lat = torch.from_numpy(rng.random((period, period))).to(device)
    x = torch.arange(n, dtype=torch.float64, device=device) * (period / n)
    i0 = torch.floor(x).long()
    f = x - i0
    u = f * f * (3.0 - 2.0 * f)
    i1 = (i0 + 1) % period
    i0 = i0 % period
    ux, uy = u[None, :], u[:, None]
    top = (1.0 - ux) * lat[i0][:, i0] + ux * lat[i0][:, i1]
    bot = (1.0 - ux) * lat[i1][:, i0] + ux * lat[i1][:, i1]
    return (1.0 - uy) * top + uy * bot


def _octaves_t(torch, n, octaves, base_period, seed, device, ridged):
    rng = np.random.default_rng(seed)
    total = torch.zeros((n, n), dtype=torch.float64, device=device)
    amp, norm, period = 1.0, 0.0, base_period
    for _ in range(octaves):
        v = 2.0 * _value_noise_t(torch, n, min(period, n), rng, device) - 1.0
        total += amp * ((1.0 - v.abs()) if ridged else v)
        norm += amp
        amp *= 0.5
        period *= 2
    return total / norm


def scene_map(n, seed, device="cuda"):
    """n x n fp32 texels of fBm(6) + 0.3 ridged(2) (units of 0.1)."""
    import torch
    base = max(8, n // 256)
    tex = _octaves_t(torch, n, 6, base, seed, device, False) + 0.3 * _octaves_t(torch, n, 2, 2 * base, seed + 1,
                                                                                 device, True)
    return tex.to(torch.float32)


def bounce_pool(texels, strength, size, n, seed):
    """Cosine-weighted rays about the displaced plane's normal at n uniform
    surface points (the host generator's construction, on the device)."""
    import torch
    dev = texels.device
    g = torch.Generator(device="cpu").manual_seed(seed)
    st = torch.rand((n, 2), generator=g, dtype=torch.float64).to(dev)
    m = texels.shape[0]
    V = texels.to(torch.float64)
    fx, fy = st[:, 0] * m, st[:, 1] * m
    i = torch.clamp(fx.long(), max=m - 1)
    j = torch.clamp(fy.long(), max=m - 1)
    s, t = fx - i, fy - j
    i1, j1 = (i + 1) % m, (j + 1) % m
    v00, v10, v01, v11 = V[j, i], V[j, i1], V[j1, i], V[j1, i1]
    h = strength * ((1 - s) * (1 - t) * v00 + s * (1 - t) * v10 + (1 - s) * t * v01 + s * t * v11)
    hs = strength * ((1 - t) * (v10 - v00) + t * (v11 - v01)) * m
    ht = strength * ((1 - s) * (v01 - v00) + s * (v11 - v10)) * m
    P = torch.stack([size * st[:, 0], size * st[:, 1], h], dim=1)
    N = torch.stack([-hs * size, -ht * size, torch.full_like(h, size * size)], dim=1)
    N = N / N.norm(dim=1, keepdim=True)
    u1 = torch.rand(n, generator=g, dtype=torch.float64).to(dev)
    u2 = torch.rand(n, generator=g, dtype=torch.float64).to(dev)
    r, phi = torch.sqrt(u1), 2 * math.pi * u2
    a = torch.where(N[:, :1].abs() < 0.9, torch.tensor([[1.0, 0.0, 0.0]], dtype=torch.float64, device=dev),
                    torch.tensor([[0.0, 1.0, 0.0]], dtype=torch.float64, device=dev))
    t1 = torch.linalg.cross(a, N)
    t1 = t1 / t1.norm(dim=1, keepdim=True)
    t2 = torch.linalg.cross(N, t1)
    d = (r * torch.cos(phi))[:, None] * t1 + (r * torch.sin(phi))[:, None] * t2 + \
        torch.sqrt(torch.clamp(1.0 - u1, min=0.0))[:, None] * N
    o = P + 1e-4 * N
    out = torch.empty((8, n), dtype=torch.float32, device=dev)
    out[0:3] = o.T.to(torch.float32)
    out[3:6] = d.T.to(torch.float32)
    out[6] = 0.0
    out[7] = float("inf")
    return out


class Scene:
    """A resident scene: tile, object and its two ray pools (pool rays each;
    pool is a power of 4, the camera is sqrt(pool) pixels square)."""

    def __init__(self, k, n, device="cuda", pool=POOL):
        import torch

        import dmap_amd
        self.k, self.n = k, n
        seed = W.NOISE_SEED + 101 * (k + 1)
        tex = scene_map(n, seed, device)
        self.tile = dmap_amd.DisplacementTile(tex, 0.1, 0.0, True, "fold")
        del tex
        self.obj = dmap_amd.ObjectIntersector(self.tile, W.quad_plane(8, 8.0))
        rng = np.random.default_rng(seed)
        eye = (4.0 + rng.uniform(-1, 1), -0.5 + rng.uniform(-0.3, 0.3), 3.5 + rng.uniform(-0.5, 0.5))
        target = (4.0 + rng.uniform(-1, 1), 4.0 + rng.uniform(-1, 1), 0.0)
        side = int(round(math.sqrt(pool)))
        o, d = W.pinhole_rays(eye, target, 60.0, side, side)
        self.primary = torch.from_numpy(W.pack_rays(o, d)).to(device)
        self.bounce = bounce_pool(self.tile._keep, 0.1, 8.0, pool, seed + 7)
        self.pool = pool


def make_scenes(sizes=SCENE_SIZES, device="cuda", pool=POOL):
    import torch
    out = []
    for k, n in enumerate(sizes):
        out.append(Scene(k, n, device, pool))
        torch.cuda.synchronize()
    return out


def make_requests(scenes, count, seed, min_rays=MIN_RAYS, max_rays=MAX_RAYS):
    """count requests: (scene index, (8, n) device rays). The rays are
    contiguous runs of the scene's pools (primaries start on an 8 x 8 block)."""
    import torch
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        k = int(rng.integers(len(scenes)))
        n = int(round(math.exp(rng.uniform(math.log(min_rays), math.log(max_rays)))))
        n_p = int(round(PRIMARY_FRACTION * n))
        n_b = n - n_p
        S = scenes[k]
        sp = int(rng.integers(0, (S.pool - n_p) // 64 + 1)) * 64
        sb = int(rng.integers(0, S.pool - n_b + 1))
        rays = torch.cat([S.primary[:, sp:sp + n_p], S.bounce[:, sb:sb + n_b]], dim=1).contiguous()
        out.append((k, rays))
    return out


def plain_rate(scenes, requests, repeats=3):
    """The plain batched kernel on the same rays: per scene, its requests'
    rays in one closest-hit call (reordered). Returns Mrays/s."""
    import torch
    batches = []
    for k, S in enumerate(scenes):
        mine = [r for (q, r) in requests if q == k]
        if mine:
            rays = torch.cat(mine, dim=1).contiguous()
            batches.append((S, rays, S.obj.alloc_hits(rays.shape[1])))
    for S, rays, h in batches:
        S.obj.intersect(rays, out=h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        for S, rays, h in batches:
            S.obj.intersect(rays, out=h)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / repeats
    n = sum(r.shape[1] for _, r, _ in batches)
    del batches
    return n / dt / 1e6


def drive(server, scenes, requests, rate_hz, duration_s, seed, log=None):
    """Poisson arrivals at rate_hz for duration_s (rate_hz None: every
    request at once). Requests cycle through `requests`. Returns per-request
    (rays, latency_s) and the achieved ray rate."""
    rng = np.random.default_rng(seed)
    if rate_hz is None:
        arrivals = np.zeros(len(requests))
    else:
        count = max(1, int(rate_hz * duration_s))
        arrivals = np.cumsum(rng.exponential(1.0 / rate_hz, count))
    R = len(arrivals)
    sizes = np.empty(R, dtype=np.int64)
    lag = np.zeros(R)
    lat = np.full(R, np.nan)
    i = done = 0
    t0 = time.perf_counter()
    last_log = t0
    while done < R:
        now = time.perf_counter() - t0
        while i < R and arrivals[i] <= now:
            k, rays = requests[i % len(requests)]
            sizes[i] = rays.shape[1]
            lag[i] = (time.perf_counter() - t0) - arrivals[i]
            server.submit(scenes[k].obj, rays, i)
            i += 1
        for tag, l in server.poll():
            lat[tag] = l + lag[tag]
            server.release(tag)
            done += 1
        if log and time.perf_counter() - last_log > 5.0:
            last_log = time.perf_counter()
            log(f"[serve] {done}/{R} requests done")
    wall = time.perf_counter() - t0
    rays = int(sizes.sum())
    return {"sizes": sizes, "latency_s": lat, "wall_s": wall, "mrays_s": rays / wall / 1e6,
            "offered_mrays_s": None if rate_hz is None else rays / float(arrivals[-1]) / 1e6}


def summarize(res):
    lat, sizes = res["latency_s"], res["sizes"]
    small = sizes <= SMALL

    def pct(x, q):
        return round(float(np.percentile(x, q)) * 1e3, 3) if x.size else None

    return {"requests": int(lat.size), "rays": int(sizes.sum()), "sustained_mrays_s": round(res["mrays_s"], 2),
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
            "small_requests": int(small.sum()), "p50_ms_small": pct(lat[small], 50),
            "p99_ms_small": pct(lat[small], 99)}
