"""Seeded synthetic workloads: fBm / ridged displacement maps, base quads
split into triangles, pinhole primaries and cosine-weighted bounce rays.

Nothing in the reference defines these generators; they produce the inputs
named by the benchmark configurations, bit-identically for the GPU path and
the CPU oracle (both read the same fp32 arrays):

  C1 "small":       one unit quad in z = 0 (normal +z), 256^2 map of 4-octave
                    fBm value noise (lacunarity 2, gain 0.5) at amplitude 0.05,
                    512 x 512 pinhole primaries from (0.5, -0.8, 0.9) towards
                    (0.5, 0.5, 0), 60 degree vertical field of view.
  C2 "incoherent":  64 quads tiling [0, 8]^2, one 2048^2 map shared by all
                    (quad (qi, qj) maps to tile coordinates
                    [qi/8, (qi+1)/8] x [qj/8, (qj+1)/8], each texel used once),
                    6-octave fBm at amplitude 0.1 plus ridged noise (abs-fold)
                    at amplitude 0.03, 12 pyramid levels (depth 11);
                    4,194,304 rays: 2048 x 1024 pinhole primaries from
                    (4, -0.5, 3.5) towards (4, 4, 0), 60 degree vertical field
                    of view (about half of them meet the plane), then 2,097,152 cosine-weighted bounce rays from
                    uniform surface points, origins lifted 1e-4 along the
                    surface normal.

  C4 "adjoint":     one unit quad, 1024^2 map of 4-octave fBm at amplitude
                    0.05, the C1 camera at 512 x 512; targets are rendered
                    from the same map plus 16 seeded Gaussian blobs of
                    sigma 8 texels and height amplitude 0.01 (0.2 in texel
                    units), wrapped with the tile.

A texel value v gives the height h = strength * v (offset 0): the maps are
stored in units of their amplitude.
"""
import numpy as np

NOISE_SEED = 2027
RAY_SEED = 4099


# ----------------------------------------------------------------- noise

def value_noise(n, period, rng):
    """Tileable value noise on an n x n grid: a period x period lattice of
    uniform values, smoothstep-interpolated; sample (i, j) at lattice
    coordinates (i, j) * period / n. Returns values in [0, 1], [j, i]."""
    lat = rng.random((period, period))
    x = np.arange(n, dtype=np.float64) * (period / n)
    i0 = np.floor(x).astype(np.int64)
    f = x - i0
    u = f * f * (3.0 - 2.0 * f)
    i1 = (i0 + 1) % period
    i0 = i0 % period
    a = lat[np.ix_(i0, i0)]
    b = lat[np.ix_(i0, i1)]
    c = lat[np.ix_(i1, i0)]
    d = lat[np.ix_(i1, i1)]
    ux, uy = u[None, :], u[:, None]
    return (1.0 - uy) * ((1.0 - ux) * a + ux * b) + uy * ((1.0 - ux) * c + ux * d)


def fbm(n, octaves, base_period, seed, lacunarity=2, gain=0.5):
    """Fractional Brownian motion of value noise, normalized to [-1, 1]."""
    rng = np.random.default_rng(seed)
    total = np.zeros((n, n))
    amp, norm, period = 1.0, 0.0, base_period
    for _ in range(octaves):
        total += amp * (2.0 * value_noise(n, min(period, n), rng) - 1.0)
        norm += amp
        amp *= gain
        period *= lacunarity
    return total / norm


def ridged(n, octaves, base_period, seed, lacunarity=2, gain=0.5):
    """Ridged noise: octaves of 1 - |2 v - 1| (the abs-fold), normalized to [0, 1]."""
    rng = np.random.default_rng(seed)
    total = np.zeros((n, n))
    amp, norm, period = 1.0, 0.0, base_period
    for _ in range(octaves):
        total += amp * (1.0 - np.abs(2.0 * value_noise(n, min(period, n), rng) - 1.0))
        norm += amp
        amp *= gain
        period *= lacunarity
    return total / norm


# -------------------------------------------------------------- geometry

def quad_plane(nq, size, z=0.0):
    """nq x nq quads tiling [0, size]^2 at height z, normal +z; quad (qi, qj)
    maps to tile coordinates [qi/nq, (qi+1)/nq] x [qj/nq, (qj+1)/nq]. Each
    quad is two triangles split along its (00)-(11) diagonal, both wound
    counter-clockwise seen from +z. Rows: p0 p1 p2, m0 m1 m2, t0 t1 t2."""
    rows = []
    step = size / nq
    nrm = [0.0, 0.0, 1.0] * 3
    for qj in range(nq):
        for qi in range(nq):
            x0, x1, y0, y1 = qi * step, (qi + 1) * step, qj * step, (qj + 1) * step
            u0, u1, v0, v1 = qi / nq, (qi + 1) / nq, qj / nq, (qj + 1) / nq
            rows.append([x0, y0, z, x1, y0, z, x1, y1, z] + nrm + [u0, v0, u1, v0, u1, v1])
            rows.append([x0, y0, z, x1, y1, z, x0, y1, z] + nrm + [u0, v0, u1, v1, u0, v1])
    return np.asarray(rows, dtype=np.float64)


def close_tile(texels, repeat=True):
    """(n+1)^2 closed node grid (texture_grid.cpp:118-132)."""
    n = texels.shape[0]
    idx = np.arange(n + 1)
    idx = idx % n if repeat else np.minimum(idx, n - 1)
    return texels[np.ix_(idx, idx)]


def field(nodes, strength, offset, repeat, x, y):
    """HeightGrid::h and grad (displaced_surface.cpp:12-47) at tile
    coordinates (x, y): height and (dh/dx, dh/dy)."""
    W = nodes.shape[0]
    V = nodes.astype(np.float64)
    if repeat:
        x = x - np.floor(x)
        y = y - np.floor(y)
    else:
        x = np.clip(x, 0.0, 1.0)
        y = np.clip(y, 0.0, 1.0)
    fx, fy = x * (W - 1), y * (W - 1)
    i = np.minimum(fx.astype(np.int64), W - 2)
    j = np.minimum(fy.astype(np.int64), W - 2)
    s, t = fx - i, fy - j
    v00, v10, v01, v11 = V[j, i], V[j, i + 1], V[j + 1, i], V[j + 1, i + 1]
    h = strength * ((1 - s) * (1 - t) * v00 + s * (1 - t) * v10 + (1 - s) * t * v01 + s * t * v11) + offset
    gs = (1 - t) * (v10 - v00) + t * (v11 - v01)
    gt = (1 - s) * (v01 - v00) + s * (v11 - v10)
    return h, strength * gs * (W - 1), strength * gt * (W - 1)


# ------------------------------------------------------------------ rays

def pinhole_rays(eye, target, fov_deg, width, height, tile=8, up=(0.0, 0.0, 1.0)):
    """One ray per pixel centre, unit directions, ordered by tile x tile
    pixel blocks (a wave of 64 lanes covers one 8 x 8 block)."""
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    th = np.tan(np.radians(fov_deg) / 2)
    aspect = width / height
    ty, tx = np.meshgrid(np.arange(height // tile), np.arange(width // tile), indexing="ij")
    ly, lx = np.meshgrid(np.arange(tile), np.arange(tile), indexing="ij")
    py = (ty[:, :, None, None] * tile + ly[None, None]).reshape(-1)
    px = (tx[:, :, None, None] * tile + lx[None, None]).reshape(-1)
    sx = (2.0 * (px + 0.5) / width - 1.0) * th * aspect
    sy = (1.0 - 2.0 * (py + 0.5) / height) * th
    d = f[None, :] + sx[:, None] * r[None, :] + sy[:, None] * u[None, :]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(eye, d.shape)
    return o.astype(np.float32), d.astype(np.float32)


def bounce_rays(nodes, strength, offset, repeat, size, n, seed, lift=1e-4):
    """Cosine-weighted hemisphere rays about the surface normal at uniform
    points of the displaced plane [0, size]^2 (P = size * (s, t, 0), N = +z)."""
    rng = np.random.default_rng(seed)
    st = rng.random((n, 2))
    h, hs, ht = field(nodes, strength, offset, repeat, st[:, 0], st[:, 1])
    S = np.stack([size * st[:, 0], size * st[:, 1], h], axis=1)
    # Su x Sv with Su = (size, 0, hs), Sv = (0, size, ht).
    nrm = np.stack([-hs * size, -ht * size, np.full(n, size * size)], axis=1)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    u1, u2 = rng.random(n), rng.random(n)
    r, phi = np.sqrt(u1), 2 * np.pi * u2
    local = np.stack([r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1.0 - u1))], axis=1)
    a = np.where(np.abs(nrm[:, :1]) < 0.9, np.array([[1.0, 0.0, 0.0]]), np.array([[0.0, 1.0, 0.0]]))
    t1 = np.cross(a, nrm)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(nrm, t1)
    d = local[:, :1] * t1 + local[:, 1:2] * t2 + local[:, 2:3] * nrm
    o = S + lift * nrm
    return o.astype(np.float32), d.astype(np.float32)


def pack_rays(o, d, tmin=0.0, tmax=np.inf):
    """(8, n) float32: o xyz, d xyz, tmin, tmax (the C ABI's SoA rays)."""
    n = o.shape[0]
    out = np.empty((8, n), dtype=np.float32)
    out[0:3] = o.T
    out[3:6] = d.T
    out[6] = tmin
    out[7] = tmax
    return out


# --------------------------------------------------------------- configs

class Workload:
    def __init__(self, name, texels, strength, offset, repeat, triangles, rays, n_primary, description):
        self.name = name
        self.texels = texels.astype(np.float32)
        self.strength, self.offset, self.repeat = strength, offset, repeat
        self.triangles = triangles
        self.rays = rays
        self.n_primary = n_primary
        self.description = description

    @property
    def nodes(self):
        return close_tile(self.texels, self.repeat)


def c1_small(n=256, width=512, height=512):
    tex = fbm(n, 4, max(4, n // 64), NOISE_SEED).astype(np.float32)
    tris = quad_plane(1, 1.0)
    o, d = pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 60.0, width, height)
    return Workload("c1_small", tex, 0.05, 0.0, True, tris, pack_rays(o, d), o.shape[0],
                    f"1 quad, {n}^2 fBm(4) amp 0.05, {width}x{height} primaries")


def c2_map(n=2048):
    base = max(8, n // 256)
    tex = fbm(n, 6, base, NOISE_SEED) + 0.3 * ridged(n, 2, 2 * base, NOISE_SEED + 1)
    return tex.astype(np.float32)


def c2_incoherent(n=2048, width=2048, height=1024, n_bounce=None, nq=8, size=8.0):
    tex = c2_map(n)
    tris = quad_plane(nq, size)
    k = size / 8.0
    o1, d1 = pinhole_rays((4.0 * k, -0.5 * k, 3.5 * k), (4.0 * k, 4.0 * k, 0.0), 60.0, width, height)
    nb = width * height if n_bounce is None else n_bounce
    nodes = close_tile(tex, True)
    o2, d2 = bounce_rays(nodes, 0.1, 0.0, True, size, nb, RAY_SEED)
    rays = pack_rays(np.concatenate([o1, o2]), np.concatenate([d1, d2]))
    return Workload("c2_incoherent", tex, 0.1, 0.0, True, tris, rays, o1.shape[0],
                    f"{nq*nq} quads on [0,{size:g}]^2, {n}^2 fBm(6)+ridged amp 0.1/0.03, "
                    f"{o1.shape[0]} primaries + {nb} cosine bounces")


def gaussian_blobs(n, count, sigma, amplitude, seed):
    """Sum of `count` Gaussian bumps (sigma in texels, peak `amplitude`) at
    seeded centres and signs, periodic over the n x n tile. [j, i]."""
    rng = np.random.default_rng(seed)
    x = np.arange(n, dtype=np.float64)
    out = np.zeros((n, n))
    for _ in range(count):
        cx, cy = rng.uniform(0, n, 2)
        sign = 1.0 if rng.random() < 0.5 else -1.0
        dx = np.abs(x - cx)
        dx = np.minimum(dx, n - dx)
        dy = np.abs(x - cy)
        dy = np.minimum(dy, n - dy)
        out += sign * amplitude * np.exp(-(dy[:, None] ** 2 + dx[None, :] ** 2) / (2 * sigma * sigma))
    return out


class AdjointWorkload(Workload):
    def __init__(self, texels, target_texels, strength, triangles, rays, description):
        super().__init__("c4_adjoint", texels, strength, 0.0, True, triangles, rays, rays.shape[1], description)
        self.target_texels = target_texels.astype(np.float32)


def c4_adjoint(n=1024, width=512, height=512, blobs=16):
    strength = 0.05
    tex = fbm(n, 4, max(4, n // 64), NOISE_SEED).astype(np.float32)
    target = tex + gaussian_blobs(n, blobs, 8.0 * n / 1024, 0.01 / strength, NOISE_SEED + 5)
    tris = quad_plane(1, 1.0)
    o, d = pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 60.0, width, height)
    return AdjointWorkload(tex, target, strength, tris, pack_rays(o, d),
                           f"1 quad, {n}^2 fBm(4) amp 0.05, {width}x{height} primaries, targets +{blobs} blobs")
