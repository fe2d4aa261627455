"""Small seeded scenes for the GPU parity tests: the flat quad of the
benchmark configurations, a quad with interpolated (non-constant) vertex
normals, and a quad with a rotated, scaled texture chart."""
import numpy as np

from dmap_amd import workload as W


def curved_quad():
    """Unit quad in z = 0 whose vertex normals lean outwards (Mu, Mv != 0)."""
    tris = W.quad_plane(1, 1.0)
    for row in tris:
        for v in range(3):
            x, y = row[3 * v], row[3 * v + 1]
            n = np.array([0.35 * (x - 0.5), 0.25 * (y - 0.5), 1.0])
            row[9 + 3 * v:12 + 3 * v] = n / np.linalg.norm(n)
    return tris


def rotated_chart_quad():
    """Unit quad whose corners map to a rotated, scaled square inside the tile."""
    tris = W.quad_plane(1, 1.0)
    c, ang, sc = np.array([0.5, 0.5]), 0.4, 0.55
    R = np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    for row in tris:
        for v in range(3):
            xy = row[3 * v:3 * v + 2]
            row[18 + 2 * v:20 + 2 * v] = c + sc * R @ (xy - 0.5)
    return tris


def span_quad(lo, hi):
    """Unit quad in z = 0 whose texture coordinates run from (lo, lo) to (hi, hi):
    a repeated tile, hi - lo tiles across the quad."""
    tris = W.quad_plane(1, 1.0)
    for row in tris:
        for v in range(3):
            row[18 + 2 * v:20 + 2 * v] = lo + (hi - lo) * row[3 * v:3 * v + 2]
    return tris


def rays(n, seed, lo=-0.2, hi=1.2):
    """Half from above towards the quad, half near the surface in random
    directions (the distributions of traversal_test_util.cpp:83-111)."""
    rng = np.random.default_rng(seed)
    m = n // 2
    o1 = np.column_stack([rng.uniform(lo, hi, m), rng.uniform(lo, hi, m), rng.uniform(0.4, 1.5, m)])
    tgt = np.column_stack([rng.uniform(0, 1, m), rng.uniform(0, 1, m), np.zeros(m)])
    d1 = tgt - o1
    o2 = np.column_stack([rng.uniform(0, 1, n - m), rng.uniform(0, 1, n - m), rng.uniform(-0.08, 0.08, n - m)])
    d2 = rng.normal(size=(n - m, 3))
    d2 /= np.linalg.norm(d2, axis=1, keepdims=True)
    return W.pack_rays(np.vstack([o1, o2]).astype(np.float32), np.vstack([d1, d2]).astype(np.float32))
