"""Comparison of GPU hits with the oracle's, for the parity tests.

The bar: hit/miss agree exactly outside the stated bands; where both hit,
t agrees to 1e-4 relative, the tile coordinates st to 1e-4 relative
(1e-4 x max(1, |st|) in tile units: coordinates of repeated tiles grow with
the uv span) and the unit normal to 1e-4. Bands where fp32 and fp64 may legitimately
disagree on hit/miss or on which root is nearest:
  * grazing: |cos(ray, normal)| < GRAZING at the nearer of the two hits;
  * domain edge: the nearer hit lies within EDGE (tile units) of its
    triangle's domain boundary.
And one band for the normal alone:
  * texel-edge crease: the hit lies within CREASE texels of a texel edge,
    where the bilinear surface is only C0 and its normal depends on which
    texel's gradient is read (the reference reads the texel that
    HeightGrid::cell picks for the final iterate, displaced_surface.cpp:12-26).
    t and st are still compared there.
"""
import numpy as np

GRAZING = 1e-2
EDGE = 1e-5
CREASE = 1e-4
REL = 1e-4


def edge_distance(tri_rows, prim, st):
    """Distance (tile units) of st to the boundary of triangle prim's domain."""
    t = tri_rows[prim, 18:24].reshape(-1, 3, 2)
    dist = np.full(len(prim), np.inf)
    for k in range(3):
        a, b = t[:, k], t[:, (k + 1) % 3]
        e = b - a
        L = np.linalg.norm(e, axis=1)
        cr = np.abs(e[:, 0] * (st[:, 1] - a[:, 1]) - e[:, 1] * (st[:, 0] - a[:, 0])) / L
        dist = np.minimum(dist, cr)
    return dist


def crease(st, n_leaf):
    f = st * n_leaf
    fr = f - np.floor(f)
    return (np.minimum(fr, 1 - fr) < CREASE).any(axis=1)


def compare(gpu, ref, d, tri_rows, n_leaf, rel=REL):
    """gpu, ref: dicts of numpy arrays t (n), st (n, 2), ng (n, 3), prim (n).
    d: (n, 3) ray directions. Returns a dict of counts and worst errors."""
    tg, tr = gpu["t"].astype(np.float64), ref["t"]
    hg, hr = np.isfinite(tg), np.isfinite(tr)
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    out = {"n": len(tg), "hits_ref": int(hr.sum()), "hits_gpu": int(hg.sum())}
    # Band membership from the nearer hit.
    g_nearer = hg & (~hr | (tg <= tr))
    cos = np.ones(len(tg))
    edge = np.full(len(tg), np.inf)
    if g_nearer.any():
        cos[g_nearer] = np.abs(np.sum(dn[g_nearer] * gpu["ng"][g_nearer], axis=1))
        edge[g_nearer] = edge_distance(tri_rows, gpu["prim"][g_nearer], gpu["st"][g_nearer].astype(np.float64))
    r_n = ~g_nearer & hr
    if r_n.any():
        cos[r_n] = np.abs(np.sum(dn[r_n] * ref["ng"][r_n], axis=1))
        edge[r_n] = edge_distance(tri_rows, ref["prim"][r_n], ref["st"][r_n])
    band = (cos < GRAZING) | (edge < EDGE)
    both = hg & hr
    dt = np.zeros(len(tg))
    dt[both] = np.abs(tg[both] - tr[both]) / np.abs(tr[both])
    dst = np.zeros(len(tg))
    dst[both] = (np.abs(gpu["st"][both].astype(np.float64) - ref["st"][both])
                 / np.maximum(1.0, np.abs(ref["st"][both]))).max(axis=1)
    dng = np.zeros(len(tg))
    dng[both] = np.abs(gpu["ng"][both].astype(np.float64) - ref["ng"][both]).max(axis=1)
    flip = hg != hr
    on_crease = np.zeros(len(tg), dtype=bool)
    on_crease[both] = crease(ref["st"][both], n_leaf)
    bad_value = both & ((dt > rel) | (dst > rel) | ((dng > rel) & ~on_crease))
    out["normal_on_crease"] = int((both & on_crease & (dng > rel)).sum())
    out["hitmiss_flips"] = int(flip.sum())
    out["flips_in_band"] = int((flip & band).sum())
    out["value_mismatch"] = int(bad_value.sum())
    out["value_mismatch_in_band"] = int((bad_value & band).sum())
    out["unexplained"] = int(((flip | bad_value) & ~band).sum())
    ok = both & ~band
    out["worst_t_rel"] = float(dt[ok].max()) if ok.any() else 0.0
    out["worst_st"] = float(dst[ok].max()) if ok.any() else 0.0
    okn = ok & ~on_crease
    out["worst_ng"] = float(dng[okn].max()) if okn.any() else 0.0
    out["bad_index"] = np.nonzero((flip | bad_value) & ~band)[0]
    # Per ray: 0 agrees within the bar, 1 differs inside a band, 2 differs unexplained.
    differs = flip | bad_value
    out["ray_flag"] = np.where(differs & ~band, 2, np.where(differs, 1, 0)).astype(np.int8)
    return out
