"""CPU oracle for the ray / displaced-surface intersection path.

TEST INFRASTRUCTURE ONLY. A ctypes view of oracle/dmap_oracle.cpp, the
double-precision restatement of the reference's CPU intersector
(code/cpp/src/{taylor_pyramid,affine_form,uv_clip,node_bounds,
displaced_intersector}.cpp). Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this package, and only as the
checker or the timed CPU baseline: the product library (dmap_amd) never
touches it.

Parity pins (each function of the restatement is checked by at least one):
  * pyramid (fold + direct): the reference's numpy golden vectors
    (tests/golden/pyramid_*.npz, written by tests/golden/make_pyramid_golden.py
    from code/python/poc/dmapref/pyramid.py), the leaf closed form, the ramp,
    exact min-max, fold/direct agreement, dense containment;
  * node bounds: dense containment of surface points in box and slab, rays
    through them inside every interval;
  * walk + leaves: brute force over all texels (two-triangle leaf vs
    Moller-Trumbore on the texel mesh; Newton leaf vs the bilinear-patch
    closed form), never-miss on aimed rays, bound-mode invariance, flat map;
  * the general-normal and general-chart branch (interpolated normals,
    rotated charts, tilted frames; tests/test_oracle_general.py): the
    surface and tangent frame against their barycentric definition, Min-Range
    normalize against dense sampling, node bounds against dense surface
    points, never-miss, the Newton residual, a texel-aligned fine mesh, and
    bound-mode invariance.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_build" / "libdmap_oracle.so"

CHANNELS = ("h0", "gu", "gv", "r", "ru", "rv", "h_min", "h_max")
STAT_NAMES = ("triangle_tests", "node_tests", "leaf_tests", "sub_squares", "newton_runs", "newton_iters",
              "node_tests_after_hit", "leaf_tests_after_hit")
MAX_LEVELS = 32  # TraversalStats::max_levels (displaced_intersector.h:91)
N_STATS = len(STAT_NAMES) + MAX_LEVELS
BOUND = {"box": 0, "slab": 1, "both": 2}
LEAF = {"newton": 0, "two_triangle": 1}

_lib = None


def build(force=False):
    """Compile the oracle with the committed Makefile (g++ only)."""
    cmd = ["make", "-s", "-C", str(_HERE)]
    if force:
        cmd.insert(2, "-B")
    subprocess.run(cmd, check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)
_I32 = ctypes.POINTER(ctypes.c_int32)
_U8 = ctypes.POINTER(ctypes.c_uint8)
_VP = ctypes.c_void_p


def _declare(L):
    c = ctypes
    L.orc_pyramid_levels.argtypes = [c.c_int]
    L.orc_pyramid_cells.argtypes = [c.c_int]
    L.orc_pyramid_cells.restype = c.c_int64
    L.orc_pyramid_build.argtypes = [_D, c.c_int, c.c_double, c.c_double, c.c_int, _D]
    L.orc_close_tile.argtypes = [_D, c.c_int, c.c_int, _D]
    L.orc_scene_create.argtypes = [_D, c.c_int, c.c_double, c.c_double, c.c_int, c.c_int, c.c_int, _D]
    L.orc_scene_create.restype = _VP
    L.orc_scene_destroy.argtypes = [_VP]
    L.orc_triangle_info.argtypes = [_VP, c.c_int, c.POINTER(c.c_int), c.POINTER(c.c_int), _I64, _D, _D, _D]
    L.orc_scene_intersect.argtypes = [_VP, c.c_int64, _D, _D, _D, _D, c.c_int, c.c_int, c.c_int, c.c_int,
                                      _D, _D, _D, _D, _I32, _I64, _I64]
    L.orc_scene_occluded.argtypes = [_VP, c.c_int64, _D, _D, _D, _D, c.c_int, c.c_int, c.c_int, c.c_int, _U8]
    L.orc_scene_records.argtypes = [_VP, c.c_int64, _D, _D, _D, _D, c.c_int, c.c_int, c.c_int, c.c_int64, _VP,
                                    _I64]
    L.orc_triangle_intersect.argtypes = [_VP, c.c_int, _D, _D, c.c_double, c.c_double, c.c_int, c.c_int,
                                         c.c_int, _D, _I64]
    L.orc_leaf_newton.argtypes = [_VP, c.c_int, _D, _D, c.c_double, c.c_double, c.c_int64, c.c_int64, _D, _I64]
    L.orc_newton_from.argtypes = [_VP, c.c_int, _D, _D, c.c_double, c.c_double, c.c_int64, c.c_int64,
                                  c.c_double, c.c_double, _D]
    L.orc_leaf_two_triangle.argtypes = [_VP, c.c_int, _D, _D, c.c_double, c.c_double, c.c_int64, c.c_int64, _D]
    L.orc_node_bound.argtypes = [_VP, c.c_int, c.c_int, c.c_int64, c.c_int64, _D, _D, _D, _D, _D]
    L.orc_node_interval.argtypes = [_VP, c.c_int, c.c_int, c.c_int64, c.c_int64, _D, _D, c.c_double,
                                    c.c_double, c.c_int, _D]
    L.orc_displaced_position.argtypes = [_VP, c.c_int, c.c_double, c.c_double, _D]
    L.orc_surface_normal.argtypes = [_VP, c.c_int, c.c_double, c.c_double, _D]
    L.orc_surface_points.argtypes = [_VP, c.c_int64, _I32, _D, _D, _D]
    L.orc_field_eval.argtypes = [_VP, c.c_int64, _D, _D, _D]
    L.orc_classify_square.argtypes = [_D, c.c_double, c.c_double, c.c_double]
    L.orc_surface_point.argtypes = [_VP, c.c_int, c.c_double, c.c_double, _D]
    L.orc_hit_record.argtypes = [_VP, c.c_int, c.c_double, c.c_double, _D, _D, c.c_int, _D]
    L.orc_affine_normalize.argtypes = [_D, _D]
    L.orc_tangent_frame.argtypes = [_VP, c.c_int, _D, _D]
    L.orc_scene_adjoint.argtypes = [_VP, c.c_int64, _D, _D, _D, _D, _D, _D, c.c_int, _D, _D, _D]


def _p(a, t=_D):
    return None if a is None else a.ctypes.data_as(t)


# orc_node_record (dmap_oracle.h): NodeTestRecord (displaced_intersector.h:77-87)
# with the ray, the triangle, the cell and the test's place among the ray's
# node tests.
NODE_RECORD = np.dtype([("ray", np.int32), ("prim", np.int32), ("level", np.int32), ("seq", np.int32),
                        ("i", np.int64), ("j", np.int64), ("box_pass", np.int32), ("slab_pass", np.int32),
                        ("box_length", np.float64), ("slab_length", np.float64), ("both_length", np.float64)])


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def pyramid_levels(nodes):
    return lib().orc_pyramid_levels(int(nodes))


def build_pyramid(values, scale=1.0, build="fold", offset=0.0):
    """TaylorPyramid over a (2^L+1)^2 node grid: a list of levels, each a
    dict channel -> (m, m) float64 array indexed [j, i]."""
    V = _f64(values)
    nodes = V.shape[0]
    if V.shape != (nodes, nodes):
        raise ValueError("square node grid expected")
    total = lib().orc_pyramid_cells(nodes) * 8
    out = np.empty(total, dtype=np.float64)
    L = lib().orc_pyramid_build(_p(V), nodes, float(scale), float(offset), 0 if build == "fold" else 1, _p(out))
    if L < 0:
        raise ValueError("pyramid needs a (2^L + 1) x (2^L + 1) node grid")
    levels, pos, m = [], 0, nodes - 1
    for _ in range(L):
        lv = {}
        for ch in CHANNELS:
            lv[ch] = out[pos:pos + m * m].reshape(m, m)
            pos += m * m
        levels.append(lv)
        m //= 2
    return levels


def close_tile(texels, repeat=True):
    T = _f64(texels)
    n = T.shape[0]
    out = np.empty((n + 1, n + 1), dtype=np.float64)
    lib().orc_close_tile(_p(T), n, 1 if repeat else 0, _p(out))
    return out


def affine_normalize(form):
    """AffineVec3::normalize on a (3, 4) array of rows (c, a, b, k)."""
    f = _f64(form).reshape(12)
    out = np.empty(12)
    lib().orc_affine_normalize(_p(f), _p(out))
    return out.reshape(3, 4)


def classify_square(t6, cx, cy, s):
    return lib().orc_classify_square(_p(_f64(t6)), float(cx), float(cy), float(s))


class Scene:
    """One node grid + pyramid shared by base triangles (the reference's
    DisplacementAsset + ObjectIntersector)."""

    def __init__(self, nodes, scale, offset, repeat, triangles, build="fold"):
        self.nodes = _f64(nodes)
        self.tri = _f64(triangles).reshape(-1, 24)
        n = self.nodes.shape[0]
        self.n_tri = self.tri.shape[0]
        self.handle = lib().orc_scene_create(_p(self.nodes), n, float(scale), float(offset), 1 if repeat else 0,
                                             0 if build == "fold" else 1, self.n_tri, _p(self.tri))
        if not self.handle:
            raise ValueError("oracle scene creation failed (bad grid or degenerate chart)")
        self.n_leaf = n - 1

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            lib().orc_scene_destroy(h)
            self.handle = None

    def triangle_info(self, k):
        c = ctypes
        rl, nr = c.c_int(), c.c_int()
        roots = np.zeros(8, dtype=np.int64)
        aabb = np.zeros(6)
        hr = np.zeros(2)
        tol = c.c_double()
        lib().orc_triangle_info(self.handle, k, c.byref(rl), c.byref(nr), _p(roots, _I64), _p(aabb), _p(hr),
                                c.byref(tol))
        return {"root_level": rl.value, "roots": roots[:2 * nr.value].reshape(-1, 2), "aabb": aabb.reshape(2, 3),
                "h_range": hr, "residual_tolerance": tol.value}

    def intersect(self, o, d, tmin=None, tmax=None, bound="both", slab_max_level=1, leaf="newton", threads=0,
                  stats=False):
        o, d = _f64(o).reshape(-1, 3), _f64(d).reshape(-1, 3)
        n = o.shape[0]
        tmin = np.zeros(n) if tmin is None else _f64(np.broadcast_to(tmin, (n,)))
        tmax = np.full(n, np.inf) if tmax is None else _f64(np.broadcast_to(tmax, (n,)))
        out = {"t": np.empty(n), "st": np.empty((n, 2)), "bary": np.empty((n, 2)), "ng": np.empty((n, 3)),
               "prim": np.empty(n, dtype=np.int32), "texel": np.empty((n, 2), dtype=np.int64)}
        st = np.empty((n, N_STATS), dtype=np.int64) if stats else None
        lib().orc_scene_intersect(self.handle, n, _p(o), _p(d), _p(tmin), _p(tmax), BOUND[bound],
                                  int(slab_max_level), LEAF[leaf], int(threads), _p(out["t"]), _p(out["st"]),
                                  _p(out["bary"]), _p(out["ng"]), _p(out["prim"], _I32), _p(out["texel"], _I64),
                                  _p(st, _I64))
        if stats:
            out["stats"] = {k: st[:, q] for q, k in enumerate(STAT_NAMES)}
            out["stats"]["node_tests_per_level"] = st[:, len(STAT_NAMES):]
        return out

    def records(self, o, d, tmin=None, tmax=None, bound="both", slab_max_level=1, leaf="newton"):
        """Every node test of the rays' closest-hit walks as NodeTestRecord
        (record_node_test, displaced_intersector.cpp:159-172), ray by ray in
        walk order: a structured array of NODE_RECORD."""
        o, d = _f64(o).reshape(-1, 3), _f64(d).reshape(-1, 3)
        n = o.shape[0]
        tmin = np.zeros(n) if tmin is None else _f64(np.broadcast_to(tmin, (n,)))
        tmax = np.full(n, np.inf) if tmax is None else _f64(np.broadcast_to(tmax, (n,)))
        count = np.zeros(1, dtype=np.int64)
        args = (self.handle, n, _p(o), _p(d), _p(tmin), _p(tmax), BOUND[bound], int(slab_max_level), LEAF[leaf])
        lib().orc_scene_records(*args, 0, None, _p(count, _I64))
        out = np.zeros(int(count[0]), dtype=NODE_RECORD)
        lib().orc_scene_records(*args, len(out), out.ctypes.data_as(_VP), _p(count, _I64))
        return out

    def occluded(self, o, d, tmin=None, tmax=None, bound="both", slab_max_level=1, leaf="newton", threads=0):
        o, d = _f64(o).reshape(-1, 3), _f64(d).reshape(-1, 3)
        n = o.shape[0]
        tmin = np.zeros(n) if tmin is None else _f64(np.broadcast_to(tmin, (n,)))
        tmax = np.full(n, np.inf) if tmax is None else _f64(np.broadcast_to(tmax, (n,)))
        occ = np.empty(n, dtype=np.uint8)
        lib().orc_scene_occluded(self.handle, n, _p(o), _p(d), _p(tmin), _p(tmax), BOUND[bound],
                                 int(slab_max_level), LEAF[leaf], int(threads), _p(occ, _U8))
        return occ.astype(bool)

    def hit_record(self, k, u, v, d, uv_scale=(1.0, 1.0), uv_offset=(0.0, 0.0), flip_v=False, twosided=False):
        """The renderer's hit record at barycentrics (u, v) of triangle k: dict of p, ng, dpdu, dpdv, uv, t, b."""
        out = np.empty(20)
        m = np.array([uv_scale[0], uv_scale[1], uv_offset[0], uv_offset[1], 1.0 if flip_v else 0.0])
        lib().orc_hit_record(self.handle, int(k), float(u), float(v), _p(_f64(d)), _p(m), 1 if twosided else 0,
                             _p(out))
        return {"p": out[0:3], "ng": out[3:6], "dpdu": out[6:9], "dpdv": out[9:12], "uv": out[12:14],
                "t": out[14:17], "b": out[17:20]}

    def adjoint(self, o, d, dl_dt, dl_dn=None, tmin=None, tmax=None, threads=0):
        """Gradient of sum_r dl_dt[r] t_r + dl_dn[r] . ng_r with respect to
        the closed grid's node values (nodes x nodes), by the implicit-function
        theorem at each closest hit's double root; also that root's t and ng."""
        o, d = _f64(o).reshape(-1, 3), _f64(d).reshape(-1, 3)
        n = o.shape[0]
        tmin = np.zeros(n) if tmin is None else _f64(np.broadcast_to(tmin, (n,)))
        tmax = np.full(n, np.inf) if tmax is None else _f64(np.broadcast_to(tmax, (n,)))
        lt = _f64(np.broadcast_to(dl_dt, (n,)))
        ln = None if dl_dn is None else _f64(np.broadcast_to(dl_dn, (n, 3)))
        m = self.nodes.shape[0]
        grad = np.zeros((m, m))
        t = np.empty(n)
        ng = np.empty((n, 3))
        lib().orc_scene_adjoint(self.handle, n, _p(o), _p(d), _p(tmin), _p(tmax), _p(lt),
                                _p(ln) if ln is not None else None, int(threads), _p(grad), _p(t), _p(ng))
        return grad, t, ng

    @staticmethod
    def _hit(ok, h):
        if not ok:
            return None
        return {"t": h[0], "st": h[1:3].copy(), "bary": h[3:5].copy(), "ng": h[5:8].copy(),
                "texel": (int(h[8]), int(h[9]))}

    def triangle_intersect(self, k, o, d, tmin=0.0, tmax=np.inf, bound="both", slab_max_level=1, leaf="newton"):
        h = np.zeros(10)
        st = np.zeros(N_STATS, dtype=np.int64)
        ok = lib().orc_triangle_intersect(self.handle, k, _p(_f64(o)), _p(_f64(d)), tmin, tmax, BOUND[bound],
                                          slab_max_level, LEAF[leaf], _p(h), _p(st, _I64))
        out = dict(zip(STAT_NAMES, st.tolist()))
        out["node_tests_per_level"] = st[len(STAT_NAMES):]
        return self._hit(ok, h), out

    def leaf_newton(self, k, o, d, i, j, tmin=0.0, tmax=np.inf):
        h = np.zeros(10)
        sub = np.zeros(1, dtype=np.int64)
        ok = lib().orc_leaf_newton(self.handle, k, _p(_f64(o)), _p(_f64(d)), tmin, tmax, i, j, _p(h), _p(sub, _I64))
        return self._hit(ok, h), int(sub[0])

    def newton_from(self, k, o, d, i, j, s, t, tmin=0.0, tmax=np.inf):
        h = np.zeros(10)
        ok = lib().orc_newton_from(self.handle, k, _p(_f64(o)), _p(_f64(d)), tmin, tmax, i, j, s, t, _p(h))
        return self._hit(ok, h)

    def leaf_two_triangle(self, k, o, d, i, j, tmin=0.0, tmax=np.inf):
        h = np.zeros(10)
        ok = lib().orc_leaf_two_triangle(self.handle, k, _p(_f64(o)), _p(_f64(d)), tmin, tmax, i, j, _p(h))
        return self._hit(ok, h)

    def node_bound(self, k, level, i, j):
        box, m, c, rect, node = np.zeros(6), np.zeros(6), np.zeros(2), np.zeros(4), np.zeros(8)
        lib().orc_node_bound(self.handle, k, level, i, j, _p(box), _p(m), _p(c), _p(rect), _p(node))
        return {"box": box.reshape(2, 3), "m": m.reshape(3, 2), "centre": c, "rect": rect.reshape(2, 2),
                "node": dict(zip(CHANNELS, node))}

    def node_interval(self, k, level, i, j, o, d, tmin=0.0, tmax=np.inf, mode="both"):
        t01 = np.zeros(2)
        ok = lib().orc_node_interval(self.handle, k, level, i, j, _p(_f64(o)), _p(_f64(d)), tmin, tmax,
                                     BOUND[mode], _p(t01))
        return bool(ok), t01

    def tangent_frame(self, k):
        """(origin q0, to_tangent 3x3) of triangle k."""
        o, m = np.zeros(3), np.zeros(9)
        lib().orc_tangent_frame(self.handle, int(k), _p(o), _p(m))
        return o, m.reshape(3, 3)

    def displaced_position(self, k, s, t):
        S = np.zeros(3)
        lib().orc_displaced_position(self.handle, k, s, t, _p(S))
        return S

    def surface_points(self, tri, st):
        tri = np.ascontiguousarray(tri, dtype=np.int32)
        st = _f64(st).reshape(-1, 2)
        n = st.shape[0]
        S, nn = np.empty((n, 3)), np.empty((n, 3))
        lib().orc_surface_points(self.handle, n, _p(tri, _I32), _p(st), _p(S), _p(nn))
        return S, nn

    def field(self, xy):
        xy = _f64(xy).reshape(-1, 2)
        n = xy.shape[0]
        h, g = np.empty(n), np.empty((n, 2))
        lib().orc_field_eval(self.handle, n, _p(xy), _p(h), _p(g))
        return h, g

    def surface_point(self, k, u, v):
        out = np.zeros(12)
        lib().orc_surface_point(self.handle, k, u, v, _p(out))
        return {"p": out[0:3], "ng": out[3:6], "dpdu": out[6:9], "dpdv": out[9:12]}


def hardware_threads():
    return os.cpu_count() or 1
