"""dmap_amd: ray / displaced-surface intersection on AMD MI355X (gfx950).

Python view of libdmap_amd.so (C ABI: include/dmap_amd.h), mirroring the
reference's intersector interface (code/cpp/src/):

  DisplacementTile      DisplacementAsset + TaylorPyramid
                        (displacement_asset.cpp:16-35, taylor_pyramid.cpp:215-251)
  ObjectIntersector     ObjectIntersector (displaced_intersector.h:176-198);
                        .intersect = closest hit (intersect_flat :578-599),
                        .occluded = any hit (displaced_geometry.cpp:79-106)
  TraversalOptions      TraversalOptions (traversal_options.h:30-38)
  displaced_hits        differentiable closest hit (t, ng) with respect to the
                        texels (no counterpart in the reference), backed by
                        the adjoint kernel

Device memory and streams are torch's (ROCm); every computation runs in the
HIP kernels of the library. There is no CPU fallback: without the built
library, or without a GPU, the calls raise.
"""
import ctypes
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "_lib" / "libdmap_amd.so"

DMAP_OK = 0
BOUND = {"box": 0, "slab": 1, "both": 2}
LEAF = {"newton": 0, "two_triangle": 1}
PRECISION = {"fp32": 0, "fp64": 1}
WALK_START = {"shell": 0, "roots": 1}
ORDER = {"given": 0, "sort": 1, "near_first": 2}  # DMAP_ORDER_*
BUILD = {"fold": 0, "direct": 1}
CHANNELS = ("h0", "gu", "gv", "r", "ru", "rv", "h_min", "h_max")
STAT_NAMES = ("triangle_tests", "node_tests", "leaf_tests", "sub_squares", "newton_runs", "newton_iters",
              "node_tests_after_hit", "leaf_tests_after_hit")
STAT_LEVELS = 16  # DMAP_STAT_LEVELS: node tests per level, levels 0 .. 15

# Exported symbols of include/dmap_amd.h (checked by the CPU tests).
EXPORTS = ("dmap_last_error", "dmap_version", "dmap_ctx_create", "dmap_ctx_destroy", "dmap_device_alloc",
           "dmap_device_free", "dmap_memcpy", "dmap_stream_synchronize", "dmap_tile_create",
           "dmap_tile_rebuild", "dmap_tile_info", "dmap_tile_read_level", "dmap_tile_read_nodes",
           "dmap_tile_destroy", "dmap_object_create", "dmap_object_create_ex", "dmap_object_triangle_info",
           "dmap_object_destroy", "dmap_intersect_closest", "dmap_intersect_closest_noalloc", "dmap_intersect_any",
           "dmap_ray_order",
           "dmap_tile_set_texels", "dmap_object_refresh", "dmap_intersect_adjoint", "dmap_hit_record",
           "dmap_hit_record_world",
           "dmap_kernel_timing", "dmap_kernel_timing_read", "dmap_server_create", "dmap_server_submit",
           "dmap_server_poll", "dmap_server_hits", "dmap_server_copy_hits", "dmap_server_release",
           "dmap_server_idle", "dmap_server_destroy", "dmap_pathtrace")
# dmap_object_create_ex's BVH builders (DMAP_BVH_*).
BVH_BUILDERS = {"auto": 0, "host": 1, "device": 2}


class DmapError(RuntimeError):
    pass


class _Options(ctypes.Structure):
    _fields_ = [("bound", ctypes.c_int), ("slab_max_level", ctypes.c_int), ("leaf", ctypes.c_int),
                ("precision", ctypes.c_int), ("reorder", ctypes.c_int), ("polish", ctypes.c_int),
                ("walk_start", ctypes.c_int)]


_FP = ctypes.c_void_p


class _Rays(ctypes.Structure):
    _fields_ = [(n, _FP) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class _Hits(ctypes.Structure):
    _fields_ = [(n, _FP) for n in ("t", "s", "tt", "bu", "bv", "nx", "ny", "nz", "prim", "texel_i", "texel_j")]


class _UvMap(ctypes.Structure):
    _fields_ = [("scale_u", ctypes.c_double), ("scale_v", ctypes.c_double), ("offset_u", ctypes.c_double),
                ("offset_v", ctypes.c_double), ("flip_v", ctypes.c_int)]


class _RayDiffs(ctypes.Structure):
    _fields_ = [(name, ctypes.c_void_p) for name in
                ("rx_ox", "rx_oy", "rx_oz", "rx_dx", "rx_dy", "rx_dz", "ry_ox", "ry_oy", "ry_oz", "ry_dx", "ry_dy",
                 "ry_dz")]


class _Stats(ctypes.Structure):
    _fields_ = [(n, _FP) for n in STAT_NAMES] + [("node_tests_per_level", _FP), ("records", ctypes.c_void_p),
                                                 ("record_capacity", ctypes.c_int64),
                                                 ("record_count", ctypes.c_void_p)]


# dmap_node_record (include/dmap_amd.h): one node test of the walk,
# NodeTestRecord (displaced_intersector.h:77-87) with its ray, triangle,
# cell and place among the ray's node tests.
NODE_RECORD = np.dtype([("ray", np.int32), ("prim", np.int32), ("level", np.int32), ("i", np.int32),
                        ("j", np.int32), ("seq", np.int32), ("box_pass", np.int32), ("slab_pass", np.int32),
                        ("box_length", np.float32), ("slab_length", np.float32), ("both_length", np.float32),
                        ("pad_", np.float32)])


class _PtScene(ctypes.Structure):
    _fields_ = [("light_corner", ctypes.c_double * 3), ("light_u", ctypes.c_double * 3),
                ("light_v", ctypes.c_double * 3), ("light_n", ctypes.c_double * 3), ("radiance", ctypes.c_double),
                ("env", ctypes.c_double), ("albedo", ctypes.c_double)]


_lib = None


def load_library():
    """The built library; raises when it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise DmapError(f"{LIB_PATH} is not built: run python -m dmap_amd.build")
        L = ctypes.CDLL(str(LIB_PATH))
        c = ctypes
        vp = c.c_void_p
        L.dmap_last_error.restype = c.c_char_p
        L.dmap_version.restype = c.c_char_p
        L.dmap_ctx_create.argtypes = [c.c_int, c.POINTER(vp)]
        L.dmap_ctx_destroy.argtypes = [vp]
        L.dmap_tile_create.argtypes = [vp, vp, c.c_int, c.c_int, c.c_int, c.c_double, c.c_double, c.c_int, vp,
                                       c.POINTER(vp)]
        L.dmap_tile_rebuild.argtypes = [vp, vp]
        L.dmap_tile_info.argtypes = [vp, c.POINTER(c.c_int), c.POINTER(c.c_int), c.POINTER(c.c_int64)]
        L.dmap_tile_read_level.argtypes = [vp, c.c_int, vp]
        L.dmap_tile_read_nodes.argtypes = [vp, vp]
        L.dmap_tile_destroy.argtypes = [vp]
        L.dmap_object_create.argtypes = [vp, vp, c.c_int, vp, vp, vp, c.POINTER(vp)]
        L.dmap_object_create_ex.argtypes = [vp, vp, c.c_int, vp, vp, vp, c.c_int, c.POINTER(vp)]
        L.dmap_object_triangle_info.argtypes = [vp, c.c_int, c.POINTER(c.c_int), c.POINTER(c.c_int), vp, vp, vp]
        L.dmap_object_destroy.argtypes = [vp]
        L.dmap_intersect_closest.argtypes = [vp, c.POINTER(_Rays), c.c_int64, c.POINTER(_Hits), c.c_int,
                                             c.POINTER(_Options), c.POINTER(_Stats), vp]
        L.dmap_intersect_closest_noalloc.argtypes = L.dmap_intersect_closest.argtypes
        L.dmap_intersect_any.argtypes = [vp, c.POINTER(_Rays), c.c_int64, vp, c.c_int, c.POINTER(_Options), vp]
        L.dmap_ray_order.argtypes = [vp, c.POINTER(_Rays), c.c_int64, c.c_int, vp, vp]
        L.dmap_tile_set_texels.argtypes = [vp, vp, c.c_int, vp]
        L.dmap_kernel_timing.argtypes = [c.c_int]
        L.dmap_hit_record.argtypes = [vp, c.POINTER(_Rays), c.c_int64, c.POINTER(_Hits), c.POINTER(_UvMap), c.c_int,
                                      vp, vp]
        L.dmap_hit_record_world.argtypes = [vp, c.POINTER(_Rays), c.c_int64, c.POINTER(_Hits), c.POINTER(_UvMap),
                                            c.c_int, vp, c.POINTER(_RayDiffs), vp, vp]
        L.dmap_kernel_timing_read.argtypes = [c.POINTER(c.c_double), c.POINTER(c.c_int64)]
        L.dmap_object_refresh.argtypes = [vp]
        L.dmap_intersect_adjoint.argtypes = [vp, c.POINTER(_Rays), c.c_int64, vp, vp, vp, vp, vp,
                                             c.POINTER(_Options), vp]
        L.dmap_pathtrace.argtypes = [vp, vp, vp, c.c_int64, c.c_int, c.c_int, c.c_uint32, c.c_int,
                                     c.POINTER(_PtScene), c.POINTER(_Options), vp, c.POINTER(c.c_int64),
                                     c.POINTER(c.c_int64), vp]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != DMAP_OK:
        msg = load_library().dmap_last_error().decode()
        raise DmapError(f"{what} failed ({rc}): {msg}")


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise DmapError("dmap_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


@dataclass
class TraversalOptions:
    """TraversalOptions (traversal_options.h:30-38) plus the device precision."""
    bound: str = "both"
    slab_max_level: int = 1
    leaf: str = "newton"
    precision: str = "fp32"
    # The order the rays are traced in (results are per ray in every order):
    # "near_first" takes the rays whose origin lies in the object's bounding
    # box (bounces, shadow rays: the longest walks) first, longest path
    # through the box first (in half-octave classes of that length, each in
    # its given order), the rest after in their given order, so a launch's
    # drain runs on cheap rays;
    # "given" traces them as given; "sort" sorts them by direction octant and
    # origin, which pays when rays of different kinds arrive shuffled.
    # (True and False stand for "sort" and "given".)
    reorder: object = "near_first"
    polish: bool = True  # False: report the search's own root, without the double polish
    # "shell": a constant-normal triangle's walk starts at the finest block of
    # cells holding the ray's clip to its height shell (the same hits, fewer
    # node tests); "roots": at the triangle's roots, as the reference's walk
    # does, for work counters comparable with the reference's.
    walk_start: str = "shell"

    def _c(self):
        if self.walk_start not in WALK_START:
            raise DmapError(f"walk_start {self.walk_start!r}: 'shell' or 'roots'")
        order = {True: "sort", False: "given"}.get(self.reorder, self.reorder)
        if order not in ORDER:
            raise DmapError(f"reorder {self.reorder!r}: one of {sorted(ORDER)}")
        return _Options(BOUND[self.bound], int(self.slab_max_level), LEAF[self.leaf], PRECISION[self.precision],
                        ORDER[order], 1 if self.polish else 0, WALK_START[self.walk_start])


class Context:
    """One HIP device (dmap_ctx)."""

    def __init__(self, device=0):
        _torch()
        self._h = ctypes.c_void_p()
        _check(load_library().dmap_ctx_create(int(device), ctypes.byref(self._h)), "dmap_ctx_create")
        self.device = device

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.dmap_ctx_destroy(self._h)
            self._h = None


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class DisplacementTile:
    """A resident displacement tile: the closed node grid and its Taylor
    pyramid, built on the device (h = strength * texel + offset)."""

    def __init__(self, texels, strength=1.0, offset=0.0, repeat=True, build="fold", ctx=None, stream=None):
        torch = _torch()
        self.ctx = ctx or default_context()
        if isinstance(texels, torch.Tensor) and texels.is_cuda:
            t = texels.to(torch.float32).contiguous()
            on_device = 1
            ptr = t.data_ptr()
        else:
            t = np.ascontiguousarray(texels, dtype=np.float32)
            on_device = 0
            ptr = t.ctypes.data
        n = t.shape[0]
        if t.shape != (n, n):
            raise DmapError("texels must be square")
        self._h = ctypes.c_void_p()
        _check(load_library().dmap_tile_create(self.ctx._h, ctypes.c_void_p(ptr), n, on_device, 1 if repeat else 0,
                                               float(strength), float(offset), BUILD[build], _stream_ptr(stream),
                                               ctypes.byref(self._h)), "dmap_tile_create")
        self._keep = t
        ni, nl, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_int64()
        _check(_lib.dmap_tile_info(self._h, ctypes.byref(ni), ctypes.byref(nl), ctypes.byref(nb)), "dmap_tile_info")
        self.n_leaf, self.n_levels, self.device_bytes = ni.value, nl.value, nb.value
        self.strength, self.offset, self.repeat, self.build = float(strength), float(offset), bool(repeat), build

    def rebuild(self, stream=None):
        _check(_lib.dmap_tile_rebuild(self._h, _stream_ptr(stream)), "dmap_tile_rebuild")

    def set_texels(self, texels, stream=None):
        """New texel values (same n x n): closes the grid and rebuilds the
        pyramid. Objects over the tile need refresh() afterwards."""
        torch = _torch()
        if isinstance(texels, torch.Tensor) and texels.is_cuda:
            t = texels.detach().to(torch.float32).contiguous()
            on_device, ptr = 1, t.data_ptr()
        else:
            t = np.ascontiguousarray(texels, dtype=np.float32)
            on_device, ptr = 0, t.ctypes.data
        if tuple(t.shape) != (self.n_leaf, self.n_leaf):
            raise DmapError(f"texels must be {self.n_leaf} x {self.n_leaf}")
        _check(_lib.dmap_tile_set_texels(self._h, ctypes.c_void_p(ptr), on_device, _stream_ptr(stream)),
               "dmap_tile_set_texels")
        self._keep = t

    def level(self, k):
        """Channels of level k as a dict of (m, m) float32 arrays [j, i]."""
        m = self.n_leaf >> k
        out = np.empty((8, m, m), dtype=np.float32)
        _check(_lib.dmap_tile_read_level(self._h, int(k), ctypes.c_void_p(out.ctypes.data)), "dmap_tile_read_level")
        return {ch: out[q] for q, ch in enumerate(CHANNELS)}

    def nodes(self):
        n = self.n_leaf + 1
        out = np.empty((n, n), dtype=np.float32)
        _check(_lib.dmap_tile_read_nodes(self._h, ctypes.c_void_p(out.ctypes.data)), "dmap_tile_read_nodes")
        return out

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.dmap_tile_destroy(self._h)
            self._h = None


class ObjectIntersector:
    """Base triangles over one tile (ObjectIntersector). triangles: (n, 24)
    float64 rows p0 p1 p2 (9), vertex normals m0 m1 m2 (9), uvs t0 t1 t2 (6).
    bvh: the top-level BVH's builder, "auto" (the device's from 16,384
    triangles), "host" or "device"."""

    def __init__(self, tile, triangles, ctx=None, bvh="auto"):
        self.tile = tile
        self.ctx = ctx or tile.ctx
        tri = np.ascontiguousarray(triangles, dtype=np.float64).reshape(-1, 24)
        self.n_tri = tri.shape[0]
        pos = np.ascontiguousarray(tri[:, 0:9])
        nrm = np.ascontiguousarray(tri[:, 9:18])
        uvs = np.ascontiguousarray(tri[:, 18:24])
        self._h = ctypes.c_void_p()
        if bvh not in BVH_BUILDERS:
            raise DmapError(f"bvh must be one of {sorted(BVH_BUILDERS)}")
        _check(load_library().dmap_object_create_ex(self.ctx._h, tile._h, self.n_tri,
                                                    ctypes.c_void_p(pos.ctypes.data), ctypes.c_void_p(nrm.ctypes.data),
                                                    ctypes.c_void_p(uvs.ctypes.data), BVH_BUILDERS[bvh],
                                                    ctypes.byref(self._h)), "dmap_object_create")

    def triangle_info(self, k):
        rl, nr = ctypes.c_int(), ctypes.c_int()
        roots = np.zeros(8, dtype=np.int32)
        aabb = np.zeros(6, dtype=np.float32)
        hr = np.zeros(2, dtype=np.float32)
        _check(_lib.dmap_object_triangle_info(self._h, int(k), ctypes.byref(rl), ctypes.byref(nr),
                                              ctypes.c_void_p(roots.ctypes.data), ctypes.c_void_p(aabb.ctypes.data),
                                              ctypes.c_void_p(hr.ctypes.data)), "dmap_object_triangle_info")
        return {"root_level": rl.value, "roots": roots[:2 * nr.value].reshape(-1, 2), "aabb": aabb.reshape(2, 3),
                "h_range": hr}

    @staticmethod
    def _rays(rays):
        torch = _torch()
        if not (isinstance(rays, torch.Tensor) and rays.is_cuda and rays.dtype == torch.float32
                and rays.dim() == 2 and rays.shape[0] == 8 and rays.is_contiguous()):
            raise DmapError("rays must be a contiguous (8, n) float32 device tensor: o xyz, d xyz, tmin, tmax")
        p = [rays[q].data_ptr() for q in range(8)]
        return _Rays(*p), rays.shape[1]

    def alloc_hits(self, n):
        torch = _torch()
        return {"f": torch.empty((8, n), dtype=torch.float32, device="cuda"),
                "i": torch.empty((3, n), dtype=torch.int32, device="cuda")}

    def intersect(self, rays, options=None, prim_mask=-1, stats=False, out=None, stream=None, records=0):
        """Closest hit of each ray; returns device tensors t, st, bary, ng,
        prim, texel (and the per-ray work counters when stats is set). With
        records = capacity > 0 (and stats), also every node test of the
        walks (TraversalStats::records): res["records"], a host array of
        NODE_RECORD (the first `capacity` appended; rays interleave, each
        ray's in seq order), and res["record_count"], how many were made."""
        torch = _torch()
        r, n = self._rays(rays)
        hits = out or self.alloc_hits(n)
        F, I = hits["f"], hits["i"]
        h = _Hits(*[F[q].data_ptr() for q in range(8)], *[I[q].data_ptr() for q in range(3)])
        st = None
        stats_t = None
        if records and not stats:
            raise DmapError("node-test records come with the work counters: set stats=True")
        if stats:
            stats_t = torch.zeros((len(STAT_NAMES) + STAT_LEVELS, n), dtype=torch.int32, device="cuda")
            st = _Stats(*[stats_t[q].data_ptr() for q in range(len(STAT_NAMES) + 1)])
            if records:
                rec_t = torch.zeros((int(records), NODE_RECORD.itemsize // 4), dtype=torch.int32, device="cuda")
                count_t = torch.zeros(1, dtype=torch.int64, device="cuda")
                st.records, st.record_capacity, st.record_count = rec_t.data_ptr(), int(records), count_t.data_ptr()
        opt = (options or TraversalOptions())._c()
        _check(_lib.dmap_intersect_closest(self._h, ctypes.byref(r), n, ctypes.byref(h), int(prim_mask),
                                           ctypes.byref(opt), ctypes.byref(st) if st is not None else None,
                                           _stream_ptr(stream)), "dmap_intersect_closest")
        res = {"t": F[0], "st": F[1:3], "bary": F[3:5], "ng": F[5:8], "prim": I[0], "texel": I[1:3]}
        if stats:
            res["stats"] = {k: stats_t[q] for q, k in enumerate(STAT_NAMES)}
            res["stats"]["node_tests_per_level"] = stats_t[len(STAT_NAMES):]
        if records:
            count = int(count_t.item())
            res["record_count"] = count
            res["records"] = rec_t[:min(count, int(records))].cpu().numpy().view(NODE_RECORD).reshape(-1)
        return res

    def ray_order(self, rays, order="near_first", stream=None):
        """The order a trace of these rays under TraversalOptions(reorder=
        order) takes them in (dmap_ray_order): an int32 device tensor of the
        ray indices in trace order (the identity under 4,096 rays)."""
        torch = _torch()
        if order not in ORDER:
            raise DmapError(f"order {order!r}: one of {sorted(ORDER)}")
        r, n = self._rays(rays)
        perm = torch.empty(n, dtype=torch.int32, device="cuda")
        _check(_lib.dmap_ray_order(self._h, ctypes.byref(r), n, ORDER[order], ctypes.c_void_p(perm.data_ptr()),
                                   _stream_ptr(stream)), "dmap_ray_order")
        return perm

    def pathtrace(self, cam_o, cam_d, spp, depth, light, radiance, env, albedo, seed=0, strategy="mis",
                  options=None, stream=None):
        """Wavefront path tracing of this object on the device
        (dmap_pathtrace): cam_o, cam_d (3, n_pix) float64 device tensors
        (copied to row-major if they are not);
        light = (corner, edge_u, edge_v, unit normal of the emitting side).
        Returns (image (n_pix,) float64 device tensor of mean radiance,
        closest-hit rays traced, any-hit rays traced)."""
        torch = _torch()
        for name, x in (("cam_o", cam_o), ("cam_d", cam_d)):
            if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float64 and x.dim() == 2
                    and x.shape[0] == 3):
                raise DmapError(f"{name} must be a (3, n_pix) float64 device tensor")
        cam_o, cam_d = cam_o.contiguous(), cam_d.contiguous()
        if cam_o.shape != cam_d.shape:
            raise DmapError("cam_o and cam_d differ in shape")
        n_pix = cam_o.shape[1]
        strategies = {"mis": 0, "bsdf": 1}
        if strategy not in strategies:
            raise DmapError(f"strategy {strategy!r}: 'mis' or 'bsdf'")
        c, u, v, n = (np.asarray(x, dtype=np.float64).reshape(3) for x in light)
        sc = _PtScene((ctypes.c_double * 3)(*c), (ctypes.c_double * 3)(*u), (ctypes.c_double * 3)(*v),
                      (ctypes.c_double * 3)(*n), float(radiance), float(env), float(albedo))
        image = torch.empty(n_pix, dtype=torch.float64, device=cam_o.device)
        n_closest, n_shadow = ctypes.c_int64(), ctypes.c_int64()
        opt = (options or TraversalOptions())._c()
        _check(_lib.dmap_pathtrace(self._h, ctypes.c_void_p(cam_o.data_ptr()), ctypes.c_void_p(cam_d.data_ptr()),
                                   n_pix, int(spp), int(depth), int(seed) & 0xFFFFFFFF, strategies[strategy],
                                   ctypes.byref(sc), ctypes.byref(opt), ctypes.c_void_p(image.data_ptr()),
                                   ctypes.byref(n_closest), ctypes.byref(n_shadow), _stream_ptr(stream)),
               "dmap_pathtrace")
        return image, n_closest.value, n_shadow.value

    def hit_record(self, rays, hits, uv_scale=(1.0, 1.0), uv_offset=(0.0, 0.0), flip_v=False, twosided=False,
                   stream=None, object_to_world=None, diffs=None):
        """The renderer's hit record (DisplacedGeometry::surface_point +
        compute_intersection's frame) for the result of intersect(rays):
        dict of device tensors p (3, n), ng, dpdu, dpdv, uv (2, n), t, b.
        Object space, or world space with object_to_world ((3, 4) array:
        world = M x + T). diffs: the rays' world-space differentials, a
        (12, n) fp32 device tensor (rx o xyz, rx d xyz, ry o xyz, ry d xyz),
        adding dpdx, dpdy (3, n) and dudx, dvdx, dudy, dvdy (n,)."""
        torch = _torch()
        r, n = self._rays(rays)
        if object_to_world is not None or diffs is not None:
            return self._hit_record_world(r, n, hits, uv_scale, uv_offset, flip_v, twosided, stream,
                                          object_to_world, diffs)
        h = _Hits(hits["t"].data_ptr(), hits["st"][0].data_ptr(), hits["st"][1].data_ptr(),
                  hits["bary"][0].data_ptr(), hits["bary"][1].data_ptr(), hits["ng"][0].data_ptr(),
                  hits["ng"][1].data_ptr(), hits["ng"][2].data_ptr(), hits["prim"].data_ptr(),
                  hits["texel"][0].data_ptr(), hits["texel"][1].data_ptr())
        m = _UvMap(float(uv_scale[0]), float(uv_scale[1]), float(uv_offset[0]), float(uv_offset[1]),
                   1 if flip_v else 0)
        rec = torch.empty((20, n), dtype=torch.float32, device="cuda")
        _check(_lib.dmap_hit_record(self._h, ctypes.byref(r), n, ctypes.byref(h), ctypes.byref(m),
                                    1 if twosided else 0, ctypes.c_void_p(rec.data_ptr()), _stream_ptr(stream)),
               "dmap_hit_record")
        return {"p": rec[0:3], "ng": rec[3:6], "dpdu": rec[6:9], "dpdv": rec[9:12], "uv": rec[12:14],
                "t": rec[14:17], "b": rec[17:20]}

    def _hit_record_world(self, r, n, hits, uv_scale, uv_offset, flip_v, twosided, stream, object_to_world, diffs):
        torch = _torch()
        h = _Hits(hits["t"].data_ptr(), hits["st"][0].data_ptr(), hits["st"][1].data_ptr(),
                  hits["bary"][0].data_ptr(), hits["bary"][1].data_ptr(), hits["ng"][0].data_ptr(),
                  hits["ng"][1].data_ptr(), hits["ng"][2].data_ptr(), hits["prim"].data_ptr(),
                  hits["texel"][0].data_ptr(), hits["texel"][1].data_ptr())
        m = _UvMap(float(uv_scale[0]), float(uv_scale[1]), float(uv_offset[0]), float(uv_offset[1]),
                   1 if flip_v else 0)
        xf = None
        if object_to_world is not None:
            xf = np.ascontiguousarray(np.asarray(object_to_world, dtype=np.float64).reshape(3, 4))
        dp = None
        if diffs is not None:
            if not (isinstance(diffs, torch.Tensor) and diffs.is_cuda and diffs.dtype == torch.float32
                    and tuple(diffs.shape) == (12, n) and diffs.is_contiguous()):
                raise DmapError(f"diffs must be a contiguous (12, {n}) float32 device tensor")
            dp = _RayDiffs(*[diffs[q].data_ptr() for q in range(12)])
        planes = 30 if diffs is not None else 20
        rec = torch.empty((planes, n), dtype=torch.float32, device="cuda")
        _check(_lib.dmap_hit_record_world(self._h, ctypes.byref(r), n, ctypes.byref(h), ctypes.byref(m),
                                          1 if twosided else 0,
                                          ctypes.c_void_p(xf.ctypes.data) if xf is not None else None,
                                          ctypes.byref(dp) if dp is not None else None,
                                          ctypes.c_void_p(rec.data_ptr()), _stream_ptr(stream)),
               "dmap_hit_record_world")
        out = {"p": rec[0:3], "ng": rec[3:6], "dpdu": rec[6:9], "dpdv": rec[9:12], "uv": rec[12:14],
               "t": rec[14:17], "b": rec[17:20]}
        if diffs is not None:
            out.update({"dpdx": rec[20:23], "dpdy": rec[23:26], "dudx": rec[26], "dvdx": rec[27], "dudy": rec[28],
                        "dvdy": rec[29]})
        return out

    def refresh(self):
        """Triangle bounds and BVH from the tile's current pyramid."""
        _check(_lib.dmap_object_refresh(self._h), "dmap_object_refresh")

    def adjoint(self, rays, dl_dt=None, dl_dn=None, grad=None, options=None, stream=None):
        """dL/d(texel) (n_leaf x n_leaf float64 device tensor, added into
        `grad` when given) for dL/dt (n,) and dL/dng (3, n) fp32 device
        tensors; the closest hits are recomputed."""
        torch = _torch()
        r, n = self._rays(rays)
        m = self.tile.n_leaf
        if grad is None:
            grad = torch.zeros((m, m), dtype=torch.float64, device="cuda")
        elif not (grad.dtype == torch.float64 and tuple(grad.shape) == (m, m) and grad.is_contiguous()):
            raise DmapError(f"grad must be a contiguous {m} x {m} float64 device tensor")

        def vec(x, shape):
            if x is None:
                return None
            if not (isinstance(x, torch.Tensor) and x.is_cuda and x.dtype == torch.float32
                    and tuple(x.shape) == shape and x.is_contiguous()):
                raise DmapError(f"upstream gradients must be contiguous float32 device tensors of shape {shape}")
            return x

        lt = vec(dl_dt, (n,))
        ln = vec(dl_dn, (3, n))
        ptr = lambda x: ctypes.c_void_p(x.data_ptr()) if x is not None else None  # noqa: E731
        opt = (options or TraversalOptions())._c()
        _check(_lib.dmap_intersect_adjoint(self._h, ctypes.byref(r), n, ptr(lt),
                                           ptr(ln[0]) if ln is not None else None,
                                           ptr(ln[1]) if ln is not None else None,
                                           ptr(ln[2]) if ln is not None else None, ctypes.c_void_p(grad.data_ptr()),
                                           ctypes.byref(opt), _stream_ptr(stream)), "dmap_intersect_adjoint")
        return grad

    def occluded(self, rays, options=None, prim_mask=-1, stream=None):
        torch = _torch()
        r, n = self._rays(rays)
        occ = torch.empty(n, dtype=torch.uint8, device="cuda")
        opt = (options or TraversalOptions())._c()
        _check(_lib.dmap_intersect_any(self._h, ctypes.byref(r), n, ctypes.c_void_p(occ.data_ptr()), int(prim_mask),
                                       ctypes.byref(opt), _stream_ptr(stream)), "dmap_intersect_any")
        return occ

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.dmap_object_destroy(self._h)
            self._h = None


class Server:
    """Continuous-batching closest-hit server over resident objects
    (dmap_server_*): submit requests, poll completions with latencies,
    read and release their hits."""

    def __init__(self, slots=3, step_rays=1 << 20, hit_pool_bytes=4 << 30, ctx=None):
        L = load_library()
        c = ctypes
        vp = c.c_void_p
        L.dmap_server_create.argtypes = [vp, c.c_int, c.c_int64, c.c_int64, c.POINTER(vp)]
        L.dmap_server_submit.argtypes = [vp, vp, c.POINTER(_Rays), c.c_int64, c.c_uint64]
        L.dmap_server_poll.argtypes = [vp, c.c_int, vp, vp, c.POINTER(c.c_int)]
        L.dmap_server_hits.argtypes = [vp, c.c_uint64, c.POINTER(_Hits)]
        L.dmap_server_copy_hits.argtypes = [vp, c.c_uint64, c.POINTER(_Hits), vp]
        L.dmap_server_release.argtypes = [vp, c.c_uint64]
        L.dmap_server_idle.argtypes = [vp, c.POINTER(c.c_int)]
        L.dmap_server_destroy.argtypes = [vp]
        self.ctx = ctx or default_context()
        self._h = ctypes.c_void_p()
        _check(L.dmap_server_create(self.ctx._h, int(slots), int(step_rays), int(hit_pool_bytes),
                                    ctypes.byref(self._h)), "dmap_server_create")
        self._keep = {}
        self._tags = np.zeros(4096, dtype=np.uint64)
        self._lat = np.zeros(4096, dtype=np.float64)

    def submit(self, obj, rays, tag):
        """rays: (8, n) float32 device tensor, kept alive until released."""
        r, n = ObjectIntersector._rays(rays)
        self._keep[int(tag)] = rays
        _check(_lib.dmap_server_submit(self._h, obj._h, ctypes.byref(r), n, ctypes.c_uint64(int(tag))),
               "dmap_server_submit")

    def poll(self):
        """[(tag, latency_s)] of requests completed since the last poll."""
        out = []
        while True:
            k = ctypes.c_int()
            _check(_lib.dmap_server_poll(self._h, len(self._tags), ctypes.c_void_p(self._tags.ctypes.data),
                                         ctypes.c_void_p(self._lat.ctypes.data), ctypes.byref(k)), "dmap_server_poll")
            out += [(int(self._tags[q]), float(self._lat[q])) for q in range(k.value)]
            if k.value < len(self._tags):
                return out

    def hits(self, tag, n, stream=None):
        """A completed request's hits, copied into fresh device tensors in
        the layout ObjectIntersector.intersect returns."""
        torch = _torch()
        F = torch.empty((8, n), dtype=torch.float32, device="cuda")
        I = torch.empty((3, n), dtype=torch.int32, device="cuda")
        h = _Hits(*[F[q].data_ptr() for q in range(8)], *[I[q].data_ptr() for q in range(3)])
        _check(_lib.dmap_server_copy_hits(self._h, ctypes.c_uint64(int(tag)), ctypes.byref(h), _stream_ptr(stream)),
               "dmap_server_copy_hits")
        return {"t": F[0], "st": F[1:3], "bary": F[3:5], "ng": F[5:8], "prim": I[0], "texel": I[1:3]}

    def release(self, tag):
        _check(_lib.dmap_server_release(self._h, ctypes.c_uint64(int(tag))), "dmap_server_release")
        self._keep.pop(int(tag), None)

    def idle(self):
        k = ctypes.c_int()
        _check(_lib.dmap_server_idle(self._h, ctypes.byref(k)), "dmap_server_idle")
        return bool(k.value)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value and _lib is not None:
            _lib.dmap_server_destroy(self._h)
            self._h = None


MAX_FILTER_PASSES = 64  # DisplacedGeometry::max_filter_passes (displaced_geometry.h:64)
# A rejected hit's ray resumes this far (relative) past the reported t: the
# reported t is the double-polished root rounded to fp32, which can sit just
# below the fp32 walk's own root, and resuming at t itself would find the
# same root again. Two crossings closer than this count as one.
FILTER_RESUME_REL = 1e-5


def filter_resume(t):
    """The t_min a rejected hit at t resumes from (fp32): FILTER_RESUME_REL
    past t, and at least the next float above it. Reported hits have t >= 0
    (newton_from rejects roots behind the origin, displaced_intersector.cpp:356),
    so the second term only matters at t = 0."""
    torch = _torch()
    return torch.maximum(t + t.abs() * FILTER_RESUME_REL, torch.nextafter(t, torch.full_like(t, float("inf"))))


def intersect_filtered(obj, rays, accept, options=None, max_passes=MAX_FILTER_PASSES):
    """Closest hit that `accept` takes: the renderer's filter-resume loop
    (intersect_callback, displaced_geometry.cpp:49-77) on a batch. `accept`
    maps a hit dict (as intersect returns, all rays) to a bool tensor (n,);
    rays whose hit it rejects are traced again with t_min raised to the
    rejected t (the walk takes t > t_min only; here filter_resume(t) just
    past it), up to max_passes times, after which they count as misses, as
    there. The reference counts the passes per triangle, this loop per ray
    over the object."""
    torch = _torch()
    out = {k: v.clone() for k, v in obj.intersect(rays, options=options).items()}
    cur = rays.clone()
    pending = torch.isfinite(out["t"])
    for _ in range(max_passes):
        rej = pending & ~accept(out).to(torch.bool)
        if not bool(rej.any()):
            return out
        idx = rej.nonzero().squeeze(1)
        cur[6, idx] = filter_resume(out["t"][idx])
        again = obj.intersect(cur[:, idx].contiguous(), options=options)
        for k, v in again.items():
            out[k][..., idx] = v
        pending = torch.zeros_like(pending)
        pending[idx] = torch.isfinite(again["t"])
    rej = pending & ~accept(out).to(torch.bool)
    if bool(rej.any()):
        out["t"][rej] = float("inf")
        out["prim"][rej] = -1
        for k in ("st", "bary", "ng", "texel"):
            out[k][:, rej] = 0
    return out


def occluded_filtered(obj, rays, accept, options=None, max_passes=MAX_FILTER_PASSES):
    """Any hit that `accept` takes (occluded_callback, displaced_geometry.cpp:79-106)."""
    res = intersect_filtered(obj, rays, accept, options=options, max_passes=max_passes)
    return _torch().isfinite(res["t"])


def kernel_timing(enable):
    """Record HIP events around every traversal kernel from now on (or stop)."""
    _check(load_library().dmap_kernel_timing(1 if enable else 0), "dmap_kernel_timing")


def kernel_timing_read():
    """(summed traversal-kernel ms, launches) since the last read."""
    ms, n = ctypes.c_double(), ctypes.c_int64()
    _check(load_library().dmap_kernel_timing_read(ctypes.byref(ms), ctypes.byref(n)), "dmap_kernel_timing_read")
    return ms.value, n.value


def displaced_hits(texels, tile, obj, rays, options=None):
    """Differentiable closest hit: (t (n,), ng (3, n)) of the rays against
    `obj` with the tile's texels set to `texels` ((n_leaf, n_leaf) float32
    device tensor). Backward runs the adjoint kernel, which recomputes the
    hits; misses (t = inf) carry no gradient."""
    return _DisplacedHits.apply(texels, tile, obj, rays, options)


def _hits_fn():
    import torch

    class DisplacedHitsFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, texels, tile, obj, rays, options):
            tile.set_texels(texels)
            obj.refresh()
            res = obj.intersect(rays, options=options)
            ctx.save_for_backward(texels, rays)
            ctx.tile, ctx.obj, ctx.options = tile, obj, options
            return res["t"].clone(), res["ng"].clone()

        @staticmethod
        def backward(ctx, g_t, g_n):
            texels, rays = ctx.saved_tensors
            tile, obj = ctx.tile, ctx.obj
            # The tile may have been set to other texels since the forward pass.
            tile.set_texels(texels)
            obj.refresh()
            lt = g_t.contiguous().to(torch.float32) if g_t is not None else None
            ln = g_n.contiguous().to(torch.float32) if g_n is not None else None
            grad = obj.adjoint(rays, lt, ln, options=ctx.options)
            return grad.to(texels.dtype), None, None, None, None

    return DisplacedHitsFn


class _LazyFn:
    _fn = None

    def apply(self, *args):
        if _LazyFn._fn is None:
            _LazyFn._fn = _hits_fn()
        return _LazyFn._fn.apply(*args)


_DisplacedHits = _LazyFn()
