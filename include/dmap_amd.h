/*
 * dmap_amd: ray / displaced-surface intersection on AMD Instinct MI355X
 * (gfx950). C ABI of libdmap_amd.so.
 *
 * Each entry point replaces a reference interface (paths relative to the
 * reference repository, code/cpp/src/):
 *   dmap_tile_create      DisplacementAsset ctor + TaylorPyramid ctor
 *                         (displacement_asset.cpp:16-35, taylor_pyramid.cpp:215-251,
 *                         texture_grid.cpp:118-132)
 *   dmap_object_create    ObjectIntersector ctor (displaced_intersector.cpp:466-479)
 *                         over BaseTriangle (displaced_surface.cpp:54-91)
 *   dmap_intersect_closest ObjectIntersector::intersect / intersect_flat
 *                         (displaced_intersector.h:193-197, displaced_intersector.cpp:556-599)
 *                         and, with prim_mask = k, DisplacedIntersector::intersect
 *                         (displaced_intersector.h:137-138) on triangle k alone
 *   dmap_intersect_any    the occluded callback (displaced_geometry.cpp:79-106)
 *   dmap_hit_record       DisplacedGeometry::surface_point + compute_intersection's
 *                         frame (displaced_geometry.cpp:126-183)
 *   dmap_hit_record_world the same, then compute_intersection's world-space end,
 *                         transform_it + compute_uv_partials (displaced_geometry.cpp:180-183)
 * The adjoint, tile texel updates, the serving loop, the wavefront path
 * tracing call, kernel timing and the device-memory helpers have no
 * counterpart in the reference (see their declarations).
 *
 * Conventions (as the reference's, displaced_intersector.h:107-114):
 *   rays are in object space, d need not be unit length, t is in units of
 *   d, hits satisfy t_min < t < t_max (open interval); a miss writes
 *   t = +inf and prim = -1. st are tile coordinates, (bu, bv) the base
 *   triangle's barycentric weights of p1 and p2, n the unit Su x Sv
 *   (unoriented), (texel_i, texel_j) the leaf texel.
 *
 * Memory: ray and hit buffers are caller-owned DEVICE pointers (SoA, fp32),
 * used asynchronously on the caller's HIP stream (NULL = the null stream).
 * Tiles and objects own their device memory and stay resident until
 * destroyed.
 *
 * Errors: every call returns DMAP_OK or an error code; dmap_last_error()
 * gives the calling thread's last message. Nothing aborts the process.
 * Threading: creation calls may run concurrently on different objects;
 * queries on one object may run concurrently from several host threads on
 * distinct streams (the reference's queries are const and reentrant).
 */
#ifndef DMAP_AMD_H
#define DMAP_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMAP_OK 0
#define DMAP_ERR_INVALID 1   /* bad argument (the reference ASSERTs) */
#define DMAP_ERR_HIP 2       /* a HIP runtime call failed */
#define DMAP_ERR_MEMORY 3    /* device allocation failed */

typedef struct dmap_ctx dmap_ctx;
typedef struct dmap_tile dmap_tile;
typedef struct dmap_object dmap_object;

/* dmap_options.reorder */
#define DMAP_ORDER_GIVEN 0
#define DMAP_ORDER_SORT 1
#define DMAP_ORDER_NEAR_FIRST 2

/* TraversalOptions (traversal_options.h:30-38). */
typedef struct
{
    int bound;          /* 0 box, 1 slab, 2 both (default) */
    int slab_max_level; /* the slab is read at levels <= this (default 1) */
    int leaf;           /* 0 certified Newton (default), 1 two-triangle */
    int precision;      /* 0 fp32 (default), 1 the same solve in fp64 on the device */
    int reorder;        /* the order the rays are traced in (results are per ray in every
                           order; batches under 4,096 rays go as given):
                           DMAP_ORDER_NEAR_FIRST (default): the rays whose origin lies in
                           the object's bounding box (bounces, shadow rays: the longest
                           walks) first, longest path through the box first (in classes
                           of half an octave of that length), the rest after in their
                           given order;
                           DMAP_ORDER_GIVEN: as given;
                           DMAP_ORDER_SORT: sorted by direction octant and origin, which
                           pays when rays of different kinds arrive shuffled */
    int polish;         /* 1 (default): the reported root is polished by two Newton steps in
                           double; 0 reports the search's own root (precision's arithmetic),
                           evaluated in double, to measure the fp32 solve's error */
    int walk_start;     /* 0 (default): a constant-normal triangle's walk starts at the finest
                           2 x 2 block of cells that holds the ray's clip to the triangle's
                           height shell (the same hits, fewer node tests); 1: at the
                           triangle's root cells, as the reference's walk does
                           (DisplacedIntersector::intersect, displaced_intersector.cpp:
                           175-237), for work counters comparable with the reference's */
} dmap_options;

/* Device pointers, n entries each. */
typedef struct
{
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
} dmap_rays;

/* Device pointers, n entries each; t, s, tt, prim, texel_i and texel_j
 * are required, bu, bv, nx, ny, nz may be NULL. */
typedef struct
{
    float *t, *s, *tt, *bu, *bv, *nx, *ny, *nz;
    int32_t *prim, *texel_i, *texel_j;
} dmap_hits;

/* Per-ray work counters, TraversalStats (displaced_intersector.h:89-105);
 * device pointers, any may be NULL. node_tests_after_hit and
 * leaf_tests_after_hit count, in each triangle's walk that found a hit, the
 * tests made after its last hit (displaced_intersector.cpp:181, 222-235).
 * node_tests_per_level holds DMAP_STAT_LEVELS planes of n counts,
 * level-major ([level * n + ray]); a triangle's root level is at most 15,
 * so no node lies above the last plane. */
#define DMAP_STAT_LEVELS 16

/* One node test as the walk made it, NodeTestRecord (displaced_intersector.h:
 * 77-87, record_node_test at displaced_intersector.cpp:159-172): the box's
 * and the slab's decisions with the walk's current t_max, and the lengths
 * of the box's, the slab's and both's intervals with t_max = inf (-1 when
 * that test fails). The slab is read at every level here, whatever the
 * options' slab_max_level, as the reference's record does. The cell
 * (level, i, j), the ray's index in the call, the triangle and the test's
 * place among the ray's node tests (seq, from 0) say which test it was;
 * records of different rays interleave in the array. */
typedef struct
{
    int32_t ray, prim, level, i, j, seq;
    int32_t box_pass, slab_pass;
    float box_length, slab_length, both_length, pad_;
} dmap_node_record;

/* records (device array of record_capacity entries) and record_count (a
 * device uint64 the caller zeroes) may be NULL together; when set, every
 * node test appends a record and adds one to *record_count, records past
 * the capacity are counted and not written. */
typedef struct
{
    int32_t *triangle_tests, *node_tests, *leaf_tests, *sub_squares, *newton_runs, *newton_iters;
    int32_t *node_tests_after_hit, *leaf_tests_after_hit, *node_tests_per_level;
    dmap_node_record *records;
    int64_t record_capacity;
    uint64_t *record_count;
} dmap_ray_stats;

const char *dmap_last_error(void);
const char *dmap_version(void);

/* One HIP device. The calls' per-call workspaces come from a memory pool of
 * the context's own, which keeps the device memory it has handed out (up to
 * the calls' largest simultaneous use) instead of returning it at every
 * synchronization; dmap_ctx_destroy waits for the device and returns it. */
int dmap_ctx_create(int device, dmap_ctx **out);
void dmap_ctx_destroy(dmap_ctx *ctx);

/* Device memory for callers without a HIP toolchain of their own (the C++
 * API in dmap_amd.hpp uses these): allocate and free on the context's
 * device, copy (kind 0 host to device, 1 device to host, 2 device to
 * device; on `stream`, synchronous when stream is NULL), wait for a stream. */
int dmap_device_alloc(dmap_ctx *ctx, int64_t bytes, void **ptr);
int dmap_device_free(dmap_ctx *ctx, void *ptr);
int dmap_memcpy(void *dst, const void *src, int64_t bytes, int kind, void *stream);
int dmap_stream_synchronize(void *stream);

/* A resident displacement tile from n x n texels (n a power of two):
 * h = strength * texel + offset (HeightGrid, displaced_surface.h:32-47),
 * repeat 1 wraps, 0 clamps; build 0 fold, 1 direct (PyramidBuild,
 * taylor_pyramid.h:41-45). texels is a host pointer, or a device pointer
 * when texels_on_device is 1. Builds the node grid and the pyramid on the
 * device, on `stream`, and waits for them. */
int dmap_tile_create(dmap_ctx *ctx, const float *texels, int n, int texels_on_device, int repeat, double strength,
                     double offset, int build, void *stream, dmap_tile **out);
/* Re-runs the pyramid build from the resident node grid (for timing). */
int dmap_tile_rebuild(dmap_tile *tile, void *stream);
/* New texel values for the tile (same size): closes the grid and rebuilds
 * the pyramid on `stream`, and waits. Objects over the tile then need
 * dmap_object_refresh (no counterpart in the reference, whose assets are
 * immutable; the differentiable path updates texels between passes). */
int dmap_tile_set_texels(dmap_tile *tile, const float *texels, int texels_on_device, void *stream);
int dmap_tile_info(const dmap_tile *tile, int *n_leaf, int *n_levels, int64_t *device_bytes);
/* The eight channels of one level (TaylorPyramid::levels, taylor_pyramid.h:51-61)
 * to host memory: out holds 8 x m x m floats, channel-major, each channel
 * row-major [j * m + i], channels h0 gu gv r ru rv h_min h_max. */
int dmap_tile_read_level(const dmap_tile *tile, int level, float *out);
/* The closed node grid ((n+1)^2 raw texel values) to host memory. */
int dmap_tile_read_nodes(const dmap_tile *tile, float *out);
void dmap_tile_destroy(dmap_tile *tile);

/* An object: n_tri (1 .. 2^30) base triangles over one tile. positions,
 * normals: 9 doubles per triangle (p0 p1 p2, m0 m1 m2); uvs: 6 doubles per
 * triangle (t0 t1 t2). Host pointers. */
int dmap_object_create(dmap_ctx *ctx, dmap_tile *tile, int n_tri, const double *positions, const double *normals,
                       const double *uvs, dmap_object **out);
/* The top-level BVH over the triangles' AABBs stands in for the reference's
 * embree scene (ObjectIntersector::build_bvh, displaced_intersector.cpp:540-554).
 * dmap_object_create_ex chooses its builder: on the host by median splits
 * (DMAP_BVH_HOST), on the device over the Morton order of the boxes' centres
 * (DMAP_BVH_DEVICE), or the device for 16,384 triangles and more
 * (DMAP_BVH_AUTO, what dmap_object_create does). The trees differ, the
 * closest hits do not (up to which of two triangles reports a hit at
 * exactly equal t). No counterpart in the reference beyond build_bvh. */
#define DMAP_BVH_AUTO 0
#define DMAP_BVH_HOST 1
#define DMAP_BVH_DEVICE 2
int dmap_object_create_ex(dmap_ctx *ctx, dmap_tile *tile, int n_tri, const double *positions, const double *normals,
                          const double *uvs, int bvh_builder, dmap_object **out);
/* Per-triangle set-up: root level, root count, roots (4 x (i, j)), the
 * object-space AABB (lo xyz, hi xyz, outward-rounded fp32) and the height
 * range over the footprint. */
int dmap_object_triangle_info(const dmap_object *obj, int k, int *root_level, int *n_roots, int32_t *roots,
                              float *aabb, float *h_range);
/* Recomputes the triangles' height ranges, AABBs and BVH from the tile's
 * current pyramid (after dmap_tile_set_texels). Synchronizes the device. */
int dmap_object_refresh(dmap_object *obj);
void dmap_object_destroy(dmap_object *obj);

/* Closest hit of n rays. prim_mask -1 tests every triangle, k >= 0 only
 * triangle k (the emitter-pdf door, displaced_area_light.cpp:174-186).
 * options and stats may be NULL. */
int dmap_intersect_closest(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits,
                           int prim_mask, const dmap_options *options, const dmap_ray_stats *stats, void *stream);

/* dmap_intersect_closest without any allocation (no counterpart in the
 * reference, which allocates nothing per ray; this is for callers that
 * cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Same arguments and
 * results; its workspace (ray-queue head, ray reorder buffers and raw
 * records) is a fixed 64 MiB device array per GPU, held from library load to process
 * exit, so one call takes at most 1,048,576 rays (larger n returns
 * DMAP_ERR_INVALID). Calls that overlap on one GPU share that array and
 * race: serialize them (one stream, or events). Kernel timing
 * (dmap_kernel_timing) does not record these calls. */
int dmap_intersect_closest_noalloc(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits,
                                   int prim_mask, const dmap_options *options, const dmap_ray_stats *stats,
                                   void *stream);

/* The order a closest-hit or any-hit call of these n rays traces them in
 * under options.reorder = reorder (DMAP_ORDER_*): perm (device memory, n
 * uint32) receives the ray indices in trace order; batches under 4,096
 * rays are traced as given (perm[k] = k). No counterpart in the reference,
 * which traces one ray per call; for callers that keep per-ray data in
 * trace order, and to check the orders. */
int dmap_ray_order(const dmap_object *obj, const dmap_rays *rays, int64_t n, int reorder, uint32_t *perm,
                   void *stream);

/* Any hit: occluded[r] = 1 when the ray meets the surface in (tmin, tmax). */
int dmap_intersect_any(const dmap_object *obj, const dmap_rays *rays, int64_t n, uint8_t *occluded, int prim_mask,
                       const dmap_options *options, void *stream);

/* The renderer's hit record (DisplacedGeometry::surface_point and the
 * geometric frame of compute_intersection, displaced_geometry.cpp:126-183),
 * in object space, for the hits of a closest-hit call (needs hits prim, bu,
 * bv): record holds 20 fp32 planes of n values, in order p xyz, ng xyz
 * (oriented along the base face's winding; towards -d when twosided), dpdu
 * xyz, dpdv xyz, uv, frame tangent xyz, frame bitangent xyz. uv_map (NULL:
 * identity) maps tile coordinates to the object's uv (DisplacedObject::
 * object_uv, scene_spec.cpp:382-388), in object space
 * (dmap_hit_record_world below adds the world transform and the ray
 * differentials). Misses get zeros. */
typedef struct dmap_uv_map {
    double scale_u, scale_v, offset_u, offset_v;
    int flip_v;
} dmap_uv_map;
int dmap_hit_record(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits,
                    const dmap_uv_map *uv_map, int twosided, float *record, void *stream);

/* The world-space half of compute_intersection (displaced_geometry.cpp:
 * 180-183: transform_it and compute_uv_partials, which live in the ks
 * renderer library, not in the reference repository) on top of
 * dmap_hit_record. object_to_world (12 host doubles, row-major 3 x 4:
 * world = M x + T, M invertible; NULL: identity) takes p and dpdu, dpdv to
 * world space, ng by M^-T (renormalized), and the frame is rebuilt there as
 * compute_intersection builds it (b = unit(ng x dpdu), t = b x ng), the
 * two-sided flip decided on the object-space ray. With diffs (the rays'
 * world-space differentials: the origins and directions of the offset rays
 * rx and ry, device fp32 SoA; NULL: none) the record gains 10 planes after
 * its 20: dp/dx xyz, dp/dy xyz, du/dx, dv/dx, du/dy, dv/dy, by the
 * tangent-plane method of PBRT (3rd ed., sec. 10.1.1): each offset ray
 * meets the plane through p with normal ng, and (du, dv) solves
 * dp = du dpdu + dv dpdv on the two axes other than ng's largest (zeros
 * where an offset ray runs parallel to the plane or the system is
 * singular). The rays and hits are the object-space ones the intersect
 * call used. */
typedef struct dmap_ray_diffs {
    const float *rx_ox, *rx_oy, *rx_oz, *rx_dx, *rx_dy, *rx_dz;
    const float *ry_ox, *ry_oy, *ry_oz, *ry_dx, *ry_dy, *ry_dz;
} dmap_ray_diffs;
int dmap_hit_record_world(const dmap_object *obj, const dmap_rays *rays, int64_t n, const dmap_hits *hits,
                          const dmap_uv_map *uv_map, int twosided, const double *object_to_world,
                          const dmap_ray_diffs *diffs, float *record, void *stream);

/* Adjoint of the closest hit (no counterpart in the reference): for the
 * loss gradients dl_dt[r] = dL/dt_r and (dl_dnx, dl_dny, dl_dnz)[r] =
 * dL/dn_r (device fp32 arrays; the normal ones all NULL for a depth-only
 * loss), ADDS dL/d(texel) into grad_texels (device, n_leaf x n_leaf
 * doubles, row-major [j][i]). The closest hits are recomputed (not kept
 * from a forward call) and the gradient follows the implicit-function
 * theorem at each hit's double root; misses contribute nothing. Newton leaf
 * only; options as dmap_intersect_closest's. */
int dmap_intersect_adjoint(const dmap_object *obj, const dmap_rays *rays, int64_t n, const float *dl_dt,
                           const float *dl_dnx, const float *dl_dny, const float *dl_dnz, double *grad_texels,
                           const dmap_options *options, void *stream);

/* Kernel timing for measurement (no counterpart in the reference): while
 * enabled, every closest/any-hit call records HIP events around its
 * traversal kernel on its stream; dmap_kernel_timing_read waits for them and
 * returns the summed kernel time (ms) and the number of launches since the
 * last read. */
int dmap_kernel_timing(int enable);
int dmap_kernel_timing_read(double *trace_ms, int64_t *launches);

/* Serving (continuous batching; no counterpart in the reference, which
 * traces one ray per call). Requests against resident objects are queued
 * and admitted (those waiting over 20 ms first, in arrival order, then by
 * fewest rays left) into the next free slot while the step's ray budget
 * (step_rays, 4 x step_rays under a backlog), its 32 requests, its 8
 * objects and the hit pool allow. A step is one persistent multi-request
 * launch (replayed from a captured HIP graph) whose workgroups are shared
 * out among the step's objects. Up to n_slots steps overlap on the device. Ray buffers stay owned by the caller and must
 * stay valid until the request completes; hits live in the server's pool
 * (44 B per ray) until released. */
typedef struct dmap_server dmap_server;
int dmap_server_create(dmap_ctx *ctx, int n_slots, int64_t step_rays, int64_t hit_pool_bytes, dmap_server **out);
int dmap_server_submit(dmap_server *server, const dmap_object *obj, const dmap_rays *rays, int64_t n, uint64_t tag);
/* Completed requests since the last poll: their tags and latencies
 * (submission to completion, seconds). */
int dmap_server_poll(dmap_server *server, int max, uint64_t *tags, double *latency_s, int *n_done);
/* The device hit arrays of a completed request (valid until released):
 * work the caller queues on them must have completed before it releases
 * the request, whose pool region a later step may then overwrite. */
int dmap_server_hits(dmap_server *server, uint64_t tag, dmap_hits *hits);
/* Copy a completed request's hits into caller-owned device arrays (NULL
 * fields are skipped), asynchronously on `stream`. The request may be
 * released right after: its pool region is reused only once the copies
 * have completed. */
int dmap_server_copy_hits(dmap_server *server, uint64_t tag, const dmap_hits *dst, void *stream);
/* Ends a completed request: its pool region returns to the server once
 * every copy queued by dmap_server_copy_hits has completed. */
int dmap_server_release(dmap_server *server, uint64_t tag);
/* idle = 1 when nothing is queued or in flight. */
int dmap_server_idle(dmap_server *server, int *idle);
void dmap_server_destroy(dmap_server *server);

/* Wavefront path tracing over one object (the path-tracing workload's
 * estimator; no entry point of the reference: there the renderer, ks's
 * SmallPT through path_trace.cpp:107-129 and direct_lighting.cpp:155-232,
 * calls the closest-hit and occluded callbacks one ray at a time, and this
 * call makes the same queries in batches through dmap_intersect_closest and
 * dmap_intersect_any). Lambertian surface of albedo `albedo`, a one-sided
 * rectangular light corner + a u + b v (a, b in [0, 1]) of radiance
 * `radiance` emitting along the unit normal light_n, and a constant
 * environment `env`. strategy 0: next-event estimation with the power
 * heuristic against the cosine-sampled BSDF; 1: BSDF sampling only.
 * Paths of at most `depth` bounces; path id = pixel * spp + sample keys the
 * random numbers with `seed`. cam_o, cam_d: device, 3 rows of n_pix doubles
 * (directions need not be unit). image: device, n_pix doubles, set to the
 * mean radiance per pixel. The counts of closest-hit and any-hit rays
 * traced go to *closest_rays and *shadow_rays when not NULL. options as
 * dmap_intersect_closest's, except that reorder is ignored: the wavefront's
 * queues stay in pixel order, which is coherent already. Returns when the
 * image is complete on `stream`'s timeline (it synchronizes with the stream
 * once per bounce to size the next batch). */
typedef struct dmap_pt_scene
{
    double light_corner[3], light_u[3], light_v[3], light_n[3];
    double radiance, env, albedo;
} dmap_pt_scene;
int dmap_pathtrace(const dmap_object *obj, const double *cam_o, const double *cam_d, int64_t n_pix, int spp,
                   int depth, uint32_t seed, int strategy, const dmap_pt_scene *scene, const dmap_options *options,
                   double *image, int64_t *closest_rays, int64_t *shadow_rays, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DMAP_AMD_H */
