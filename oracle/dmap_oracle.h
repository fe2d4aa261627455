/*
 * CPU oracle for the ray / displaced-surface intersection path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a double-precision restatement of the
 * reference's CPU path (code/cpp/src/{displaced_surface,taylor_pyramid,
 * affine_form,uv_clip,node_bounds,displaced_intersector}.cpp and
 * descent_sampler.cpp:15-27,66-83), used by tests/ as the parity checker,
 * by __graft_entry__.smoke() as its checker and by bench.py's cpu_baseline
 * leg as the timed host-core baseline. The product library never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it follows. Eigen is
 * replaced by explicit 2- and 3-vectors and a 3x3 adjugate inverse; embree's
 * top level by the reference's own flat AABB loop
 * (displaced_intersector.cpp:578-599).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Channel order of a pyramid node, as taylor_pyramid.h:63-66. */
enum { ORC_H0 = 0, ORC_GU, ORC_GV, ORC_R, ORC_RU, ORC_RV, ORC_HMIN, ORC_HMAX, ORC_CHANNELS };

/* Number of levels of the pyramid over a (n+1)^2 node grid (taylor_pyramid.cpp:224-229). */
int orc_pyramid_levels(int nodes);
/* Total cells over all levels: sum over k of (n >> k)^2. */
int64_t orc_pyramid_cells(int nodes);

/* TaylorPyramid(values, nodes, scale, build, offset) (taylor_pyramid.cpp:215-251).
 * build: 0 fold, 1 direct. out receives, level by level from the leaves up,
 * the 8 channels of that level, each an m*m row-major [j*m + i] array.
 * Returns the level count, or -1 for a node grid that is not 2^L + 1. */
int orc_pyramid_build(const double *values, int nodes, double scale, double offset, int build, double *out);

/* close_tile (texture_grid.cpp:118-132): n*n texels -> (n+1)^2 nodes. */
void orc_close_tile(const double *texels, int n, int repeat, double *nodes_out);

/* A scene: one node grid and its pyramid shared by n_tri base triangles
 * (DisplacementAsset + ObjectIntersector, displacement_asset.cpp:16-35,
 * displaced_intersector.cpp:466-479). tri_data holds per triangle 24
 * doubles: p0 p1 p2 (9), m0 m1 m2 vertex normals (9), t0 t1 t2 texture
 * coordinates (6). Returns NULL on a degenerate input. */
void *orc_scene_create(const double *node_values, int nodes, double scale, double offset, int repeat, int build,
                       int n_tri, const double *tri_data);
void orc_scene_destroy(void *scene);

/* Per-triangle set-up: root level, root count, the 4 roots (i, j pairs),
 * the object-space AABB (lo xyz, hi xyz), the height range, the residual
 * tolerance. */
int orc_triangle_info(void *scene, int k, int *root_level, int *n_roots, int64_t *roots_ij, double *aabb,
                      double *h_range, double *residual_tolerance);

/* Closest hit over the object, ObjectIntersector::intersect_flat
 * (displaced_intersector.cpp:578-599), one ray per entry, on nthreads host
 * threads (0 = all). o, d: n*3; tmin, tmax: n. Outputs (any may be NULL
 * except t): t (inf on a miss), st n*2, bary n*2, ng n*3, prim n (-1 on a
 * miss), texel n*2, stats n*ORC_STATS. bound: 0 box, 1 slab, 2 both;
 * leaf: 0 newton, 1 two-triangle. */
/* Per-ray stats (TraversalStats, displaced_intersector.h:89-105): the six
 * counts, the node and leaf tests made after the walk's final hit in walks
 * that hit, then node tests per level (levels >= ORC_MAX_LEVELS - 1 in the
 * last entry). */
enum { ORC_ST_TRIANGLE = 0, ORC_ST_NODE, ORC_ST_LEAF, ORC_ST_SUB, ORC_ST_NEWTON_RUNS, ORC_ST_NEWTON_ITERS,
       ORC_ST_NODE_AFTER_HIT, ORC_ST_LEAF_AFTER_HIT, ORC_ST_LEVEL0 };
enum { ORC_MAX_LEVELS = 32, ORC_STATS = ORC_ST_LEVEL0 + ORC_MAX_LEVELS };
void orc_scene_intersect(void *scene, int64_t n, const double *o, const double *d, const double *tmin,
                         const double *tmax, int bound, int slab_max_level, int leaf, int nthreads, double *t,
                         double *st, double *bary, double *ng, int32_t *prim, int64_t *texel, int64_t *stats);

/* One node test as the walk made it, NodeTestRecord (displaced_intersector.h:
 * 77-87; record_node_test, displaced_intersector.cpp:159-172): the ray's
 * index, the triangle, the cell, the test's place among the ray's node tests
 * (seq, from 0), the box's and the slab's decisions with the walk's t_max,
 * and the three interval lengths with t_max = inf (-1 when that test fails). */
typedef struct
{
    int32_t ray, prim, level, seq;
    int64_t i, j;
    int32_t box_pass, slab_pass;
    double box_length, slab_length, both_length;
} orc_node_record;

/* The records of every node test of the closest-hit walks of n rays, in
 * order (ray by ray, one thread), as TraversalStats::records collects them
 * (test_traversal_cost.cpp:377-386). Writes at most capacity records into
 * out and the total into *count. */
void orc_scene_records(void *scene, int64_t n, const double *o, const double *d, const double *tmin,
                       const double *tmax, int bound, int slab_max_level, int leaf, int64_t capacity,
                       orc_node_record *out, int64_t *count);

/* Any-hit: the occluded callback's answer (displaced_geometry.cpp:79-106)
 * with no filter, over the flat loop with an early exit. */
void orc_scene_occluded(void *scene, int64_t n, const double *o, const double *d, const double *tmin,
                        const double *tmax, int bound, int slab_max_level, int leaf, int nthreads, uint8_t *occluded);

/* One triangle's walk, DisplacedIntersector::intersect (displaced_intersector.cpp:175-237). */
int orc_triangle_intersect(void *scene, int k, const double *o, const double *d, double tmin, double tmax, int bound,
                           int slab_max_level, int leaf, double *hit10, int64_t *stats);

/* The Newton leaf on texel (i, j) and one Newton run from (s, t)
 * (displaced_intersector.cpp:239-264, 325-388). hit10 = t, st(2), bary(2), ng(3), i, j. */
int orc_leaf_newton(void *scene, int k, const double *o, const double *d, double tmin, double tmax, int64_t i,
                    int64_t j, double *hit10, int64_t *sub_squares);
int orc_newton_from(void *scene, int k, const double *o, const double *d, double tmin, double tmax, int64_t i,
                    int64_t j, double s, double t, double *hit10);
int orc_leaf_two_triangle(void *scene, int k, const double *o, const double *d, double tmin, double tmax, int64_t i,
                          int64_t j, double *hit10);

/* node_bound (node_bounds.cpp:83-109): box lo/hi (6), m intervals (6:
 * lo x, hi x, lo y, hi y, lo z, hi z), centre (2), rect lo/hi (4), the
 * node's 8 channels. */
void orc_node_bound(void *scene, int k, int level, int64_t i, int64_t j, double *box6, double *m6, double *centre2,
                    double *rect4, double *node8);
/* node_interval and the three tests (node_bounds.cpp:130-211) on a ray in
 * object space. mode: 0 box, 1 slab, 2 both. */
int orc_node_interval(void *scene, int k, int level, int64_t i, int64_t j, const double *o, const double *d,
                      double tmin, double tmax, int mode, double *t01);

/* Pointwise surface (displaced_surface.cpp:162-166, 99-108, 29-47): the
 * displaced position S(s, t), and the unit normal of Su x Sv, for
 * triangle k. */
void orc_displaced_position(void *scene, int k, double s, double t, double *S3);
void orc_surface_normal(void *scene, int k, double s, double t, double *n3);
/* Batched position and normal for n (triangle, s, t) samples. */
void orc_surface_points(void *scene, int64_t n, const int32_t *tri, const double *st, double *S3, double *n3);
/* The height field h and its gradient at texture coordinates (x, y). */
void orc_field_eval(void *scene, int64_t n, const double *xy, double *h, double *grad2);

/* AffineVec3::normalize (affine_form.cpp:87-95; Min-Range rsqrt :64-76,
 * sq_length :78-85) of one affine vector: in12/out12 hold x, y, z as
 * (c, a, b, k), the form c + a eu + b ev +- k. */
void orc_affine_normalize(const double *in12, double *out12);
/* tangent_frame (node_bounds.cpp:13-23) of triangle k: the origin q0 and the
 * row-major 3x3 to_tangent = [e1 e2 ng]^-1. */
void orc_tangent_frame(void *scene, int k, double *origin3, double *to_tangent9);

/* classify_square (uv_clip.cpp:27-44): 0 outside, 1 inside, 2 straddle. */
int orc_classify_square(const double *t6, double cx, double cy, double s);

/* Renderer hit record, DisplacedGeometry::surface_point
 * (displaced_geometry.cpp:126-152): from (prim, u, v) base barycentrics,
 * position (3), oriented unit normal (3), dpdu (3), dpdv (3) with unit uv
 * scale. */
void orc_surface_point(void *scene, int k, double u, double v, double *out12);
/* The renderer's hit record at barycentrics (u, v) of triangle k
 * (DisplacedGeometry::surface_point + compute_intersection's frame, object
 * space): p, ng, dpdu, dpdv, uv, frame t, frame b (20 doubles). uv_map =
 * (scale_u, scale_v, offset_u, offset_v, flip_v). */
void orc_hit_record(void *scene, int k, double u, double v, const double *dir3, const double *uv_map, int twosided,
                    double *out20);

/* Adjoint of the closest hit (no counterpart in the reference): adds into
 * grad_nodes (nodes x nodes, row-major [j][i]) the gradient of
 * sum_r dl_dt[r] t_r + dl_dn[r] . ng_r with respect to the closed grid's node
 * values, by the implicit-function theorem at each hit's double root. Also
 * writes that root's t and ng (misses: inf and 0). dl_dn (n x 3) may be NULL. */
void orc_scene_adjoint(void *scene, int64_t n, const double *o, const double *d, const double *tmin,
                       const double *tmax, const double *dl_dt, const double *dl_dn, int nthreads, double *grad_nodes,
                       double *t_out, double *ng_out);

#ifdef __cplusplus
}
#endif
