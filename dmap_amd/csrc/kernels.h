// Launch wrappers of the HIP kernels, called by the C-ABI layer (api.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "layout.h"

namespace dmap_amd
{

// Writable view of a pyramid during its build.
struct PyramidOut
{
    float2 *mm;
    float4 *tp;
    float2 *rr;
    int n_leaf, n_levels;
    double scale, offset;
};

hipError_t launch_close_tile(const float *texels, float *nodes, int n, int repeat, hipStream_t stream);
size_t fold_scratch_doubles(int n);
hipError_t launch_build_fold(const PyramidOut &P, const float *nodes, double *scratch, hipStream_t stream);
hipError_t launch_build_direct(const PyramidOut &P, const float *nodes, hipStream_t stream);

// triangle_height_range (node_bounds.cpp:218-249) for every triangle: one
// workgroup per triangle walks the min-max channel over its footprint, the
// walk split into cells a few levels below the root cells.
hipError_t launch_height_range(const PyramidView &P, const TriRecord *tris, int n_tri, float *h_range,
                               hipStream_t stream);

// Ray and hit buffers (device pointers, SoA).
struct RaysDev
{
    const float *ox, *oy, *oz, *dx, *dy, *dz, *tmin, *tmax;
};
struct HitsDev
{
    float *t, *s, *tt, *bu, *bv, *nx, *ny, *nz;
    int32_t *prim, *texel_i, *texel_j;
};
// dmap_node_record's layout (include/dmap_amd.h), device side.
struct NodeRecordDev
{
    int32_t ray, prim, level, i, j, seq;
    int32_t box_pass, slab_pass;
    float box_length, slab_length, both_length, pad_;
};
struct StatsDev
{
    int32_t *triangle_tests, *node_tests, *leaf_tests, *sub_squares, *newton_runs, *newton_iters;
    int32_t *node_tests_after_hit, *leaf_tests_after_hit;
    int32_t *node_tests_per_level; // STAT_LEVELS planes of n_rays, level-major
    NodeRecordDev *records;        // NULL: no node-test records
    int64_t record_capacity;
    unsigned long long *record_count;
};
constexpr int STAT_LEVELS = 16; // levels 0 .. MAX_ROOT_LEVEL

// The object's top-level BVH over the triangles' boxes: a complete binary
// tree in heap order with P = 2^k leaves. Node 1 is the root, node i has
// children 2i and 2i + 1, and each of the leaves i in [P, 2P) holds one
// triangle or none (bvh_leaf[i - P] = triangle or -1). An internal node
// holds its two children's boxes, so one 64-byte read serves both tests;
// node 0 holds the root's own box in ctr[0], rad[0]. The traversal needs no
// parent links or stack: the path to a node is its index. A box is stored
// as centre +- radius per axis, rounded so that it holds the box it stands
// for, which is the form the ray test uses.
struct BvhNode
{
    float ctr[2][3], rad[2][3]; // the boxes of children 2i (lower centres along axis) and 2i + 1
    int axis;                   // the split axis
    int empty;                  // bit c: child 2i + c holds no triangle
    int pad[2];
};
static_assert(sizeof(BvhNode) == 64, "one BVH node is 64 bytes");

// What the traversal reads of one object: its tile and its triangles.
struct ObjectDev
{
    PyramidView pyr;
    const TriRecord *tris;
    const float *boxes; // n_tri x 6: lo xyz, hi xyz
    const BvhNode *bvh;  // P nodes: 0 (the root's box), 1 .. P - 1 (internal)
    const int *bvh_leaf; // P leaves
    int bvh_leaves;      // P
    int n_tri;
    int const_normal; // every triangle has a constant interpolated normal
};

// One request's share of a serving step: the step slots from `start` up to
// the next segment's start hold the request's rays `first`, first + 1, ...
// The request's ray planes are at `rays` + q * rays_stride (q = ox .. tmax)
// when they lie at one stride, which is the usual case; with rays_stride 0,
// `rays` points at a RaysDev record in device memory. Its hits are its
// region of the server's pool: 8 fp32 planes of hits_n values (t, s, tt,
// bu, bv, nx, ny, nz), then 3 int32 planes (prim, texel_i, texel_j).
struct Segment
{
    int64_t start, first;
    const float *rays;
    int64_t rays_stride;
    float *hits;
    int64_t hits_n;
};
static_assert(sizeof(Segment) == 48, "a step segment is 48 bytes");
__host__ __device__ inline RaysDev segment_rays(const Segment &g)
{
    if (g.rays_stride == 0)
        return *(const RaysDev *)g.rays;
    const float *r = g.rays;
    const int64_t k = g.rays_stride;
    return RaysDev{r, r + k, r + 2 * k, r + 3 * k, r + 4 * k, r + 5 * k, r + 6 * k, r + 7 * k};
}
__host__ __device__ inline HitsDev segment_hits(const Segment &g)
{
    float *f = g.hits;
    const int64_t n = g.hits_n;
    int32_t *i = (int32_t *)(f + 8 * n);
    return HitsDev{f, f + n, f + 2 * n, f + 3 * n, f + 4 * n, f + 5 * n, f + 6 * n, f + 7 * n, i, i + n, i + 2 * n};
}
// A step's descriptor travels in the kernels' arguments (no upload before
// the launch); with at most MAX_SEGMENTS requests and MAX_GROUPS objects per
// step it stays, with TraceParams, under the 4 KB of kernel arguments.
constexpr int MAX_SEGMENTS = 32;
constexpr int MAX_GROUPS = 8;
constexpr int64_t MAX_STEP_RAYS = (int64_t)1 << 30;
// One object's share of a serving step: the requests seg[seg_lo, seg_hi),
// whose rays fill the step slots [start, end), traced by the trace kernel's
// workgroups [wg_lo, wg_hi) and finalized by the finalize kernel's
// workgroups [fwg_lo, fwg_hi). A workgroup serves one object only, so the
// object's fields are uniform across its waves and its BVH fits in LDS.
struct StepGroup
{
    ObjectDev obj;
    int64_t start, end;
    int wg_lo, wg_hi, fwg_lo, fwg_hi;
    int seg_lo, seg_hi;
};
struct StepDesc
{
    int64_t total;
    int n_seg, n_grp;
    Segment seg[MAX_SEGMENTS];
    StepGroup grp[MAX_GROUPS];
};

// A closest-hit trace's result for one ray, one aligned 32-byte sector:
// k_trace writes it whole at the ray's index (the lanes of a wave hold rays
// far apart in index, so the six 4-byte fields stored into separate arrays
// would each write a fraction of a sector) and k_finalize reads it back in
// index order. a, b: the root in texel-local coordinates.
struct alignas(32) RawHit
{
    float t;
    int32_t prim, ti, tj;
    float a, b, pad_[2];
};

struct TraceParams
{
    ObjectDev obj;        // single-object launches
    int prim_mask; // -1: every triangle; k: triangle k only
    int bound;     // 0 box, 1 slab, 2 both
    int slab_max_level;
    int leaf; // 0 newton, 1 two-triangle
    int polish; // 1: the reported Newton root is polished in double (two steps)
    int walk_start; // 0: constant-normal walks start at the shell's block (start_block); 1: at the roots
    int any_hit;
    int64_t n_rays;
    const uint32_t *perm; // lane order of the rays (NULL: as given)
    unsigned *queue;      // the launch's ray queue head (reset by launch_trace); a serving step's
                          // group g takes queue[g]
    RaysDev rays;
    HitsDev hits;
    RawHit *raw; // n_rays records for k_finalize (NULL: the raw record goes into the hit arrays)
    uint8_t *occluded;
    StatsDev stats;
};

static_assert(sizeof(TraceParams) + sizeof(StepDesc) <= 4096, "a step's kernel arguments exceed 4 KB");

// The top-level BVH built on the device (bvh.hip) from n triangle boxes
// already in device memory (n x 6 fp32: lo xyz, hi xyz) into P = 2^k >= n
// leaves: nodes (P BvhNodes) and leaf (P triangle indices, -1 for none).
// tree_box (host memory) receives the whole tree's box. Waits for the
// stream.
size_t bvh_build_workspace_bytes(int n, int P);
hipError_t bvh_build_device(const float *boxes, int n, int P, BvhNode *nodes, int *leaf, float tree_box[6],
                            void *workspace, size_t workspace_bytes, hipStream_t stream);

// Ray orders (reorder.hip), each a permutation of 0..n-1 written into the
// workspace and returned in *perm: near_first_rays puts the rays whose
// origin lies in [lo, hi] first, longest path through the box first (stable
// within each class of path length); reorder_rays sorts by direction octant
// and origin for coherence.
hipError_t identity_order(uint32_t *perm, int64_t n, hipStream_t stream);
size_t near_first_workspace_bytes(int64_t n);
hipError_t near_first_rays(const RaysDev &rays, int64_t n, const float lo[3], const float hi[3], void *workspace,
                           size_t workspace_bytes, uint32_t **perm, hipStream_t stream);
size_t reorder_workspace_bytes(int64_t n);
hipError_t reorder_rays(const RaysDev &rays, int64_t n, const float lo[3], const float hi[3], void *workspace,
                        size_t workspace_bytes, uint32_t **perm, hipStream_t stream);

// Adjoint of the closest hit: from the hits of a closest-hit launch and the
// upstream gradients dL/dt and dL/dng, adds dL/d(texel) into grad (n_leaf^2
// doubles, row-major [j][i]) by the implicit-function theorem at each hit's
// double root.
struct AdjointParams
{
    ObjectDev obj;
    RaysDev rays;
    HitsDev hits;
    const float *dl_dt;    // NULL: no depth term
    const float *dl_dn[3]; // NULL: no normal term
    double *grad;
    int64_t n_rays;
};
hipError_t launch_adjoint(const AdjointParams &p, hipStream_t stream);

// The renderer's hit record (DisplacedGeometry::surface_point and
// compute_intersection's frame, displaced_geometry.cpp:126-183) for the hits
// of a closest-hit launch: 20 fp32 planes of n values each, p xyz, ng xyz,
// dpdu xyz, dpdv xyz, uv, frame t xyz, frame b xyz (misses: zeros).
struct UvMap
{
    double scale_u, scale_v, offset_u, offset_v;
    int flip_v;
};
// The world-space part of the record: the object-to-world transform
// (world = m x + t, row-major 3 x 3; nt = m^-T for normals) and the rays'
// world-space differentials (origins and directions of the offset rays rx
// and ry, 12 fp32 planes: rx o xyz, rx d xyz, ry o xyz, ry d xyz).
struct Instance
{
    double m[9], t[3], nt[9];
    int on;
};
struct RayDiffs
{
    const float *p[12];
    int on;
};
hipError_t launch_hit_record(const ObjectDev &obj, const RaysDev &rays, const HitsDev &hits, int64_t n,
                             const UvMap &uv, int twosided, const Instance &inst, const RayDiffs &diffs,
                             float *record, hipStream_t stream);

// Closest hit / any hit over the object; precision 0 = fp32, 1 = fp64.
// A closest-hit launch is followed by the finalize kernel, which turns the
// raw record (t, prim, texel, texel-local root in s/tt) into the hit.
// ev_start / ev_end, when given, are recorded on `stream` right before and
// after the traversal kernel (for its own duration).
hipError_t launch_trace(const TraceParams &p, int precision, hipStream_t stream, hipEvent_t ev_start = nullptr,
                        hipEvent_t ev_end = nullptr);
// The multi-request closest hit of serving as a HIP graph: reset of the
// step's ray queue, trace and finalize (fp32, Newton leaf, no counters),
// with grids that do not depend on the step. Each launch sets the trace and
// finalize nodes' arguments to the step's descriptor and replays the graph.
// Every object of a step has constant normals (const_normal) or none has.
struct StepGraph
{
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t trace = nullptr, finalize = nullptr;
    bool const_normal = false;
};
// The trace and finalize workgroups of a step, shared out among its groups
// in proportion to their rays (at least one each): fills the groups'
// wg_lo/wg_hi and fwg_lo/fwg_hi from their start/end.
void step_partition(StepDesc &d);
hipError_t step_graph_create(unsigned *queue, bool const_normal, StepGraph &g);
hipError_t step_graph_launch(StepGraph &g, const TraceParams &p, const StepDesc &desc, hipStream_t stream);
void step_graph_destroy(StepGraph &g);

} // namespace dmap_amd
