// Shared host/device layout of the resident tile and of the per-triangle
// records the traversal reads. Plain structs, no constructors: they are
// copied to the device as bytes.
#pragma once
#include <stddef.h>
#include <stdint.h>

namespace dmap_amd
{

// A resident displacement tile: the closed node grid ((n + 1)^2 fp32
// texel values, close_tile's output, texture_grid.cpp:118-132) and the
// Taylor pyramid (taylor_pyramid.h:51-100) in three channel groups, each
// a fp32 array over every level:
//   mm = (h_min, h_max)        float2 per cell, read by every box test;
//   tp = (h0, gu, gv, r)       float4 per cell, read by the slab test;
//   rr = (ru, rv)              float2 per cell, read by the fold and export.
// Levels are stored root first: level k (m = n >> k cells per side) starts
// at cell (m^2 - 1) / 3. Inside a level cells are in Morton order (i in the
// even bits, j in the odd bits), so the four children of a cell are four
// consecutive entries: one 32 B sector of mm, one 64 B sector of tp.
// Channels are the double-precision build rounded outward: h_min down,
// h_max up; h0, gu, gv to nearest with the rounding folded into r, ru, rv
// (rounded up), so every fp32 node still bounds the bilinear field.
struct PyramidView
{
    const float2 *mm;
    const float4 *tp;
    const float2 *rr;
    const float *nodes; // (n + 1)^2 closed grid, raw texel values (scale not applied)
    int n_leaf;         // n = 2^L texels per side
    int n_levels;       // L + 1
    int repeat;         // wrap (1) or clamp (0), HeightGrid::repeat
    float scale, offset;
    double scale_d, offset_d;
    float root_hmin, root_hmax; // root min-max node, served above the root (taylor_pyramid.cpp:261-266)
};

// Morton interleave of two coordinates below 2^16.
__host__ __device__ inline uint32_t part1by1(uint32_t x)
{
    x &= 0x0000ffffu;
    x = (x | (x << 8)) & 0x00ff00ffu;
    x = (x | (x << 4)) & 0x0f0f0f0fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}
__host__ __device__ inline uint32_t morton2(uint32_t i, uint32_t j) { return part1by1(i) | (part1by1(j) << 1); }
// The even bits of x, compacted (inverse of part1by1).
__host__ __device__ inline uint32_t compact1by1(uint32_t x)
{
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0f0f0f0fu;
    x = (x | (x >> 4)) & 0x00ff00ffu;
    x = (x | (x >> 8)) & 0x0000ffffu;
    return x;
}
// First entry of the level of m x m cells. Levels follow each other root
// first, after LEVEL_PAD unused entries: (m^2 - 1) / 3 = 1 + 4 + ... + m^2 / 4
// is 1 mod 4 below the root, so the pad puts every level, and with it every
// Morton group of four children, on a multiple of four entries; a group is
// then one aligned 32-byte min-max sector and one aligned 64-byte Taylor
// line, which the walk reads as whole vectors.
constexpr uint32_t LEVEL_PAD = 3;
__host__ __device__ inline uint32_t level_base(uint32_t m) { return (m * m - 1u) / 3u + LEVEL_PAD; }
// Entries of a pyramid over n x n leaves, the pad included.
__host__ __device__ inline size_t pyramid_cells(uint32_t n) { return ((size_t)4 * n * n - 1) / 3 + LEVEL_PAD; }

// What the traversal needs of one base triangle. Built on the host in
// double from the reference's BaseTriangle (displaced_surface.cpp:54-91),
// TangentFrame (node_bounds.cpp:13-23), UvTriangle (uv_clip.cpp:10-25),
// footprint_roots (descent_sampler.cpp:66-83) and the DisplacedIntersector
// constructor (displaced_intersector.cpp:135-151).
// The walk keeps four trail bits per level, from the root block's level down
// to the leaves, in one 64-bit word: a triangle's root level is at most 15.
constexpr int MAX_ROOT_LEVEL = 15;

struct TriRecord
{
    // Surface in the chart: P = q0 + s e1 + t e2, M = m0 + s Mu + t Mv.
    double q0[3], e1[3], e2[3], m0[3], Mu[3], Mv[3];
    // to_tangent = [e1 e2 ng]^-1, row-major.
    double T[9];
    // Chart: (u, v) = J ((s, t) - t0).
    double J[4], t0[2];
    // Texture-space domain: edge normals n[k], offsets d[k] (inside iff
    // n . p >= d), bounding box.
    float dn[6], dd[3];
    float dlo[2], dhi[2];
    double dn_d[6], dd_d[3];
    double dlo_d[2], dhi_d[2];
    // Walk roots: root_level and up to four (i, j) cells.
    int root_level, n_roots;
    int root_i[4], root_j[4];
    // Constant interpolated normal (Mu = Mv = 0): the normalized normal's
    // affine form in tangent space (centre and remainder) is the same for
    // every cell; so is the object-space unit normal.
    int const_normal;
    float mc[3], mk[3];     // to_tangent N_hat, centre and remainder per axis
    float nhat[3], nhat_k;  // object-space N_hat (centre) and its remainder bound
    double mc_d[3], mk_d[3];
    double nhat_d[3], nhat_kd;
    float f_m0[3], f_Mu[3], f_Mv[3]; // normal fields in fp32 (general path)
    float f_T[9];
    double residual_tolerance; // 1e-10 x mean edge (displaced_intersector.cpp:150)
    float aabb_lo[3], aabb_hi[3];
    float mck[3]; // |mc| + mk per axis, rounded up
    // The height range over the footprint (triangle_height_range,
    // node_bounds.cpp:238-249), outward-rounded: every leaf texel the walk
    // can reach has its heights inside it.
    float h_lo, h_hi;
    float pad_[3];
};

} // namespace dmap_amd
