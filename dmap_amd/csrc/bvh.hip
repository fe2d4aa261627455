// The top-level BVH built on the device, for large base meshes (no
// counterpart in the reference, which hands the triangles' boxes to embree:
// ObjectIntersector::build_bvh, displaced_intersector.cpp:540-554).
//
// The tree has the layout the traversal reads (kernels.h: BvhNode): a
// complete binary tree in heap order over P leaves (the power of two at or
// above the triangle count), node i's children 2i and 2i + 1, leaves at
// P .. 2P - 1, each internal node holding its two children's boxes as fp32
// centre and radius. It is the host builder's tree (api.cpp,
// build_bvh_node): each node splits its triangles at the median of their box
// centres along the longest axis of the centres' spread, the lower
// (m + 1) / 2 going to child 2i, so the node ranges in the triangle order
// depend on the count alone. The host finds each median with nth_element,
// node by node; here every level is one pass over all triangles:
//  * the triangles are sorted once per axis by lo + hi of their boxes (the
//    host's comparison key; a stable radix sort, so ties keep index order);
//  * a level's nodes each read their axis from the ends of their ranges in
//    the three orders, mark the first half of their range in their axis's
//    order as going left, and all three orders are stably partitioned by
//    that mark (one scan), so each stays sorted inside the children's
//    ranges.
// Then one launch per level, from the deepest internal level up, forms each
// node's two boxes from its children's subtree boxes. Boxes are the
// triangles' own fp32 boxes, unions are exact in fp32, and the centre and
// radius are rounded as the host does (bvh_centre_radius), so every node box
// holds its subtree's triangles' boxes. Only equal keys can go to different
// children than on the host (nth_element leaves their order open).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "kernels.h"

namespace dmap_amd
{

namespace
{

// An unsigned key that sorts as the float does (finite values).
__device__ inline uint32_t float_key(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ void k_bvh_axis_keys(const float *boxes, int n, int a, uint32_t *keys, uint32_t *idx)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    keys[k] = float_key(boxes[6 * (size_t)k + a] + boxes[6 * (size_t)k + 3 + a]);
    idx[k] = (uint32_t)k;
}

__device__ inline float centre(const float *boxes, uint32_t tri, int a)
{
    return 0.5f * (boxes[6 * (size_t)tri + a] + boxes[6 * (size_t)tri + 3 + a]);
}

// The node of one level (nodes first .. 2 first - 1) whose range holds
// position p: the last one starting at or before p (an empty node is
// followed by one with the same start, or ends the level).
__device__ inline int node_at(const int *start, int first, int p)
{
    int lo = 0, hi = first - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[first + mid] <= p)
            lo = mid;
        else
            hi = mid - 1;
    }
    return first + lo;
}

// Each node of the level: the axis of the longest spread of its triangles'
// box centres (the first of equal spreads; axis 0 below two triangles).
__global__ void k_bvh_split_axis(int first, const int *start, const int *cnt, const uint32_t *ord, int n,
                                 const float *boxes, int *axis_of)
{
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * first)
        return;
    const int lo = start[i], c = cnt[i];
    int axis = 0;
    if (c >= 2) {
        float best = 0.f;
        for (int a = 0; a < 3; ++a) {
            const uint32_t *o = ord + (size_t)a * n;
            const float spread = centre(boxes, o[lo + c - 1], a) - centre(boxes, o[lo], a);
            if (a == 0 || spread > best) {
                best = spread;
                axis = a;
            }
        }
    }
    axis_of[i] = axis;
}

// Marks the triangles that go to the left children: the first (c + 1) / 2
// of each node's range in its axis's order.
__global__ void k_bvh_mark(int first, const int *start, const int *cnt, const int *axis_of, const uint32_t *ord,
                           int n, uint8_t *left)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n)
        return;
    const int i = node_at(start, first, p);
    const int lo = start[i], c = cnt[i];
    const uint32_t tri = ord[(size_t)axis_of[i] * n + p];
    left[tri] = p < lo + (c + 1) / 2 ? 1 : 0;
}

__global__ void k_bvh_flags(const uint32_t *ord, int n, const uint8_t *left, uint32_t *flag)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * (int64_t)n)
        return;
    flag[q] = left[ord[q]];
}

// The stable partition of each axis's order inside each node's range: the
// marked triangles first, in order, then the rest. scan is the exclusive
// prefix sum of flag over all three orders.
__global__ void k_bvh_partition(int first, const int *start, const int *cnt, const uint32_t *ord,
                                const uint32_t *flag, const uint32_t *scan, int n, uint32_t *out)
{
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3 * (int64_t)n)
        return;
    const int a = (int)(q / n), p = (int)(q - (int64_t)a * n);
    const size_t base = (size_t)a * n;
    const int i = node_at(start, first, p);
    const int lo = start[i], c = cnt[i];
    const int mid = lo + (c + 1) / 2;
    const int lefts_before = (int)(scan[q] - scan[base + lo]);
    const int dst = flag[q] ? lo + lefts_before : mid + (p - lo - lefts_before);
    out[base + dst] = ord[q];
}

// Leaf j holds the one triangle of node P + j's range, or none; its subtree
// box is the triangle's box, or empty (lo > hi).
__global__ void k_bvh_leaves(const float *boxes, const uint32_t *ord, const int *start, const int *cnt, int P,
                             int *leaf, float *sub)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P)
        return;
    const int tri = cnt[P + j] > 0 ? (int)ord[start[P + j]] : -1;
    leaf[j] = tri;
    float *out = sub + 6 * ((size_t)P + j);
    for (int a = 0; a < 3; ++a) {
        out[a] = tri >= 0 ? boxes[6 * (size_t)tri + a] : INFINITY;
        out[3 + a] = tri >= 0 ? boxes[6 * (size_t)tri + 3 + a] : -INFINITY;
    }
}

// [lo, hi] as fp32 centre c and radius r with [c - r, c + r] holding it: the
// centre rounded to nearest, the radius formed in double, rounded up and
// raised by one more ulp (bvh_centre_radius on the host does the same).
__device__ inline void centre_radius(float lo, float hi, float &c, float &r)
{
    c = __double2float_rn(0.5 * ((double)lo + (double)hi));
    const double rd = fmax((double)hi - (double)c, (double)c - (double)lo);
    float rf = __double2float_rn(rd);
    if ((double)rf < rd)
        rf = nextafterf(rf, INFINITY);
    r = nextafterf(rf, INFINITY);
}

// Internal nodes first .. 2 first - 1 (one level): each node's two child
// boxes, which child is empty, its split axis, and its subtree box for the
// level above. Node 0 (first = 0) holds the whole tree's box as its child 0.
__global__ void k_bvh_boxes(int first, const int *axis_of, float *sub, BvhNode *nodes)
{
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (first == 0 ? i > 0 : i >= 2 * first)
        return;
    float lo[2][3], hi[2][3];
    bool empty[2];
    for (int c = 0; c < 2; ++c) {
        // node 0: child 0 is the tree (node 1), child 1 does not exist
        const int child = first == 0 ? 1 : 2 * i + c;
        const float *s = sub + 6 * (size_t)child;
        empty[c] = (first == 0 && c == 1) || !(s[0] <= s[3]);
        for (int a = 0; a < 3; ++a) {
            lo[c][a] = s[a];
            hi[c][a] = s[3 + a];
        }
    }
    BvhNode nd;
    nd.empty = (empty[0] ? 1 : 0) | (empty[1] ? 2 : 0);
    nd.axis = first == 0 ? 0 : axis_of[i];
    nd.pad[0] = nd.pad[1] = 0;
    for (int c = 0; c < 2; ++c)
        for (int a = 0; a < 3; ++a) {
            if (empty[c]) {
                nd.ctr[c][a] = 0.f;
                nd.rad[c][a] = 0.f;
            } else {
                centre_radius(lo[c][a], hi[c][a], nd.ctr[c][a], nd.rad[c][a]);
            }
        }
    nodes[i] = nd;
    if (first != 0) {
        float *out = sub + 6 * (size_t)i;
        for (int a = 0; a < 3; ++a) {
            out[a] = fminf(empty[0] ? INFINITY : lo[0][a], empty[1] ? INFINITY : lo[1][a]);
            out[3 + a] = fmaxf(empty[0] ? -INFINITY : hi[0][a], empty[1] ? -INFINITY : hi[1][a]);
        }
    }
}

// Carves the workspace: every array 256-byte aligned.
struct Carve
{
    char *at;
    template <class T> T *take(size_t count)
    {
        T *p = (T *)at;
        at += (count * sizeof(T) + 255) & ~(size_t)255;
        return p;
    }
};

size_t temp_bytes_for(int n)
{
    size_t sort = 0, scan = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                       (const uint32_t *)nullptr, (uint32_t *)nullptr, n, 0, 32);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, scan, (const uint32_t *)nullptr, (uint32_t *)nullptr, 3 * n);
    return std::max(sort, scan);
}

unsigned blocks(int64_t n) { return (unsigned)((n + 255) / 256); }

size_t rounded(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

} // namespace

size_t bvh_build_workspace_bytes(int n, int P)
{
    return 256 + 7 * rounded(3 * (size_t)n * 4) + rounded((size_t)n) + 3 * rounded(2 * (size_t)P * 4) +
           rounded(2 * (size_t)P * 6 * 4) + rounded(temp_bytes_for(n));
}

hipError_t bvh_build_device(const float *boxes, int n, int P, BvhNode *nodes, int *leaf, float tree_box[6],
                            void *workspace, size_t workspace_bytes, hipStream_t stream)
{
    // The node ranges in the triangle order: node 1 holds [0, n), and a
    // node of c triangles gives its first (c + 1) / 2 to child 2i.
    std::vector<int> start(2 * (size_t)P, 0), cnt(2 * (size_t)P, 0);
    cnt[1] = n;
    for (int i = 1; i < P; ++i) {
        const int half = (cnt[i] + 1) / 2;
        start[2 * i] = start[i];
        cnt[2 * i] = half;
        start[2 * i + 1] = start[i] + half;
        cnt[2 * i + 1] = cnt[i] - half;
    }
    Carve w{(char *)(((uintptr_t)workspace + 255) & ~(uintptr_t)255)};
    uint32_t *ord = w.take<uint32_t>(3 * (size_t)n);
    uint32_t *ord2 = w.take<uint32_t>(3 * (size_t)n);
    uint32_t *flag = w.take<uint32_t>(3 * (size_t)n);
    uint32_t *scan = w.take<uint32_t>(3 * (size_t)n);
    uint32_t *keys = w.take<uint32_t>(3 * (size_t)n);
    uint32_t *keys2 = w.take<uint32_t>(3 * (size_t)n);
    uint32_t *idx = w.take<uint32_t>(3 * (size_t)n);
    uint8_t *left = w.take<uint8_t>((size_t)n);
    int *d_start = w.take<int>(2 * (size_t)P);
    int *d_cnt = w.take<int>(2 * (size_t)P);
    int *axis_of = w.take<int>(2 * (size_t)P);
    float *sub = w.take<float>(2 * (size_t)P * 6);
    const size_t temp_bytes = temp_bytes_for(n);
    void *temp = w.take<char>(temp_bytes);
    if ((size_t)(w.at - (char *)workspace) > workspace_bytes)
        return hipErrorInvalidValue;
    hipError_t e;
    if ((e = hipMemcpyAsync(d_start, start.data(), sizeof(int) * 2 * (size_t)P, hipMemcpyHostToDevice, stream)) !=
            hipSuccess ||
        (e = hipMemcpyAsync(d_cnt, cnt.data(), sizeof(int) * 2 * (size_t)P, hipMemcpyHostToDevice, stream)) !=
            hipSuccess)
        return e;
    // The three sorted orders.
    for (int a = 0; a < 3; ++a) {
        hipLaunchKernelGGL(k_bvh_axis_keys, dim3(blocks(n)), dim3(256), 0, stream, boxes, n, a, keys + (size_t)a * n,
                           idx + (size_t)a * n);
        size_t tb = temp_bytes;
        if ((e = hipcub::DeviceRadixSort::SortPairs(temp, tb, keys + (size_t)a * n, keys2 + (size_t)a * n,
                                                    idx + (size_t)a * n, ord + (size_t)a * n, n, 0, 32, stream)) !=
            hipSuccess)
            return e;
    }
    // The levels, top down: level nodes first .. 2 first - 1.
    for (int first = 1; first < P; first *= 2) {
        hipLaunchKernelGGL(k_bvh_split_axis, dim3(blocks(first)), dim3(256), 0, stream, first, d_start, d_cnt, ord, n,
                           boxes, axis_of);
        hipLaunchKernelGGL(k_bvh_mark, dim3(blocks(n)), dim3(256), 0, stream, first, d_start, d_cnt, axis_of, ord, n,
                           left);
        hipLaunchKernelGGL(k_bvh_flags, dim3(blocks(3 * (int64_t)n)), dim3(256), 0, stream, ord, n, left, flag);
        size_t tb = temp_bytes;
        if ((e = hipcub::DeviceScan::ExclusiveSum(temp, tb, flag, scan, 3 * n, stream)) != hipSuccess)
            return e;
        hipLaunchKernelGGL(k_bvh_partition, dim3(blocks(3 * (int64_t)n)), dim3(256), 0, stream, first, d_start, d_cnt,
                           ord, flag, scan, n, ord2);
        std::swap(ord, ord2);
    }
    // Leaves and boxes, bottom up.
    hipLaunchKernelGGL(k_bvh_leaves, dim3(blocks(P)), dim3(256), 0, stream, boxes, ord, d_start, d_cnt, P, leaf, sub);
    for (int first = P / 2; first >= 1; first /= 2)
        hipLaunchKernelGGL(k_bvh_boxes, dim3(blocks(first)), dim3(256), 0, stream, first, axis_of, sub, nodes);
    hipLaunchKernelGGL(k_bvh_boxes, dim3(1), dim3(64), 0, stream, 0, axis_of, sub, nodes);
    if ((e = hipGetLastError()) != hipSuccess)
        return e;
    // The tree's box: node 1's subtree (a leaf's when P is 1). The host
    // vectors above must outlive the copies: wait for the stream.
    if ((e = hipMemcpyAsync(tree_box, sub + 6, 6 * sizeof(float), hipMemcpyDeviceToHost, stream)) != hipSuccess)
        return e;
    return hipStreamSynchronize(stream);
}

} // namespace dmap_amd
