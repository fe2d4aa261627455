// Ray orders before a trace (neither changes a result, only the order of
// work). The sort: rays are sorted by a 24-bit key, the
// direction octant over a 21-bit Morton code of the origin quantized to
// 128 steps per axis of the object's bounding box, so that a wave's 64
// lanes hold rays that start close together and head the same way and walk
// the same cells. (9 bits per axis sorted no better and cost one more
// radix pass.) The sort
// is stable: rays with one origin (a pinhole camera's) keep their given
// order within an octant, so pixel-tiled primaries stay tiled. The trace
// reads rays through the permutation and writes each result at the ray's
// own index; only the order of work changes, never a result.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "kernels.h"

namespace dmap_amd
{

__device__ inline uint32_t part1by2(uint32_t x)
{
    x &= 0x000003ffu;
    x = (x | (x << 16)) & 0xff0000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

__device__ inline uint32_t quantize(float v, float lo, float scale)
{
    float q = (v - lo) * scale;
    q = fminf(fmaxf(q, 0.f), 127.f);
    return (uint32_t)q;
}

__global__ void k_ray_keys(RaysDev rays, int64_t n, float lx, float ly, float lz, float sx, float sy, float sz,
                           uint32_t *keys, uint32_t *index)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n)
        return;
    float dx = rays.dx[r], dy = rays.dy[r], dz = rays.dz[r];
    uint32_t oct = (dx < 0.f ? 1u : 0u) | (dy < 0.f ? 2u : 0u) | (dz < 0.f ? 4u : 0u);
    uint32_t m = part1by2(quantize(rays.ox[r], lx, sx)) | (part1by2(quantize(rays.oy[r], ly, sy)) << 1) |
                 (part1by2(quantize(rays.oz[r], lz, sz)) << 2);
    keys[r] = (oct << 21) | m;
    index[r] = (uint32_t)r;
}

__global__ void k_identity(uint32_t *perm, int64_t n)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n)
        perm[r] = (uint32_t)r;
}

hipError_t identity_order(uint32_t *perm, int64_t n, hipStream_t stream)
{
    hipLaunchKernelGGL(k_identity, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, perm, n);
    return hipGetLastError();
}

size_t reorder_workspace_bytes(int64_t n)
{
    size_t temp = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                       (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, 24);
    return 4 * (size_t)n * 4 + temp + 256;
}

hipError_t reorder_rays(const RaysDev &rays, int64_t n, const float lo[3], const float hi[3], void *workspace,
                        size_t workspace_bytes, uint32_t **perm, hipStream_t stream)
{
    uint32_t *keys_in = (uint32_t *)workspace;
    uint32_t *keys_out = keys_in + n;
    uint32_t *idx_in = keys_out + n;
    uint32_t *idx_out = idx_in + n;
    void *temp = idx_out + n;
    size_t temp_bytes = workspace_bytes - 4 * (size_t)n * 4;
    float s[3];
    for (int a = 0; a < 3; ++a)
        s[a] = hi[a] > lo[a] ? 128.f / (hi[a] - lo[a]) : 0.f;
    unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_ray_keys, dim3(blocks), dim3(256), 0, stream, rays, n, lo[0], lo[1], lo[2], s[0], s[1],
                       s[2], keys_in, idx_in);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, idx_in, idx_out, (int)n, 0, 24,
                                           stream);
    *perm = idx_out;
    return e;
}

// Near-first order: the rays whose origin lies inside the object's bounding
// box (rays that start at or near the displaced surface: bounces, shadow
// rays) go first, the longest first, and the others after in their given
// order. Rays that start near the surface walk long and cost the most; a
// persistent launch that takes them last ends on them, with its waves
// draining one expensive ray at a time. Taking them first leaves the cheap
// ones (rays from afar, many of which leave the object's boxes at once) for
// the drain: on 2M pinhole primaries followed by 2M cosine bounces from the
// surface, the trace took 2.55 ms near-first against 2.95 ms as given.
// Among the inside rays the cost follows the length of the ray's path
// through the box (a ray that leaves the surface at a low angle crosses many
// texels inside the displacement shell): they go in NEAR_BUCKETS classes of
// that length over the box's diagonal, half an octave each, longest first,
// each class in its given order (2.29 against 2.46 ms on the same rays,
// with the bounces in one class; a full sort by length was no faster, and
// ordering the outside rays by length gained nothing).
// One block of PART_BLOCK rays counts its rays per class and leaves each
// ray's class and rank in the block (2 bytes), a scan over the class-major
// counts places each class's share of each block, and a second pass
// scatters each ray's index into the permutation (stable in every class).
constexpr int PART_BLOCK = 256;
constexpr int NEAR_BUCKETS = 16;
constexpr int NEAR_CLASSES = NEAR_BUCKETS + 1; // the inside buckets, then the outside rays

struct NearBox
{
    float lo[3], hi[3];
    float inv_diag; // 1 / |hi - lo|, 0 for a degenerate box
};

// The class of ray r: 0 .. NEAR_BUCKETS - 1 for a ray whose origin is in
// the box, by -2 log2 of its path length in the box over the diagonal (the
// path from t_min to where it leaves the box, or t_max); NEAR_BUCKETS for
// the others (and for any ray past n). A ray whose length is not a positive
// finite number goes in the last inside bucket.
__device__ inline int near_class(const RaysDev &rays, int64_t r, int64_t n, const NearBox &B)
{
    if (r >= n)
        return NEAR_BUCKETS;
    const float o[3] = {rays.ox[r], rays.oy[r], rays.oz[r]};
    if (!(o[0] >= B.lo[0] && o[0] <= B.hi[0] && o[1] >= B.lo[1] && o[1] <= B.hi[1] && o[2] >= B.lo[2] &&
          o[2] <= B.hi[2]))
        return NEAR_BUCKETS;
    // (The class only orders work: the hardware reciprocal, square root and
    // log2, each within a few ulps, are exact enough.)
    const float d[3] = {rays.dx[r], rays.dy[r], rays.dz[r]};
    float t_exit = rays.tmax[r];
    for (int a = 0; a < 3; ++a)
        if (d[a] != 0.f)
            t_exit = fminf(t_exit, ((d[a] > 0.f ? B.hi[a] : B.lo[a]) - o[a]) * __builtin_amdgcn_rcpf(d[a]));
    const float len = (t_exit - rays.tmin[r]) * __builtin_amdgcn_sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const float x = len * B.inv_diag;
    if (!(x > 0.f) || !(x < INFINITY))
        return NEAR_BUCKETS - 1;
    const float b = floorf(-2.f * __builtin_amdgcn_logf(x));
    return b <= 0.f ? 0 : b >= (float)(NEAR_BUCKETS - 1) ? NEAR_BUCKETS - 1 : (int)b;
}

// Each ray's class and its rank among the block's rays of that class
// (earlier waves, then earlier lanes), as code[r] = class << 8 | rank; and
// the block's rays per class, class-major: counts[c * blocks + block].
__global__ void __launch_bounds__(PART_BLOCK) k_near_count(RaysDev rays, int64_t n, NearBox B, int blocks,
                                                           uint16_t *code, uint32_t *counts)
{
    static_assert(PART_BLOCK <= 256 && NEAR_CLASSES <= 255, "class and rank share 16 bits");
    __shared__ uint32_t wave_count[NEAR_CLASSES][PART_BLOCK / 64];
    for (int q = threadIdx.x; q < NEAR_CLASSES * (PART_BLOCK / 64); q += PART_BLOCK)
        (&wave_count[0][0])[q] = 0u;
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    const int c = r < n ? near_class(rays, r, n, B) : -1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // One ballot per class present in the wave (a wave of rays from afar
    // holds one), led by the first lane not yet counted.
    uint64_t mine = 0, todo = __ballot(r < n);
    while (todo) {
        const int lead = __ffsll((unsigned long long)todo) - 1;
        const int k = __shfl(c, lead);
        const uint64_t ball = __ballot(c == k);
        if (c == k)
            mine = ball;
        if (lane == lead)
            wave_count[k][wave] = (uint32_t)__popcll(ball);
        todo &= ~ball;
    }
    __syncthreads();
    if (threadIdx.x < NEAR_CLASSES) {
        uint32_t total = 0;
        for (int w = 0; w < PART_BLOCK / 64; ++w)
            total += wave_count[threadIdx.x][w];
        counts[(int64_t)threadIdx.x * blocks + blockIdx.x] = total;
    }
    if (r >= n)
        return;
    uint32_t rank = (uint32_t)__popcll(mine & ((1ull << lane) - 1));
    for (int w = 0; w < wave; ++w)
        rank += wave_count[c][w];
    code[r] = (uint16_t)((uint32_t)c << 8 | rank);
}

// Each ray's place: its class's offset for the block (the exclusive scan
// of the class-major counts) plus its rank in the block.
__global__ void __launch_bounds__(PART_BLOCK) k_near_scatter(int64_t n, int blocks, const uint16_t *code,
                                                             const uint32_t *offsets, uint32_t *perm)
{
    const int64_t r = (int64_t)blockIdx.x * PART_BLOCK + threadIdx.x;
    if (r >= n)
        return;
    const uint32_t k = code[r];
    perm[offsets[(int64_t)(k >> 8) * blocks + blockIdx.x] + (k & 255u)] = (uint32_t)r;
}

static size_t round256(size_t b) { return (b + 255) / 256 * 256; }

size_t near_first_workspace_bytes(int64_t n)
{
    const int64_t cells = (n + PART_BLOCK - 1) / PART_BLOCK * NEAR_CLASSES;
    size_t temp = 0;
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, temp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                           (int)cells);
    return round256((size_t)n * 4) + round256((size_t)n * 2) + 2 * round256((size_t)cells * 4) + temp + 256;
}

hipError_t near_first_rays(const RaysDev &rays, int64_t n, const float lo[3], const float hi[3], void *workspace,
                           size_t workspace_bytes, uint32_t **perm, hipStream_t stream)
{
    const int blocks = (int)((n + PART_BLOCK - 1) / PART_BLOCK);
    const int cells = blocks * NEAR_CLASSES;
    NearBox B;
    double diag2 = 0.0;
    for (int a = 0; a < 3; ++a) {
        B.lo[a] = lo[a];
        B.hi[a] = hi[a];
        diag2 += ((double)hi[a] - lo[a]) * ((double)hi[a] - lo[a]);
    }
    B.inv_diag = diag2 > 0.0 && std::isfinite(diag2) ? (float)(1.0 / std::sqrt(diag2)) : 0.f;
    char *at = (char *)workspace;
    uint32_t *out = (uint32_t *)at;
    at += round256((size_t)n * 4);
    uint16_t *code = (uint16_t *)at;
    at += round256((size_t)n * 2);
    uint32_t *counts = (uint32_t *)at;
    at += round256((size_t)cells * 4);
    uint32_t *offsets = (uint32_t *)at;
    at += round256((size_t)cells * 4);
    size_t temp_bytes = workspace_bytes - (size_t)(at - (char *)workspace);
    hipLaunchKernelGGL(k_near_count, dim3(blocks), dim3(PART_BLOCK), 0, stream, rays, n, B, blocks, code, counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    if ((e = hipcub::DeviceScan::ExclusiveSum(at, temp_bytes, counts, offsets, cells, stream)) != hipSuccess)
        return e;
    hipLaunchKernelGGL(k_near_scatter, dim3(blocks), dim3(PART_BLOCK), 0, stream, n, blocks, code, offsets, out);
    *perm = out;
    return hipGetLastError();
}

} // namespace dmap_amd
