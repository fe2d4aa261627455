// Ray reordering before a trace: rays are sorted by a 24-bit key, the
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

} // namespace dmap_amd
