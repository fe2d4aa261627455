// Internal handles shared by the C ABI (api.cpp) and the serving runtime
// (serve.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/dmap_amd.h"
#include "kernels.h"
#include "layout.h"

using namespace dmap_amd;

struct dmap_ctx
{
    int device = 0;
    // The device address of the no-allocation entry point's scratch array
    // on this context's device (hipGetSymbolAddress at creation).
    void *scratch = nullptr;
    // The calls' stream-ordered workspaces come from this pool, which keeps
    // the memory it has handed out (release threshold at its maximum)
    // instead of returning it at every synchronization, until the context
    // is destroyed.
    hipMemPool_t pool = nullptr;
};

struct dmap_tile
{
    dmap_ctx *ctx = nullptr;
    int n = 0, n_levels = 0, repeat = 1, build = 0;
    double strength = 1.0, offset = 0.0;
    float *texels = nullptr, *nodes = nullptr;
    float2 *mm = nullptr;
    float4 *tp = nullptr;
    float2 *rr = nullptr;
    double *scratch = nullptr;
    size_t cells = 0;
    float root_hmin = 0.f, root_hmax = 0.f;
    int64_t bytes = 0;
};

struct dmap_object
{
    dmap_ctx *ctx = nullptr;
    dmap_tile *tile = nullptr;
    int n_tri = 0;
    int const_normal = 0;
    float bound_lo[3] = {0, 0, 0}, bound_hi[3] = {0, 0, 0};
    std::vector<TriRecord> host;
    std::vector<float> boxes_host, h_range_host;
    int bvh_leaves = 1;
    int bvh_builder = DMAP_BVH_AUTO; // dmap_object_create_ex's choice, kept for dmap_object_refresh
    TriRecord *tris = nullptr;
    float *boxes = nullptr;
    BvhNode *bvh = nullptr; // bvh_leaves nodes, then bvh_leaves leaf entries (int)
    ObjectDev view{}; // what a launch passes in its arguments
    std::vector<double> positions, normals, uvs; // the input triangles, for dmap_object_refresh
};


namespace dmap_amd_internal
{
int fail(int code, const char *fmt, ...);
}
using dmap_amd_internal::fail;

#define HIP_TRY(call)                                                                                                 \
    do {                                                                                                              \
        hipError_t e_ = (call);                                                                                       \
        if (e_ != hipSuccess)                                                                                         \
            return fail(e_ == hipErrorOutOfMemory ? DMAP_ERR_MEMORY : DMAP_ERR_HIP, "%s: %s", #call,                  \
                        hipGetErrorString(e_));                                                                       \
    } while (0)
