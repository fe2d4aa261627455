// dmap_amd C++ API: the reference's intersector interface over the C ABI of
// libdmap_amd.so (dmap_amd.h), for C++ callers such as the reference's own
// renderer and validation tools. Header only; needs no HIP headers (device
// memory goes through dmap_device_alloc / dmap_memcpy).
//
// Reference types mirrored (code/cpp/src/):
//   TraversalOptions, BoundMode, LeafMode   traversal_options.h:18-38
//   DisplacedHit                            displaced_intersector.h:107-114
//   TraversalStats (per-ray counters)       displaced_intersector.h:89-105
//   RayQuery, TraceResult, trace()          traversal_test_util.h:20-35
//   DisplacementTile                        DisplacementAsset + TaylorPyramid
//                                           (displacement_asset.cpp:16-35)
//   ObjectIntersector                       displaced_intersector.h:176-198
// The difference: calls take batches of rays (host vectors or device
// arrays), not one ray, and run on the GPU.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "dmap_amd.h"

namespace dmap_amd
{

class Error : public std::runtime_error
{
  public:
    Error(int code, const std::string &what) : std::runtime_error(what), code(code) {}
    int code;
};

inline void check(int rc, const char *call)
{
    if (rc != DMAP_OK)
        throw Error(rc, std::string(call) + ": " + dmap_last_error());
}

enum class BoundMode { Box = 0, Slab = 1, Both = 2 };
enum class LeafMode { Newton = 0, TwoTriangle = 1 };

struct TraversalOptions
{
    BoundMode bound = BoundMode::Both;
    int slab_max_level = 1;
    LeafMode leaf = LeafMode::Newton;
    bool fp64 = false;   // the same solve in double on the device
    // the ray order: DMAP_ORDER_NEAR_FIRST (rays starting inside the object's bounds first,
    // longest path through the bounds first),
    // DMAP_ORDER_GIVEN, or DMAP_ORDER_SORT (by octant and origin: pays for shuffled kinds of rays)
    int reorder = DMAP_ORDER_NEAR_FIRST;
    bool polish = true;  // polish the reported root in double (false: the search's own root)
    bool from_roots = false; // every walk starts at the triangle's roots, as the reference's does

    dmap_options c() const
    {
        return dmap_options{(int)bound, slab_max_level, (int)leaf, fp64 ? 1 : 0, reorder, polish ? 1 : 0,
                            from_roots ? 1 : 0};
    }
};

struct DisplacedHit
{
    double t = INFINITY;
    double st[2] = {0, 0};   // tile coordinates
    double bary[2] = {0, 0}; // weights of p1, p2
    double ng[3] = {0, 0, 0};
    int64_t i = 0, j = 0; // the leaf texel
};

struct TraversalStats
{
    static constexpr int max_levels = DMAP_STAT_LEVELS;
    int64_t triangle_tests = 0, node_tests = 0, leaf_tests = 0, sub_squares = 0;
    int64_t node_tests_per_level[max_levels] = {};
    int64_t node_tests_after_hit = 0, leaf_tests_after_hit = 0;
};

struct RayQuery
{
    double o[3], d[3];
    double t_min = 0.0, t_max = INFINITY;
};

struct TraceResult
{
    bool hit = false;
    DisplacedHit h;
    int triangle = -1;
    TraversalStats stats;
};

// One device, the library's context.
class Context
{
  public:
    explicit Context(int device = 0) { check(dmap_ctx_create(device, &h_), "dmap_ctx_create"); }
    ~Context() { dmap_ctx_destroy(h_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    dmap_ctx *get() const { return h_; }

  private:
    dmap_ctx *h_ = nullptr;
};

// A device array of T (owning).
template <class T> class DeviceArray
{
  public:
    DeviceArray(Context &ctx, size_t n) : ctx_(&ctx), n_(n)
    {
        void *p = nullptr;
        check(dmap_device_alloc(ctx.get(), (int64_t)(n * sizeof(T)), &p), "dmap_device_alloc");
        p_ = static_cast<T *>(p);
    }
    DeviceArray(Context &ctx, const std::vector<T> &host) : DeviceArray(ctx, host.size()) { upload(host); }
    ~DeviceArray()
    {
        if (p_)
            dmap_device_free(ctx_->get(), p_);
    }
    DeviceArray(DeviceArray &&o) noexcept : ctx_(o.ctx_), p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;

    void upload(const std::vector<T> &host)
    {
        if (host.size() != n_)
            throw Error(DMAP_ERR_INVALID, "DeviceArray::upload: size mismatch");
        check(dmap_memcpy(p_, host.data(), (int64_t)(n_ * sizeof(T)), 0, nullptr), "dmap_memcpy");
    }
    std::vector<T> download() const
    {
        std::vector<T> host(n_);
        check(dmap_memcpy(host.data(), p_, (int64_t)(n_ * sizeof(T)), 1, nullptr), "dmap_memcpy");
        return host;
    }
    T *data() const { return p_; }
    size_t size() const { return n_; }

  private:
    Context *ctx_;
    T *p_ = nullptr;
    size_t n_ = 0;
};

// A batch of rays in device SoA (fp32, as the ABI takes them).
struct DeviceRays
{
    DeviceArray<float> a; // 8 planes: ox oy oz dx dy dz tmin tmax
    size_t n;
    DeviceRays(Context &ctx, const std::vector<RayQuery> &rays) : a(ctx, 8 * rays.size()), n(rays.size())
    {
        std::vector<float> h(8 * n);
        for (size_t r = 0; r < n; ++r) {
            for (int q = 0; q < 3; ++q) {
                h[q * n + r] = (float)rays[r].o[q];
                h[(3 + q) * n + r] = (float)rays[r].d[q];
            }
            h[6 * n + r] = (float)rays[r].t_min;
            h[7 * n + r] = (float)rays[r].t_max;
        }
        a.upload(h);
    }
    dmap_rays c() const
    {
        float *f = a.data();
        return dmap_rays{f, f + n, f + 2 * n, f + 3 * n, f + 4 * n, f + 5 * n, f + 6 * n, f + 7 * n};
    }
};

// Hits of a batch in device SoA: 8 fp32 planes (t s tt bu bv nx ny nz) + 3
// int32 planes (prim texel_i texel_j).
struct DeviceHits
{
    DeviceArray<float> f;
    DeviceArray<int32_t> i;
    size_t n;
    DeviceHits(Context &ctx, size_t n) : f(ctx, 8 * n), i(ctx, 3 * n), n(n) {}
    dmap_hits c() const
    {
        float *a = f.data();
        int32_t *b = i.data();
        return dmap_hits{a, a + n, a + 2 * n, a + 3 * n, a + 4 * n, a + 5 * n, a + 6 * n, a + 7 * n, b, b + n, b + 2 * n};
    }
    std::vector<TraceResult> download() const
    {
        std::vector<float> a = f.download();
        std::vector<int32_t> b = i.download();
        std::vector<TraceResult> out(n);
        for (size_t r = 0; r < n; ++r) {
            TraceResult &x = out[r];
            x.triangle = b[r];
            x.hit = x.triangle >= 0;
            x.h.t = a[r];
            x.h.st[0] = a[n + r];
            x.h.st[1] = a[2 * n + r];
            x.h.bary[0] = a[3 * n + r];
            x.h.bary[1] = a[4 * n + r];
            for (int q = 0; q < 3; ++q)
                x.h.ng[q] = a[(5 + q) * n + r];
            x.h.i = b[n + r];
            x.h.j = b[2 * n + r];
        }
        return out;
    }
};

// A resident tile: texels (n x n, row-major [j][i]) with h = strength *
// (texel - midlevel), as DisplacementAsset builds it.
class DisplacementTile
{
  public:
    DisplacementTile(Context &ctx, const std::vector<float> &texels, int n, double strength, double midlevel = 0.0,
                     bool repeat = true, bool direct_build = false)
    {
        if (texels.size() != (size_t)n * n)
            throw Error(DMAP_ERR_INVALID, "DisplacementTile: texels must be n x n");
        check(dmap_tile_create(ctx.get(), texels.data(), n, 0, repeat ? 1 : 0, strength, -strength * midlevel,
                               direct_build ? 1 : 0, nullptr, &h_),
              "dmap_tile_create");
        int nl = 0;
        int64_t bytes = 0;
        check(dmap_tile_info(h_, &n_leaf, &nl, &bytes), "dmap_tile_info");
    }
    ~DisplacementTile() { dmap_tile_destroy(h_); }
    DisplacementTile(const DisplacementTile &) = delete;
    DisplacementTile &operator=(const DisplacementTile &) = delete;
    dmap_tile *get() const { return h_; }
    // The closed node grid ((n+1)^2 raw texel values).
    std::vector<float> nodes() const
    {
        std::vector<float> out((size_t)(n_leaf + 1) * (n_leaf + 1));
        check(dmap_tile_read_nodes(h_, out.data()), "dmap_tile_read_nodes");
        return out;
    }
    int n_leaf = 0;

  private:
    dmap_tile *h_ = nullptr;
};

// Base triangles over one tile: 24 doubles each (p0 p1 p2, m0 m1 m2, t0 t1 t2).
class ObjectIntersector
{
  public:
    ObjectIntersector(Context &ctx, DisplacementTile &tile, const std::vector<double> &triangles) : ctx_(&ctx)
    {
        if (triangles.size() % 24)
            throw Error(DMAP_ERR_INVALID, "ObjectIntersector: 24 doubles per triangle");
        n_tri = (int)(triangles.size() / 24);
        std::vector<double> pos(9 * n_tri), nrm(9 * n_tri), uvs(6 * n_tri);
        for (int k = 0; k < n_tri; ++k) {
            std::memcpy(&pos[9 * k], &triangles[24 * k], 9 * sizeof(double));
            std::memcpy(&nrm[9 * k], &triangles[24 * k + 9], 9 * sizeof(double));
            std::memcpy(&uvs[6 * k], &triangles[24 * k + 18], 6 * sizeof(double));
        }
        check(dmap_object_create(ctx.get(), tile.get(), n_tri, pos.data(), nrm.data(), uvs.data(), &h_),
              "dmap_object_create");
    }
    ~ObjectIntersector() { dmap_object_destroy(h_); }
    ObjectIntersector(const ObjectIntersector &) = delete;
    ObjectIntersector &operator=(const ObjectIntersector &) = delete;
    dmap_object *get() const { return h_; }

    // Closest hits of a device batch (asynchronous on `stream`).
    void intersect(const DeviceRays &rays, DeviceHits &hits, const TraversalOptions &options = {}, int prim_mask = -1,
                   void *stream = nullptr) const
    {
        dmap_rays r = rays.c();
        dmap_hits h = hits.c();
        dmap_options o = options.c();
        check(dmap_intersect_closest(h_, &r, (int64_t)rays.n, &h, prim_mask, &o, nullptr, stream),
              "dmap_intersect_closest");
    }

    // Closest hits of host rays (ObjectIntersector::intersect over a batch;
    // with stats, the per-ray work counters). With records (and stats),
    // every node test of the walks is appended to *records, as
    // TraversalStats::records collects them (displaced_intersector.h:
    // 101-102), up to record_capacity; rays interleave, each ray's records
    // in seq order.
    std::vector<TraceResult> intersect(const std::vector<RayQuery> &rays, const TraversalOptions &options = {},
                                       bool stats = false, int prim_mask = -1,
                                       std::vector<dmap_node_record> *records = nullptr,
                                       int64_t record_capacity = 0) const
    {
        if (rays.empty())
            return {};
        DeviceRays r(*ctx_, rays);
        DeviceHits h(*ctx_, rays.size());
        dmap_rays rc = r.c();
        dmap_hits hc = h.c();
        dmap_options o = options.c();
        const size_t n = rays.size();
        constexpr int planes = 6 + DMAP_STAT_LEVELS;
        DeviceArray<int32_t> counters(*ctx_, stats ? planes * n : 0);
        const bool rec = stats && records && record_capacity > 0;
        DeviceArray<dmap_node_record> rec_buf(*ctx_, rec ? (size_t)record_capacity : 0);
        DeviceArray<uint64_t> rec_count(*ctx_, rec ? 1 : 0);
        dmap_ray_stats st{};
        if (stats) {
            int32_t *c = counters.data();
            st = dmap_ray_stats{c,       c + n,     c + 2 * n, c + 3 * n, nullptr,
                                nullptr, c + 4 * n, c + 5 * n, c + 6 * n};
        }
        if (rec) {
            rec_count.upload({0});
            st.records = rec_buf.data();
            st.record_capacity = record_capacity;
            st.record_count = rec_count.data();
        }
        check(dmap_intersect_closest(h_, &rc, (int64_t)n, &hc, prim_mask, &o, stats ? &st : nullptr, nullptr),
              "dmap_intersect_closest");
        check(dmap_stream_synchronize(nullptr), "dmap_stream_synchronize");
        std::vector<TraceResult> out = h.download();
        if (stats) {
            std::vector<int32_t> c = counters.download();
            for (size_t q = 0; q < n; ++q) {
                out[q].stats.triangle_tests = c[q];
                out[q].stats.node_tests = c[n + q];
                out[q].stats.leaf_tests = c[2 * n + q];
                out[q].stats.sub_squares = c[3 * n + q];
                out[q].stats.node_tests_after_hit = c[4 * n + q];
                out[q].stats.leaf_tests_after_hit = c[5 * n + q];
                for (int l = 0; l < DMAP_STAT_LEVELS; ++l)
                    out[q].stats.node_tests_per_level[l] = c[(6 + l) * n + q];
            }
        }
        if (rec) {
            const uint64_t made = rec_count.download()[0];
            std::vector<dmap_node_record> all = rec_buf.download();
            all.resize((size_t)std::min<uint64_t>(made, (uint64_t)record_capacity));
            records->insert(records->end(), all.begin(), all.end());
        }
        return out;
    }

    // Any hit of host rays (the occluded callback).
    std::vector<uint8_t> occluded(const std::vector<RayQuery> &rays, const TraversalOptions &options = {},
                                  int prim_mask = -1) const
    {
        if (rays.empty())
            return {};
        DeviceRays r(*ctx_, rays);
        DeviceArray<uint8_t> occ(*ctx_, rays.size());
        dmap_rays rc = r.c();
        dmap_options o = options.c();
        check(dmap_intersect_any(h_, &rc, (int64_t)rays.size(), occ.data(), prim_mask, &o, nullptr),
              "dmap_intersect_any");
        check(dmap_stream_synchronize(nullptr), "dmap_stream_synchronize");
        return occ.download();
    }

    // Wavefront path tracing of this object (dmap_pathtrace): camera rays
    // as host queries (origin, direction per pixel), the light as corner,
    // edges and the unit normal of its emitting side. Returns the mean
    // radiance per pixel; ray counts go to *closest_rays / *shadow_rays.
    struct PathScene
    {
        double light_corner[3] = {3, 3, 3}, light_u[3] = {0, 2, 0}, light_v[3] = {2, 0, 0}, light_n[3] = {0, 0, -1};
        double radiance = 10.0, env = 0.2, albedo = 0.7;
    };
    std::vector<double> pathtrace(const std::vector<RayQuery> &camera, int spp, int depth, const PathScene &scene,
                                  uint32_t seed = 0, bool mis = true, const TraversalOptions &options = {},
                                  int64_t *closest_rays = nullptr, int64_t *shadow_rays = nullptr) const
    {
        const size_t n = camera.size();
        if (n == 0)
            return {};
        std::vector<double> o(3 * n), d(3 * n);
        for (size_t r = 0; r < n; ++r)
            for (int q = 0; q < 3; ++q) {
                o[q * n + r] = camera[r].o[q];
                d[q * n + r] = camera[r].d[q];
            }
        DeviceArray<double> o_dev(*ctx_, o), d_dev(*ctx_, d), image(*ctx_, n);
        dmap_pt_scene sc;
        for (int q = 0; q < 3; ++q) {
            sc.light_corner[q] = scene.light_corner[q];
            sc.light_u[q] = scene.light_u[q];
            sc.light_v[q] = scene.light_v[q];
            sc.light_n[q] = scene.light_n[q];
        }
        sc.radiance = scene.radiance;
        sc.env = scene.env;
        sc.albedo = scene.albedo;
        dmap_options opt = options.c();
        check(dmap_pathtrace(h_, o_dev.data(), d_dev.data(), (int64_t)n, spp, depth, seed, mis ? 0 : 1, &sc, &opt,
                             image.data(), closest_rays, shadow_rays, nullptr),
              "dmap_pathtrace");
        check(dmap_stream_synchronize(nullptr), "dmap_stream_synchronize");
        return image.download();
    }

    int n_tri = 0;

  private:
    Context *ctx_;
    dmap_object *h_ = nullptr;
};

// trace (traversal_test_util.h:32-35): one ray from t = 0 to infinity.
inline TraceResult trace(const ObjectIntersector &obj, const RayQuery &r, const TraversalOptions &options)
{
    return obj.intersect(std::vector<RayQuery>{r}, options).front();
}

} // namespace dmap_amd
