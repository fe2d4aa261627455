// Wavefront path tracing of one displaced object on the device: the
// estimator of the path-tracing workload (a Lambertian displaced surface, one
// one-sided rectangular light, a constant environment, NEE with the power
// heuristic against cosine-sampled BSDF rays), with the intersections done by
// the library's closest-hit and any-hit calls.
//
// The reference renders through ks's SmallPT one ray at a time
// (path_trace.cpp:107-129, direct_lighting.cpp:155-232; ks is not part of
// the reference repository). Here every path of a chunk advances one bounce
// per round:
//
//   trace closest (all live paths)
//   k_pt_shade: light / environment reached, then for the surviving paths the
//               next-event ray and the cosine-sampled continuation, both
//               compacted into fresh queues with one atomic per workgroup
//   trace any-hit (the next-event rays), k_pt_shadow: add the unoccluded ones
//
// Each path sums its radiance in its own slot (one lane owns a path at a
// time, so no atomics), and k_pt_gather averages a pixel's samples once the
// chunk is done; chunks hold whole pixels.
//
// All shading arithmetic is in double, with contraction off, so that the
// same paths through another intersector give the same image; rays go to
// the intersector in fp32, as its ABI takes them. Random numbers are a
// counter-based hash of (path, bounce, dimension): path = pixel * spp +
// sample.
#include <cmath>
#include <cstring>

#include "internal.h"

#pragma clang fp contract(off)

namespace
{

constexpr double LIGHT_EPS = 1e-4; // shadow rays stop this short of the light
constexpr double SPAWN_EPS = 1e-4; // new rays start this far off the surface
constexpr double PI = 3.141592653589793;
constexpr int BLOCK = 256;
// The shading kernel's workgroups: 16 waves share one queue atomic.
constexpr int SHADE_BLOCK = 1024;
// Paths per wavefront chunk: about 270 B of queues each.
constexpr int64_t MAX_CHUNK = (int64_t)1 << 22;

struct Light
{
    double c[3], u[3], v[3], n[3];
    double area, radiance, env, albedo;
};

// One queue of paths: pid, throughput, the previous BSDF pdf (0: camera
// ray), origin and direction in double.
struct PathQueue
{
    uint32_t *pid;
    double *beta, *pdf_prev;
    double *o[3], *d[3];
};

// The intersector's fp32 ray planes.
struct RayPlanes
{
    float *p[8];
};

__device__ __forceinline__ uint32_t mix32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// Uniform [0, 1) for (path, bounce, dimension); the same integers as the
// host-side estimator's hash.
__device__ __forceinline__ double uniform01(uint32_t pid, uint32_t bounce, uint32_t dim, uint32_t seed)
{
    uint32_t key = mix32(pid * 0x9E3779B1u + seed);
    key = mix32(key ^ (bounce * 0x85EBCA77u + dim * 0xC2B2AE3Du));
    return ((double)key + 0.5) / 4294967296.0;
}

// Slots in two compacted queues for the lanes of a workgroup that want
// them (want0, want1): one atomic per queue per workgroup, the workgroup's
// waves in order and each wave's lanes in order after it. Every thread of
// the workgroup calls it. (One atomic per wave on the one counter serialized
// at about 8 ns each: 1.1 ms for the 4M paths of a bounce.)
__device__ __forceinline__ void block_claim(bool want0, bool want1, uint32_t *counters, uint32_t &slot0,
                                            uint32_t &slot1)
{
    __shared__ uint32_t s_cnt[2][SHADE_BLOCK / 64];
    __shared__ uint32_t s_base[2];
    const uint64_t m0 = __ballot(want0), m1 = __ballot(want1);
    const int lane = __lane_id(), wv = threadIdx.x >> 6;
    if (lane == 0) {
        s_cnt[0][wv] = (uint32_t)__popcll(m0);
        s_cnt[1][wv] = (uint32_t)__popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < 2; ++q) {
            uint32_t tot = 0;
            for (int w = 0; w < SHADE_BLOCK / 64; ++w)
                tot += s_cnt[q][w];
            s_base[q] = tot ? atomicAdd(counters + q, tot) : 0u;
        }
    }
    __syncthreads();
    uint32_t b0 = s_base[0], b1 = s_base[1];
    for (int w = 0; w < wv; ++w) {
        b0 += s_cnt[0][w];
        b1 += s_cnt[1][w];
    }
    const uint64_t below = (1ull << lane) - 1ull;
    slot0 = b0 + (uint32_t)__popcll(m0 & below);
    slot1 = b1 + (uint32_t)__popcll(m1 & below);
}

__device__ __forceinline__ double dot3(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double r[3])
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ void put_ray(const RayPlanes &R, int64_t k, const double o[3], const double d[3],
                                       double tmax)
{
    for (int q = 0; q < 3; ++q) {
        R.p[q][k] = (float)o[q];
        R.p[3 + q][k] = (float)d[q];
    }
    R.p[6][k] = 0.0f;
    R.p[7][k] = (float)tmax;
}

__global__ void k_pt_camera(int64_t first, int64_t count, int spp, const double *cam_o, const double *cam_d,
                            int64_t n_pix, PathQueue Q, RayPlanes R, double *radiance)
{
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= count)
        return;
    int64_t pid = first + i;
    int64_t pix = pid / spp;
    double o[3], d[3];
    for (int q = 0; q < 3; ++q) {
        o[q] = cam_o[q * n_pix + pix];
        d[q] = cam_d[q * n_pix + pix];
    }
    double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    for (int q = 0; q < 3; ++q) {
        d[q] = d[q] / len;
        Q.o[q][i] = o[q];
        Q.d[q][i] = d[q];
    }
    Q.pid[i] = (uint32_t)pid;
    Q.beta[i] = 1.0;
    Q.pdf_prev[i] = 0.0;
    radiance[i] = 0.0;
    put_ray(R, i, o, d, INFINITY);
}

struct ShadeArgs
{
    PathQueue in, out;
    RayPlanes next_rays, shadow_rays;
    const float *hit_t, *hit_n[3];
    double *shadow_contrib;
    uint32_t *shadow_pid;
    uint32_t *counters; // [0] next paths, [1] shadow rays
    double *radiance;   // per path of the chunk, indexed pid - first
    int64_t first;
    Light L;
    int64_t n;
    int spp, bounce, last, mis;
    uint32_t seed;
};

__global__ void __launch_bounds__(SHADE_BLOCK) k_pt_shade(ShadeArgs a)
{
    int64_t i = (int64_t)blockIdx.x * SHADE_BLOCK + threadIdx.x;
    bool live = i < a.n;
    bool cont = false, nee = false;
    uint32_t pid = 0;
    double beta = 0, origin[3] = {0, 0, 0}, nd[3] = {0, 0, 0}, pdf_next = 0, contrib = 0, wi[3] = {0, 0, 0},
           dist = 0;
    const Light &L = a.L;
    if (live) {
        pid = a.in.pid[i];
        double &Lp = a.radiance[pid - a.first];
        beta = a.in.beta[i];
        double pdf_prev = a.in.pdf_prev[i];
        double o[3], d[3];
        for (int q = 0; q < 3; ++q) {
            o[q] = a.in.o[q][i];
            d[q] = a.in.d[q][i];
        }
        double t_surf = (double)a.hit_t[i];
        // the light's front face along the ray
        double dn = dot3(d, L.n);
        double cm[3] = {L.c[0] - o[0], L.c[1] - o[1], L.c[2] - o[2]};
        double t_light = dot3(cm, L.n) / dn;
        double p[3];
        for (int q = 0; q < 3; ++q)
            p[q] = o[q] + t_light * d[q] - L.c[q];
        double uu = dot3(p, L.u) / dot3(L.u, L.u);
        double vv = dot3(p, L.v) / dot3(L.v, L.v);
        bool ok_l = (dn < 0) && (t_light > 0) && (uu >= 0) && (uu <= 1) && (vv >= 0) && (vv <= 1);
        if (!ok_l)
            t_light = INFINITY;
        bool on_light = t_light < t_surf;
        bool hit = isfinite(t_surf);
        if (on_light) {
            double w = 1.0;
            if (a.mis && pdf_prev > 0) {
                double cos_l = -dot3(d, L.n);
                double pdf_l = (t_light * t_light) / (L.area * fmax(cos_l, 1e-12));
                w = pdf_prev * pdf_prev / (pdf_prev * pdf_prev + pdf_l * pdf_l);
            }
            Lp += beta * w * L.radiance;
        } else if (!hit) {
            Lp += beta * L.env;
        } else if (!a.last) {
            double ng[3];
            for (int q = 0; q < 3; ++q)
                ng[q] = (double)a.hit_n[q][i];
            double ps[3];
            for (int q = 0; q < 3; ++q)
                ps[q] = o[q] + t_surf * d[q];
            // shade on the side the ray came from
            if (dot3(ng, d) > 0)
                for (int q = 0; q < 3; ++q)
                    ng[q] = -ng[q];
            double ax[3] = {0, 0, 0};
            ax[fabs(ng[0]) < 0.9 ? 0 : 1] = 1.0;
            double t1[3], t2[3];
            cross3(ax, ng, t1);
            double l1 = sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
            for (int q = 0; q < 3; ++q)
                t1[q] = t1[q] / l1;
            cross3(ng, t1, t2);
            double f = L.albedo / PI;
            for (int q = 0; q < 3; ++q)
                origin[q] = ps[q] + SPAWN_EPS * ng[q];
            if (a.mis) {
                double ua = uniform01(pid, a.bounce, 0, a.seed);
                double va = uniform01(pid, a.bounce, 1, a.seed);
                double to_l[3];
                for (int q = 0; q < 3; ++q)
                    to_l[q] = L.c[q] + ua * L.u[q] + va * L.v[q] - origin[q];
                double dist2 = dot3(to_l, to_l);
                dist = sqrt(dist2);
                for (int q = 0; q < 3; ++q)
                    wi[q] = to_l[q] / dist;
                double cos_s = dot3(wi, ng);
                double cos_l = -dot3(wi, L.n);
                if (cos_s > 0 && cos_l > 0) {
                    nee = true;
                    double pdf_l = dist2 / (L.area * fmax(cos_l, 1e-12));
                    double pdf_b = fmax(cos_s, 0.0) / PI;
                    double w = pdf_l * pdf_l / (pdf_l * pdf_l + pdf_b * pdf_b);
                    contrib = beta * f * cos_s * L.radiance * w / pdf_l;
                }
            }
            // cosine-weighted continuation: beta *= f cos / pdf = albedo
            double u1 = uniform01(pid, a.bounce, 2, a.seed);
            double u2 = uniform01(pid, a.bounce, 3, a.seed);
            double r = sqrt(u1), phi = 2.0 * PI * u2;
            double lx = r * cos(phi), ly = r * sin(phi), lz = sqrt(fmax(1.0 - u1, 0.0));
            for (int q = 0; q < 3; ++q)
                nd[q] = lx * t1[q] + ly * t2[q] + lz * ng[q];
            beta = beta * L.albedo;
            pdf_next = lz / PI;
            cont = true;
        }
    }
    uint32_t kc, ks;
    block_claim(cont, nee, a.counters, kc, ks);
    if (nee) {
        put_ray(a.shadow_rays, ks, origin, wi, dist - LIGHT_EPS);
        a.shadow_contrib[ks] = contrib;
        a.shadow_pid[ks] = pid;
    }
    if (cont) {
        a.out.pid[kc] = pid;
        a.out.beta[kc] = beta;
        a.out.pdf_prev[kc] = pdf_next;
        for (int q = 0; q < 3; ++q) {
            a.out.o[q][kc] = origin[q];
            a.out.d[q][kc] = nd[q];
        }
        put_ray(a.next_rays, kc, origin, nd, INFINITY);
    }
}

__global__ void k_pt_shadow(int64_t n, const uint8_t *occluded, const double *contrib, const uint32_t *pid,
                            int64_t first, double *radiance)
{
    int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n || occluded[i])
        return;
    radiance[pid[i] - first] += contrib[i];
}

// The mean over each pixel's samples: pixels first_pix .. first_pix + n - 1,
// whose samples are consecutive paths of the chunk.
__global__ void k_pt_gather(int64_t first_pix, int64_t n, int spp, const double *radiance, double *image)
{
    int64_t p = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
    if (p >= n)
        return;
    double sum = 0.0;
    for (int k = 0; k < spp; ++k)
        sum += radiance[p * spp + k];
    image[first_pix + p] = sum / spp;
}

unsigned grid_of(int64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

// Carves the workspace into the chunk's queues (256-byte aligned arrays of
// cap entries); with base NULL it only measures.
struct Workspace
{
    PathQueue q[2];
    RayPlanes rays, shadow;
    float *hit_f[8];   // t, s, tt, bu, bv, nx, ny, nz (bu, bv unused)
    int32_t *hit_i[3]; // prim, texel_i, texel_j
    double *contrib;
    uint32_t *shadow_pid;
    uint8_t *occluded;
    double *radiance;
    uint32_t *counters;

    size_t carve(char *base, int64_t cap)
    {
        size_t off = 0;
        auto take = [&](size_t each) {
            char *p = base ? base + off : nullptr;
            off += ((size_t)cap * each + 255) & ~(size_t)255;
            return p;
        };
        for (int k = 0; k < 2; ++k) {
            q[k].beta = (double *)take(8);
            q[k].pdf_prev = (double *)take(8);
            for (int c = 0; c < 3; ++c) {
                q[k].o[c] = (double *)take(8);
                q[k].d[c] = (double *)take(8);
            }
            q[k].pid = (uint32_t *)take(4);
        }
        for (int c = 0; c < 8; ++c)
            rays.p[c] = (float *)take(4);
        for (int c = 0; c < 8; ++c)
            shadow.p[c] = (float *)take(4);
        for (int c = 0; c < 8; ++c)
            hit_f[c] = (float *)take(4);
        for (int c = 0; c < 3; ++c)
            hit_i[c] = (int32_t *)take(4);
        contrib = (double *)take(8);
        shadow_pid = (uint32_t *)take(4);
        occluded = (uint8_t *)take(1);
        radiance = (double *)take(8);
        counters = base ? (uint32_t *)(base + off) : nullptr;
        return off + 256;
    }
};

dmap_rays as_rays(const RayPlanes &R)
{
    return dmap_rays{R.p[0], R.p[1], R.p[2], R.p[3], R.p[4], R.p[5], R.p[6], R.p[7]};
}

} // namespace

extern "C" {

int dmap_pathtrace(const dmap_object *obj, const double *cam_o, const double *cam_d, int64_t n_pix, int spp,
                   int depth, uint32_t seed, int strategy, const dmap_pt_scene *scene, const dmap_options *options,
                   double *image, int64_t *closest_rays, int64_t *shadow_rays, void *stream)
{
    if (!obj || !scene || (n_pix > 0 && (!image || !cam_o || !cam_d)))
        return fail(DMAP_ERR_INVALID, "dmap_pathtrace: NULL argument");
    if (n_pix < 0 || spp < 1 || depth < 0)
        return fail(DMAP_ERR_INVALID, "dmap_pathtrace: n_pix %lld, spp %d, depth %d", (long long)n_pix, spp, depth);
    if (strategy != 0 && strategy != 1)
        return fail(DMAP_ERR_INVALID, "dmap_pathtrace: strategy %d (0 NEE + MIS, 1 BSDF only)", strategy);
    const int64_t n_paths = n_pix * (int64_t)spp;
    if (n_paths >= ((int64_t)1 << 32))
        return fail(DMAP_ERR_INVALID, "dmap_pathtrace: n_pix * spp must stay below 2^32");
    Light L;
    for (int q = 0; q < 3; ++q) {
        L.c[q] = scene->light_corner[q];
        L.u[q] = scene->light_u[q];
        L.v[q] = scene->light_v[q];
        L.n[q] = scene->light_n[q];
    }
    double un[3] = {L.u[1] * L.v[2] - L.u[2] * L.v[1], L.u[2] * L.v[0] - L.u[0] * L.v[2],
                    L.u[0] * L.v[1] - L.u[1] * L.v[0]};
    L.area = std::sqrt(un[0] * un[0] + un[1] * un[1] + un[2] * un[2]);
    L.radiance = scene->radiance;
    L.env = scene->env;
    L.albedo = scene->albedo;
    if (!(L.area > 0))
        return fail(DMAP_ERR_INVALID, "dmap_pathtrace: the light's edges span no area");
    // The queues stay in pixel order (camera rays by pixel, later bounces
    // compacted wave by wave in that order), which is coherent already:
    // sorting each batch cost more than it saved (20.4 against 17.2 ms for
    // 1024^2 x 4 spp on the 4096^2 plane), and so did the near-first order,
    // whose length classes scatter neighbouring pixels' bounces (12.2-12.4
    // against 11.6 ms, the same image), so the calls below trace as given.
    dmap_options opt{2, 1, 0, 0, 0, 1, 0};
    if (options)
        opt = *options;
    opt.reorder = 0;
    HIP_TRY(hipSetDevice(obj->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    int64_t cnt_closest = 0, cnt_shadow = 0;
    // Chunks of whole pixels.
    const int64_t cap = n_paths < MAX_CHUNK ? n_paths : (MAX_CHUNK / spp) * spp;
    if (n_paths > 0 && cap <= 0)
        return fail(DMAP_ERR_INVALID, "dmap_pathtrace: spp %d exceeds the %lld paths of one wavefront", spp,
                    (long long)MAX_CHUNK);
    char *ws = nullptr;
    Workspace W;
    if (cap > 0)
        HIP_TRY(hipMallocFromPoolAsync((void **)&ws, W.carve(nullptr, cap), obj->ctx->pool, s));
    if (ws)
        W.carve(ws, cap);
    uint32_t *pinned = nullptr;
    int rc = DMAP_OK;
    if (cap > 0 && hipHostMalloc((void **)&pinned, 2 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
        (void)hipFreeAsync(ws, s);
        return fail(DMAP_ERR_MEMORY, "dmap_pathtrace: pinned counters");
    }
    // A bounce's next-event rays are traced on a second stream while the
    // next bounce's closest-hit batch runs on `s` (the tails of the two
    // persistent launches overlap); their contributions are added on `s`
    // before the next shade, in the same order as without the overlap, so
    // the image does not change. Both streams keep the default priority.
    hipStream_t s2 = nullptr;
    hipEvent_t ev_rays = nullptr, ev_occluded = nullptr;
    if (cap > 0 && (hipStreamCreateWithFlags(&s2, hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreateWithFlags(&ev_rays, hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&ev_occluded, hipEventDisableTiming) != hipSuccess))
        rc = fail(DMAP_ERR_HIP, "dmap_pathtrace: side stream");
    int64_t pending = 0; // next-event rays traced on s2, not yet added on s
    auto add_pending = [&](int64_t first) {
        if (pending == 0)
            return true;
        if (hipStreamWaitEvent(s, ev_occluded, 0) != hipSuccess)
            return false;
        hipLaunchKernelGGL(k_pt_shadow, dim3(grid_of(pending)), dim3(BLOCK), 0, s, pending, W.occluded, W.contrib,
                           W.shadow_pid, first, W.radiance);
        pending = 0;
        return true;
    };
    for (int64_t first = 0; first < n_paths && rc == DMAP_OK; first += cap) {
        int64_t count = n_paths - first < cap ? n_paths - first : cap;
        hipLaunchKernelGGL(k_pt_camera, dim3(grid_of(count)), dim3(BLOCK), 0, s, first, count, spp, cam_o, cam_d,
                           n_pix, W.q[0], W.rays, W.radiance);
        int cur = 0;
        int64_t live = count;
        for (int b = 0; b <= depth && live > 0; ++b) {
            dmap_rays R = as_rays(W.rays);
            dmap_hits H{W.hit_f[0], W.hit_f[1], W.hit_f[2], nullptr, nullptr, W.hit_f[5], W.hit_f[6], W.hit_f[7],
                        W.hit_i[0], W.hit_i[1], W.hit_i[2]};
            if ((rc = dmap_intersect_closest(obj, &R, live, &H, -1, &opt, nullptr, stream)) != DMAP_OK)
                break;
            cnt_closest += live;
            // The previous bounce's next-event contributions, before this
            // bounce's shade adds to the same paths and refills the
            // next-event queue.
            if (!add_pending(first)) {
                rc = fail(DMAP_ERR_HIP, "dmap_pathtrace: stream wait");
                break;
            }
            bool last = b == depth;
            if (hipMemsetAsync(W.counters, 0, 2 * sizeof(uint32_t), s) != hipSuccess) {
                rc = fail(DMAP_ERR_HIP, "dmap_pathtrace: counter reset");
                break;
            }
            ShadeArgs a;
            a.in = W.q[cur];
            a.out = W.q[cur ^ 1];
            a.next_rays = W.rays; // the closest call above has consumed them
            a.shadow_rays = W.shadow;
            a.hit_t = W.hit_f[0];
            for (int q = 0; q < 3; ++q)
                a.hit_n[q] = W.hit_f[5 + q];
            a.shadow_contrib = W.contrib;
            a.shadow_pid = W.shadow_pid;
            a.counters = W.counters;
            a.radiance = W.radiance;
            a.first = first;
            a.L = L;
            a.n = live;
            a.spp = spp;
            a.bounce = b;
            a.last = last ? 1 : 0;
            a.mis = strategy == 0 ? 1 : 0;
            a.seed = seed;
            hipLaunchKernelGGL(k_pt_shade, dim3((unsigned)((live + SHADE_BLOCK - 1) / SHADE_BLOCK)), dim3(SHADE_BLOCK),
                               0, s, a);
            if (last)
                break;
            if (hipMemcpyAsync(pinned, W.counters, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess) {
                rc = fail(DMAP_ERR_HIP, "dmap_pathtrace: counter read");
                break;
            }
            int64_t n_next = pinned[0], n_shadow = pinned[1];
            if (n_shadow > 0) {
                dmap_rays SR = as_rays(W.shadow);
                if (hipEventRecord(ev_rays, s) != hipSuccess || hipStreamWaitEvent(s2, ev_rays, 0) != hipSuccess) {
                    rc = fail(DMAP_ERR_HIP, "dmap_pathtrace: stream wait");
                    break;
                }
                if ((rc = dmap_intersect_any(obj, &SR, n_shadow, W.occluded, -1, &opt, s2)) != DMAP_OK)
                    break;
                if (hipEventRecord(ev_occluded, s2) != hipSuccess) {
                    rc = fail(DMAP_ERR_HIP, "dmap_pathtrace: event");
                    break;
                }
                cnt_shadow += n_shadow;
                pending = n_shadow;
            }
            live = n_next;
            cur ^= 1;
        }
        if (rc == DMAP_OK && !add_pending(first))
            rc = fail(DMAP_ERR_HIP, "dmap_pathtrace: stream wait");
        if (rc == DMAP_OK)
            hipLaunchKernelGGL(k_pt_gather, dim3(grid_of(count / spp)), dim3(BLOCK), 0, s, first / spp, count / spp,
                               spp, W.radiance, image);
    }
    hipError_t e = hipGetLastError();
    if (s2) {
        // Work still queued on s2 (after an error) ends before its buffers go.
        (void)hipStreamSynchronize(s2);
        (void)hipStreamDestroy(s2);
    }
    if (ev_rays)
        (void)hipEventDestroy(ev_rays);
    if (ev_occluded)
        (void)hipEventDestroy(ev_occluded);
    if (ws)
        (void)hipFreeAsync(ws, s);
    if (pinned) {
        (void)hipStreamSynchronize(s);
        (void)hipHostFree(pinned);
    }
    if (rc != DMAP_OK)
        return rc;
    if (e != hipSuccess)
        return fail(DMAP_ERR_HIP, "dmap_pathtrace: launch: %s", hipGetErrorString(e));
    if (closest_rays)
        *closest_rays = cnt_closest;
    if (shadow_rays)
        *shadow_rays = cnt_shadow;
    return DMAP_OK;
}

} // extern "C"
