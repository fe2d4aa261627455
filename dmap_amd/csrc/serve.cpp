// Serving runtime: continuous batching of closest-hit requests against
// resident objects (the reference has no counterpart: it traces one ray per
// call on host threads).
//
// A request is (object, device rays, count). The server keeps the submitted
// requests and `slots` steps in flight, each on its own stream. Whenever a
// slot is free it forms the next step from the waiting requests:
//  * order: requests that have waited longer than AGING_S first, in arrival
//    order; then the rest by rays still to trace, fewest first (shortest
//    remaining work first keeps small requests from queueing behind large
//    ones);
//  * memory-based admission: a request enters its first step only when the
//    hit pool has room for all its results (44 B per ray); a request that
//    does not fit waits, and an aged one that does not fit stops admission
//    so that it is next;
//  * chunking: a step holds at most step_rays rays; a larger request is
//    traced over several steps, possibly on several slots at once;
//  * one kind per step: objects with constant normals and objects with
//    interpolated normals use different kernel specialisations, so a step
//    takes requests of the kind of its first request only.
// A step is one persistent multi-request launch (reset, trace, finalize)
// captured once per slot and kind as a HIP graph and replayed with the
// step's descriptor, so requests that arrive while others run are admitted
// into the next free slot and overlap on the device with the steps already
// running. Each step records an event; the server's calls (submit, poll,
// release, idle) query the events, and a request's latency is the time its
// last step was seen complete minus its submission, so a caller that polls
// continuously measures it to within one poll.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "internal.h"

namespace
{

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one device arena (offsets in 256 B units).
struct Arena
{
    char *base = nullptr;
    size_t bytes = 0;
    std::map<size_t, size_t> free_; // offset -> size

    void init(char *b, size_t n)
    {
        base = b;
        bytes = n;
        free_.clear();
        free_[0] = n;
    }
    bool alloc(size_t n, size_t &off)
    {
        n = (n + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                size_t rest = it->second - n;
                free_.erase(it);
                if (rest)
                    free_[off + n] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = (n + 255) & ~(size_t)255;
        auto it = free_.emplace(off, n).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// Requests waiting longer than this go first, in arrival order.
constexpr double AGING_S = 0.02;
// Hit pool bytes per ray: 8 fp32 + 3 int32 hit fields.
constexpr size_t POOL_BYTES_PER_RAY = 44;

struct Request
{
    uint64_t tag = 0;
    uint64_t seq = 0; // arrival order
    const dmap_object *obj = nullptr;
    dmap_rays rays{};
    int64_t n = 0;
    int64_t next = 0; // rays handed to steps so far
    int inflight = 0; // steps holding a chunk of this request
    double t_submit = 0, t_done = 0;
    bool admitted = false; // has its hit arrays
    size_t off = 0, bytes = 0;
    dmap_hits hits{};
    bool done = false;
    std::vector<hipEvent_t> copies; // recorded after each dmap_server_copy_hits on the caller's stream
};

// A released request's pool region whose copies may still be queued on
// the callers' streams: it returns to the arena once all have completed.
struct Deferred
{
    size_t off = 0, bytes = 0;
    std::vector<hipEvent_t> copies;
};

struct Slot
{
    hipStream_t stream = nullptr;
    StepDesc *desc_host = nullptr; // pinned
    StepDesc *desc_dev = nullptr;
    unsigned *queue = nullptr;
    hipGraphExec_t graph[2] = {nullptr, nullptr}; // by kind: interpolated, constant normals
    hipEvent_t done = nullptr;                     // recorded after the step's launches
    std::vector<Request *> reqs;
    bool busy = false;
};

} // namespace

struct dmap_server
{
    dmap_ctx *ctx = nullptr;
    int64_t step_rays = 0;
    std::vector<Slot *> slots;
    Arena arena;
    char *arena_dev = nullptr;
    std::vector<Request *> pending; // submitted, rays left to hand to a step
    std::map<uint64_t, Request *> live; // submitted, not released
    std::vector<Request *> finished;    // completed, not yet polled
    std::vector<Deferred> deferred;     // released regions waiting for their copies
    std::mutex mu;
    uint64_t submitted = 0;
    int64_t completed_rays = 0;
};

namespace
{

// Hit arrays of one request inside its arena chunk: 8 float + 3 int32 SoA.
dmap_hits hits_in(char *base, int64_t n)
{
    dmap_hits h{};
    float *f = (float *)base;
    h.t = f;
    h.s = f + n;
    h.tt = f + 2 * n;
    h.bu = f + 3 * n;
    h.bv = f + 4 * n;
    h.nx = f + 5 * n;
    h.ny = f + 6 * n;
    h.nz = f + 7 * n;
    int32_t *i = (int32_t *)(f + 8 * n);
    h.prim = i;
    h.texel_i = i + n;
    h.texel_j = i + 2 * n;
    return h;
}

HitsDev hits_dev(const dmap_hits &h)
{
    return HitsDev{h.t, h.s, h.tt, h.bu, h.bv, h.nx, h.ny, h.nz, h.prim, h.texel_i, h.texel_j};
}

RaysDev rays_dev(const dmap_rays &r)
{
    return RaysDev{r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, r.tmin, r.tmax};
}

// The pool bytes of a request of n rays, as the arena rounds them.
size_t request_bytes(int64_t n) { return ((size_t)n * POOL_BYTES_PER_RAY + 255) & ~(size_t)255; }

// Returns released regions to the arena once the copies out of them have
// completed; a copy that failed leaves its region reclaimed as well (the
// error surfaces on the caller's stream).
void reclaim(dmap_server *S)
{
    auto &D = S->deferred;
    for (size_t k = 0; k < D.size();) {
        auto &ev = D[k].copies;
        ev.erase(std::remove_if(ev.begin(), ev.end(),
                                [](hipEvent_t e) {
                                    if (hipEventQuery(e) == hipErrorNotReady)
                                        return false;
                                    (void)hipEventDestroy(e);
                                    return true;
                                }),
                 ev.end());
        if (ev.empty()) {
            S->arena.release(D[k].off, D[k].bytes);
            D[k] = D.back();
            D.pop_back();
        } else {
            ++k;
        }
    }
}

// Collect finished slots; then fill every free slot with a new step.
int pump(dmap_server *S)
{
    reclaim(S);
    for (Slot *sl : S->slots) {
        if (!sl->busy)
            continue;
        hipError_t q = hipEventQuery(sl->done);
        if (q == hipErrorNotReady)
            continue;
        if (q != hipSuccess)
            return fail(DMAP_ERR_HIP, "serving step: %s", hipGetErrorString(q));
        const double t = now_s();
        for (Request *r : sl->reqs) {
            if (--r->inflight == 0 && r->next == r->n) {
                r->t_done = t;
                r->done = true;
                S->finished.push_back(r);
                S->completed_rays += r->n;
            }
        }
        sl->reqs.clear();
        sl->busy = false;
    }
    const double now = now_s();
    for (Slot *sl : S->slots) {
        if (sl->busy || S->pending.empty())
            continue;
        std::vector<Request *> order(S->pending);
        std::sort(order.begin(), order.end(), [now](const Request *a, const Request *b) {
            bool aa = now - a->t_submit > AGING_S, ab = now - b->t_submit > AGING_S;
            if (aa != ab)
                return aa;
            if (!aa) {
                int64_t ra = a->n - a->next, rb = b->n - b->next;
                if (ra != rb)
                    return ra < rb;
            }
            return a->seq < b->seq;
        });
        StepDesc &D = *sl->desc_host;
        const int kind = order.front()->obj->const_normal ? 1 : 0;
        int64_t total = 0;
        int n_seg = 0;
        for (Request *r : order) {
            if (n_seg == MAX_SEGMENTS || total >= S->step_rays)
                break;
            if ((r->obj->const_normal ? 1 : 0) != kind)
                continue;
            if (!r->admitted) {
                size_t bytes = request_bytes(r->n);
                size_t off;
                if (!S->arena.alloc(bytes, off)) {
                    if (now - r->t_submit > AGING_S)
                        break; // an aged request waits for room first
                    continue;
                }
                r->admitted = true;
                r->off = off;
                r->bytes = bytes;
                r->hits = hits_in(S->arena.base + off, r->n);
            }
            int64_t take = std::min(r->n - r->next, S->step_rays - total);
            Segment &g = D.seg[n_seg++];
            g.start = total;
            g.count = take;
            g.first = r->next;
            g.obj = r->obj->view_dev;
            g.rays = rays_dev(r->rays);
            g.hits = hits_dev(r->hits);
            r->next += take;
            ++r->inflight;
            total += take;
            sl->reqs.push_back(r);
        }
        if (n_seg == 0)
            break; // nothing fits the hit pool yet
        S->pending.erase(std::remove_if(S->pending.begin(), S->pending.end(),
                                        [](const Request *r) { return r->next == r->n; }),
                         S->pending.end());
        D.total = total;
        D.n_seg = n_seg;
        size_t desc_bytes = offsetof(StepDesc, seg) + sizeof(Segment) * n_seg;
        sl->busy = true;
        HIP_TRY(hipMemcpyAsync(sl->desc_dev, sl->desc_host, desc_bytes, hipMemcpyHostToDevice, sl->stream));
        HIP_TRY(hipGraphLaunch(sl->graph[kind], sl->stream));
        HIP_TRY(hipEventRecord(sl->done, sl->stream));
    }
    return DMAP_OK;
}

} // namespace

extern "C" {

int dmap_server_create(dmap_ctx *ctx, int n_slots, int64_t step_rays, int64_t hit_pool_bytes, dmap_server **out)
{
    if (!ctx || !out || n_slots < 1 || n_slots > 16 || step_rays < 1 || step_rays > MAX_STEP_RAYS ||
        hit_pool_bytes < 4096)
        return fail(DMAP_ERR_INVALID, "dmap_server_create: bad argument");
    HIP_TRY(hipSetDevice(ctx->device));
    dmap_server *S = new dmap_server;
    S->ctx = ctx;
    S->step_rays = step_rays;
    hipError_t e = hipMalloc(&S->arena_dev, (size_t)hit_pool_bytes);
    if (e != hipSuccess) {
        delete S;
        return fail(DMAP_ERR_MEMORY, "dmap_server_create: hit pool of %lld bytes: %s", (long long)hit_pool_bytes,
                    hipGetErrorString(e));
    }
    // The arena hands out whole 256 B units.
    S->arena.init(S->arena_dev, (size_t)hit_pool_bytes & ~(size_t)255);
    for (int k = 0; k < n_slots; ++k) {
        Slot *sl = new Slot;
        S->slots.push_back(sl);
        e = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking);
        if (e == hipSuccess)
            e = hipHostMalloc((void **)&sl->desc_host, sizeof(StepDesc), hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipMalloc(&sl->desc_dev, sizeof(StepDesc));
        if (e == hipSuccess)
            e = hipMalloc(&sl->queue, 256);
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&sl->done, hipEventDisableTiming);
        if (e != hipSuccess)
            break;
        // Capture the step's three launches once; they read the descriptor
        // from desc_dev, so a replay after a new copy runs the new step.
        TraceParams p;
        std::memset(&p, 0, sizeof p);
        p.step = sl->desc_dev;
        p.queue = sl->queue;
        p.prim_mask = -1;
        p.bound = 2;
        p.slab_max_level = 1;
        p.leaf = 0;
        p.polish = 1;
        for (int kind = 0; kind < 2 && e == hipSuccess; ++kind) {
            hipGraph_t g = nullptr;
            e = hipStreamBeginCapture(sl->stream, hipStreamCaptureModeThreadLocal);
            if (e == hipSuccess) {
                hipError_t le = launch_trace_multi(p, kind == 1, sl->stream);
                e = hipStreamEndCapture(sl->stream, &g);
                if (le != hipSuccess)
                    e = le;
            }
            if (e == hipSuccess)
                e = hipGraphInstantiate(&sl->graph[kind], g, nullptr, nullptr, 0);
            if (g)
                (void)hipGraphDestroy(g);
        }
        if (e != hipSuccess)
            break;
    }
    if (e != hipSuccess) {
        dmap_server_destroy(S);
        return fail(DMAP_ERR_HIP, "dmap_server_create: %s", hipGetErrorString(e));
    }
    *out = S;
    return DMAP_OK;
}

int dmap_server_submit(dmap_server *S, const dmap_object *obj, const dmap_rays *rays, int64_t n, uint64_t tag)
{
    if (!S || !obj || !rays || n <= 0)
        return fail(DMAP_ERR_INVALID, "dmap_server_submit: bad argument");
    if (request_bytes(n) > S->arena.bytes)
        return fail(DMAP_ERR_INVALID, "dmap_server_submit: %lld rays need more than the whole hit pool",
                    (long long)n);
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->live.count(tag))
        return fail(DMAP_ERR_INVALID, "dmap_server_submit: tag %llu is in use", (unsigned long long)tag);
    Request *r = new Request;
    r->tag = tag;
    r->seq = S->submitted;
    r->obj = obj;
    r->rays = *rays;
    r->n = n;
    r->t_submit = now_s();
    S->live[tag] = r;
    S->pending.push_back(r);
    ++S->submitted;
    HIP_TRY(hipSetDevice(S->ctx->device));
    return pump(S);
}

int dmap_server_poll(dmap_server *S, int max, uint64_t *tags, double *latency_s, int *n_done)
{
    if (!S || !n_done || max < 0)
        return fail(DMAP_ERR_INVALID, "dmap_server_poll: bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    HIP_TRY(hipSetDevice(S->ctx->device));
    int rc = pump(S);
    if (rc != DMAP_OK)
        return rc;
    int k = 0;
    while (k < max && !S->finished.empty()) {
        Request *r = S->finished.back();
        S->finished.pop_back();
        if (tags)
            tags[k] = r->tag;
        if (latency_s)
            latency_s[k] = r->t_done - r->t_submit;
        ++k;
    }
    *n_done = k;
    return DMAP_OK;
}

int dmap_server_hits(dmap_server *S, uint64_t tag, dmap_hits *hits)
{
    if (!S || !hits)
        return fail(DMAP_ERR_INVALID, "dmap_server_hits: bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->live.find(tag);
    if (it == S->live.end() || !it->second->done)
        return fail(DMAP_ERR_INVALID, "dmap_server_hits: no completed request %llu", (unsigned long long)tag);
    *hits = it->second->hits;
    return DMAP_OK;
}

int dmap_server_copy_hits(dmap_server *S, uint64_t tag, const dmap_hits *dst, void *stream)
{
    if (!S || !dst)
        return fail(DMAP_ERR_INVALID, "dmap_server_copy_hits: bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->live.find(tag);
    if (it == S->live.end() || !it->second->done)
        return fail(DMAP_ERR_INVALID, "dmap_server_copy_hits: no completed request %llu", (unsigned long long)tag);
    const Request *r = it->second;
    HIP_TRY(hipSetDevice(S->ctx->device));
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)r->n;
    void *const to[11] = {dst->t, dst->s, dst->tt, dst->bu, dst->bv, dst->nx, dst->ny, dst->nz,
                          dst->prim, dst->texel_i, dst->texel_j};
    const void *const from[11] = {r->hits.t, r->hits.s, r->hits.tt, r->hits.bu, r->hits.bv, r->hits.nx,
                                  r->hits.ny, r->hits.nz, r->hits.prim, r->hits.texel_i, r->hits.texel_j};
    for (int q = 0; q < 11; ++q)
        if (to[q])
            HIP_TRY(hipMemcpyAsync(to[q], from[q], 4 * n, hipMemcpyDeviceToDevice, st));
    // The region must outlive the copies: release defers its reuse until
    // this event has completed.
    hipEvent_t ev = nullptr;
    HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    hipError_t e = hipEventRecord(ev, st);
    if (e != hipSuccess) {
        (void)hipEventDestroy(ev);
        return fail(DMAP_ERR_HIP, "dmap_server_copy_hits: %s", hipGetErrorString(e));
    }
    it->second->copies.push_back(ev);
    return DMAP_OK;
}

int dmap_server_release(dmap_server *S, uint64_t tag)
{
    if (!S)
        return fail(DMAP_ERR_INVALID, "dmap_server_release: NULL server");
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->live.find(tag);
    if (it == S->live.end() || !it->second->done)
        return fail(DMAP_ERR_INVALID, "dmap_server_release: no completed request %llu", (unsigned long long)tag);
    Request *r = it->second;
    HIP_TRY(hipSetDevice(S->ctx->device));
    if (r->admitted) {
        if (r->copies.empty())
            S->arena.release(r->off, r->bytes);
        else
            S->deferred.push_back(Deferred{r->off, r->bytes, std::move(r->copies)});
    }
    r->copies.clear();
    S->live.erase(it);
    delete r;
    return pump(S);
}

int dmap_server_idle(dmap_server *S, int *idle)
{
    if (!S || !idle)
        return fail(DMAP_ERR_INVALID, "dmap_server_idle: bad argument");
    std::lock_guard<std::mutex> lk(S->mu);
    HIP_TRY(hipSetDevice(S->ctx->device));
    int rc = pump(S);
    bool any = !S->pending.empty();
    for (Slot *sl : S->slots)
        any = any || sl->busy;
    *idle = any ? 0 : 1;
    return rc;
}

void dmap_server_destroy(dmap_server *S)
{
    if (!S)
        return;
    (void)hipSetDevice(S->ctx->device);
    for (Slot *sl : S->slots) {
        if (sl->stream)
            (void)hipStreamSynchronize(sl->stream);
        for (hipGraphExec_t g : sl->graph)
            if (g)
                (void)hipGraphExecDestroy(g);
        if (sl->done)
            (void)hipEventDestroy(sl->done);
        if (sl->stream)
            (void)hipStreamDestroy(sl->stream);
        (void)hipHostFree(sl->desc_host);
        (void)hipFree(sl->desc_dev);
        (void)hipFree(sl->queue);
        delete sl;
    }
    for (auto &kv : S->live) {
        for (hipEvent_t e : kv.second->copies) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
        delete kv.second;
    }
    for (Deferred &d : S->deferred)
        for (hipEvent_t e : d.copies) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
    (void)hipFree(S->arena_dev);
    delete S;
}

} // extern "C"
