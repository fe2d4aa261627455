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
//  * backlog steps: while the rays waiting would fill every slot's next
//    BACKLOG_STEPS steps of BACKLOG_FACTOR x step_rays, a step holds up to
//    BACKLOG_FACTOR x step_rays. Each step's start, drain and finalize cost
//    about the same at any size, so fewer, larger steps raise the rate under
//    a backlog; at a light load the steps stay at step_rays, and a small
//    request that arrives finds the slots running steps that short;
//  * one kind per step: objects with constant normals and objects with
//    interpolated normals use different kernel specialisations, so a step
//    takes requests of the kind of its first request only;
//  * object groups: a step holds requests of up to MAX_GROUPS objects, its
//    slots grouped by object, and the step kernels' workgroups are shared
//    out among the groups in proportion to their rays. A workgroup serves
//    one object only, reading it from the kernel arguments (uniform across
//    its waves, in scalar registers) with the object's BVH in LDS, as a
//    single-object launch does. (With the objects mixed inside waves each
//    lane read its object's fields through its own pointer and the BVH from
//    global memory: the step kernel was 24% slower per ray than the
//    single-object launch on the same rays. Steps of one object each were
//    as fast per ray but needed a step per object, and small requests of
//    different objects waited for steps in turn.)
// A step is one persistent multi-request launch (reset, trace, finalize)
// built once per slot and kind as a HIP graph and replayed with the step's
// descriptor in the kernels' arguments (no upload kernel before it), so
// requests that arrive while others run are admitted into the next free
// slot and overlap on the device with the steps already running. (A second
// step queued behind the running one on each slot's stream left the
// saturated rate unchanged and doubled the median latency at a fixed load,
// so each slot runs one step at a time; STEPS_PER_SLOT keeps the choice
// open.) Each step records an event; the server's calls
// (submit, poll, release, idle) query the events, and a request's latency
// is the time its last step was seen complete minus its submission, so a
// caller that polls continuously measures it to within one poll.
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
constexpr int64_t BACKLOG_FACTOR = 4;
constexpr int64_t BACKLOG_STEPS = 1;
// Hit pool bytes per ray: 8 fp32 + 3 int32 hit fields.
constexpr size_t POOL_BYTES_PER_RAY = 44;

struct Request
{
    uint64_t tag = 0;
    uint64_t seq = 0; // arrival order
    const dmap_object *obj = nullptr;
    dmap_rays rays{};
    // The ray planes as a step segment takes them: base + q * stride, or
    // (stride 0) a RaysDev record in the server's device table, entry `table`.
    const float *rays_base = nullptr;
    int64_t rays_stride = 0;
    int table = -1;
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

// One step in flight on a slot's stream.
struct StepBuf
{
    unsigned *queue = nullptr; // the step's ray queue head (device)
    StepGraph graph[2];        // by kind: interpolated, constant normals
    hipEvent_t done = nullptr; // recorded after the step's launches
    std::vector<Request *> reqs;
    bool busy = false;
};

// Steps queued on one stream: at most STEPS_PER_SLOT.
constexpr int STEPS_PER_SLOT = 1;
struct Slot
{
    hipStream_t stream = nullptr;
    StepBuf step[STEPS_PER_SLOT];
};

// RaysDev records of requests whose ray planes do not lie at one stride.
constexpr int RAY_TABLE = 1024;

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
    RaysDev *ray_table = nullptr;       // device, RAY_TABLE records
    std::vector<int> table_free;
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

// Collects finished steps; then forms new steps for free step buffers: first
// on streams with nothing in flight, then (STEPS_PER_SLOT > 1) queued behind
// a running step.
int pump(dmap_server *S)
{
    reclaim(S);
    for (Slot *sl : S->slots) {
        for (StepBuf &b : sl->step) {
            if (!b.busy)
                continue;
            hipError_t q = hipEventQuery(b.done);
            if (q == hipErrorNotReady)
                continue;
            if (q != hipSuccess)
                return fail(DMAP_ERR_HIP, "serving step: %s", hipGetErrorString(q));
            const double t = now_s();
            for (Request *r : b.reqs) {
                if (--r->inflight == 0 && r->next == r->n) {
                    r->t_done = t;
                    r->done = true;
                    S->finished.push_back(r);
                    S->completed_rays += r->n;
                }
            }
            b.reqs.clear();
            b.busy = false;
        }
    }
    for (int depth = 0; depth < STEPS_PER_SLOT; ++depth) {
        for (Slot *sl : S->slots) {
            if (S->pending.empty())
                return DMAP_OK;
            int inflight = 0;
            StepBuf *b = nullptr;
            for (StepBuf &c : sl->step) {
                if (c.busy)
                    ++inflight;
                else if (!b)
                    b = &c;
            }
            if (inflight != depth || !b)
                continue;
            const double now = now_s();
            std::vector<Request *> order(S->pending);
            std::sort(order.begin(), order.end(), [now](const Request *a, const Request *c) {
                bool aa = now - a->t_submit > AGING_S, ac = now - c->t_submit > AGING_S;
                if (aa != ac)
                    return aa;
                if (!aa) {
                    int64_t ra = a->n - a->next, rc = c->n - c->next;
                    if (ra != rc)
                        return ra < rc;
                }
                return a->seq < c->seq;
            });
            StepDesc D;
            const int kind = order.front()->obj->const_normal ? 1 : 0;
            int64_t waiting = 0;
            for (const Request *r : order)
                waiting += r->n - r->next;
            const int64_t big = BACKLOG_FACTOR * S->step_rays;
            const int64_t cap =
                waiting >= BACKLOG_STEPS * (int64_t)S->slots.size() * big && big <= MAX_STEP_RAYS ? big : S->step_rays;
            // The step's requests in order, each with its object's group.
            struct Pick
            {
                Request *r;
                int grp;
                int64_t first, take;
            };
            std::vector<Pick> picks;
            const dmap_object *objs[MAX_GROUPS];
            int n_grp = 0;
            int64_t total = 0;
            for (Request *r : order) {
                if ((int)picks.size() == MAX_SEGMENTS || total >= cap)
                    break;
                if ((r->obj->const_normal ? 1 : 0) != kind)
                    continue;
                int grp = 0;
                while (grp < n_grp && objs[grp] != r->obj)
                    ++grp;
                if (grp == MAX_GROUPS)
                    continue; // a ninth object waits for a later step
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
                if (grp == n_grp)
                    objs[n_grp++] = r->obj;
                int64_t take = std::min(r->n - r->next, cap - total);
                picks.push_back(Pick{r, grp, r->next, take});
                r->next += take;
                ++r->inflight;
                total += take;
                b->reqs.push_back(r);
            }
            if (picks.empty())
                return DMAP_OK; // nothing fits the hit pool yet
            S->pending.erase(std::remove_if(S->pending.begin(), S->pending.end(),
                                            [](const Request *r) { return r->next == r->n; }),
                             S->pending.end());
            // The step slots group by object: group g's requests fill
            // [grp[g].start, grp[g].end), in the order they were picked.
            int n_seg = 0;
            int64_t at = 0;
            for (int g = 0; g < n_grp; ++g) {
                StepGroup &G = D.grp[g];
                G.obj = objs[g]->view;
                G.start = at;
                G.seg_lo = n_seg;
                for (const Pick &k : picks) {
                    if (k.grp != g)
                        continue;
                    Segment &sg = D.seg[n_seg++];
                    sg.start = at;
                    sg.first = k.first;
                    sg.rays = k.r->rays_base;
                    sg.rays_stride = k.r->rays_stride;
                    sg.hits = k.r->hits.t;
                    sg.hits_n = k.r->n;
                    at += k.take;
                }
                G.end = at;
                G.seg_hi = n_seg;
            }
            D.total = total;
            D.n_seg = n_seg;
            D.n_grp = n_grp;
            step_partition(D);
            TraceParams p;
            std::memset(&p, 0, sizeof p);
            p.queue = b->queue;
            p.prim_mask = -1;
            p.bound = 2;
            p.slab_max_level = 1;
            p.leaf = 0;
            p.polish = 1;
            b->busy = true;
            HIP_TRY(step_graph_launch(b->graph[kind], p, D, sl->stream));
            HIP_TRY(hipEventRecord(b->done, sl->stream));
        }
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
    e = hipMalloc(&S->ray_table, sizeof(RaysDev) * RAY_TABLE);
    for (int k = RAY_TABLE - 1; k >= 0; --k)
        S->table_free.push_back(k);
    for (int k = 0; k < n_slots && e == hipSuccess; ++k) {
        Slot *sl = new Slot;
        S->slots.push_back(sl);
        e = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking);
        for (StepBuf &b : sl->step) {
            if (e == hipSuccess)
                e = hipMalloc(&b.queue, 256);
            if (e == hipSuccess)
                e = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
            for (int kind = 0; kind < 2 && e == hipSuccess; ++kind)
                e = step_graph_create(b.queue, kind == 1, b.graph[kind]);
        }
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
    const float *planes[8] = {rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin, rays->tmax};
    for (const float *pl : planes)
        if (!pl)
            return fail(DMAP_ERR_INVALID, "dmap_server_submit: a ray array is NULL");
    HIP_TRY(hipSetDevice(S->ctx->device));
    // The step segments take the ray planes as base + q * stride; planes
    // elsewhere go through a RaysDev record in the device table.
    const int64_t stride = planes[1] - planes[0];
    bool uniform = stride != 0;
    for (int q = 2; q < 8 && uniform; ++q)
        uniform = planes[q] - planes[0] == q * stride;
    int table = -1;
    if (!uniform) {
        if (S->table_free.empty())
            return fail(DMAP_ERR_INVALID, "dmap_server_submit: more than %d live requests whose ray arrays do not "
                        "lie at one stride", RAY_TABLE);
        table = S->table_free.back();
        RaysDev rec{planes[0], planes[1], planes[2], planes[3], planes[4], planes[5], planes[6], planes[7]};
        HIP_TRY(hipMemcpy(S->ray_table + table, &rec, sizeof rec, hipMemcpyHostToDevice));
        S->table_free.pop_back();
    }
    Request *r = new Request;
    r->tag = tag;
    r->seq = S->submitted;
    r->obj = obj;
    r->rays = *rays;
    r->rays_base = uniform ? planes[0] : (const float *)(S->ray_table + table);
    r->rays_stride = uniform ? stride : 0;
    r->table = table;
    r->n = n;
    r->t_submit = now_s();
    S->live[tag] = r;
    S->pending.push_back(r);
    ++S->submitted;
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
    if (r->table >= 0)
        S->table_free.push_back(r->table);
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
        for (const StepBuf &b : sl->step)
            any = any || b.busy;
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
        for (StepBuf &b : sl->step) {
            for (StepGraph &g : b.graph)
                step_graph_destroy(g);
            if (b.done)
                (void)hipEventDestroy(b.done);
            (void)hipFree(b.queue);
        }
        if (sl->stream)
            (void)hipStreamDestroy(sl->stream);
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
    (void)hipFree(S->ray_table);
    (void)hipFree(S->arena_dev);
    delete S;
}

} // extern "C"
