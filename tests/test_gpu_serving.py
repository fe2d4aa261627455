"""The serving loop: every request's hits are the same alone and in the mix,
and the same as the single-request entry point's; chunked requests, steps
replayed from the captured graphs many times with eager work between them,
memory-based admission, and objects of both normal kinds in one server."""
import numpy as np
import pytest

from dmap_amd import serving
from dmap_amd import workload as W
from tests import scenes as S

pytestmark = pytest.mark.gpu

FIELDS = ("t", "st", "bary", "ng", "prim", "texel")


def _host(h):
    return {k: h[k].cpu().numpy() for k in FIELDS}


def _same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in FIELDS)


def _run(server, jobs, release=True):
    """Submit every (obj, rays) at once, wait for all; returns host hits by index."""
    import torch
    for q, (obj, rays) in enumerate(jobs):
        server.submit(obj, rays, q)
    out, lat = {}, {}
    while len(out) < len(jobs):
        for tag, l in server.poll():
            out[tag] = _host(server.hits(tag, jobs[tag][1].shape[1]))
            lat[tag] = l
            if release:
                server.release(tag)
    torch.cuda.synchronize()
    assert server.idle()
    assert all(l >= 0 for l in lat.values())
    return out


@pytest.fixture(scope="module")
def small_scenes(torch_gpu):
    print("[serving] building 3 scenes (128^2, 256^2, 512^2, pools of 16384 rays)", flush=True)
    return serving.make_scenes((128, 256, 512), pool=1 << 14)


def test_mix_equals_alone_and_plain(small_scenes):
    import dmap_amd
    scn = small_scenes
    reqs = serving.make_requests(scn, 40, seed=11, min_rays=100, max_rays=12000)
    jobs = [(scn[k].obj, r) for k, r in reqs]
    # A small step budget splits the larger requests over several steps.
    srv = dmap_amd.Server(slots=3, step_rays=8192, hit_pool_bytes=64 << 20)
    mix = _run(srv, jobs)
    for q, (obj, rays) in enumerate(jobs):
        alone = _run(srv, [(obj, rays)])[0]
        assert _same(mix[q], alone), f"request {q}: mix and alone differ"
        plain = _host(obj.intersect(rays))
        assert _same(mix[q], plain), f"request {q}: server and single-request entry point differ"
        if q % 8 == 0:
            # eager work on the default stream between graph replays
            obj.intersect(rays, options=dmap_amd.TraversalOptions(reorder=False))
    hit = np.concatenate([np.isfinite(m["t"]) for m in mix.values()])
    assert 0.2 < hit.mean() < 1.0


def test_backlog_steps_equal_plain(small_scenes):
    """60 requests of 2,000-12,000 rays on 2 slots of 4,096-ray steps: the
    waiting rays (over four times slots x 4 x 4,096, the backlog
    threshold) make the first steps hold up to 16,384 rays, and the later ones,
    once the backlog has drained, 4,096; every request's hits equal the
    single-request entry point's bit for bit."""
    import dmap_amd
    scn = small_scenes
    reqs = serving.make_requests(scn, 60, seed=31, min_rays=2000, max_rays=12000)
    jobs = [(scn[k].obj, r) for k, r in reqs]
    assert sum(r.shape[1] for _, r in jobs) > 2 * 4 * 4096 * 4
    srv = dmap_amd.Server(slots=2, step_rays=4096, hit_pool_bytes=64 << 20)
    mix = _run(srv, jobs)
    for q, (obj, rays) in enumerate(jobs):
        assert _same(mix[q], _host(obj.intersect(rays))), f"request {q}"


def test_memory_admission(small_scenes):
    """A hit pool that holds ~3 requests at once: the rest wait for releases."""
    import dmap_amd
    scn = small_scenes
    reqs = serving.make_requests(scn, 12, seed=5, min_rays=4000, max_rays=9000)
    jobs = [(scn[k].obj, r) for k, r in reqs]
    srv = dmap_amd.Server(slots=2, step_rays=1 << 20, hit_pool_bytes=44 * 30000)
    mix = _run(srv, jobs)
    for q, (obj, rays) in enumerate(jobs):
        assert _same(mix[q], _host(obj.intersect(rays))), f"request {q}"
    import torch
    big = torch.cat([scn[0].primary, scn[0].bounce], dim=1).contiguous()  # 32768 rays: more than the pool holds
    with pytest.raises(dmap_amd.DmapError):
        srv.submit(scn[0].obj, big, 999)


def test_unreleased_hits_block_admission(small_scenes):
    """Hits stay valid until released; admission waits for the pool."""
    import torch

    import dmap_amd
    scn = small_scenes
    rays = [scn[0].primary[:, 64 * k:64 * k + 5000].contiguous() for k in range(3)]
    srv = dmap_amd.Server(slots=2, step_rays=1 << 20, hit_pool_bytes=44 * 11000)
    srv.submit(scn[0].obj, rays[0], 0)
    srv.submit(scn[0].obj, rays[1], 1)
    srv.submit(scn[0].obj, rays[2], 2)
    done = set()
    while len(done) < 2:
        done |= {t for t, _ in srv.poll()}
    torch.cuda.synchronize()
    assert done == {0, 1}
    assert not srv.idle()  # the third waits for room
    h0 = _host(srv.hits(0, 5000))
    srv.release(0)
    while 2 not in done:
        done |= {t for t, _ in srv.poll()}
    assert _same(h0, _host(scn[0].obj.intersect(rays[0])))
    assert _same(_host(srv.hits(2, 5000)), _host(scn[0].obj.intersect(rays[2])))
    with pytest.raises(dmap_amd.DmapError):
        srv.submit(scn[0].obj, rays[0], 1)  # tag 1 is still held
    srv.release(1)
    srv.release(2)
    with pytest.raises(dmap_amd.DmapError):
        srv.release(2)


def test_both_normal_kinds(small_scenes, torch_gpu):
    """Interpolated-normal and constant-normal objects share one server."""
    import dmap_amd
    torch = torch_gpu
    tex = np.random.default_rng(3).random((64, 64)).astype(np.float32)
    tile = dmap_amd.DisplacementTile(tex, 0.05)
    curved = dmap_amd.ObjectIntersector(tile, S.curved_quad())
    o, d = W.pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 60.0, 64, 64)
    rc = torch.from_numpy(W.pack_rays(o, d)).cuda()
    scn = small_scenes
    jobs = []
    for q in range(10):
        if q % 2:
            jobs.append((curved, rc[:, 256 * q:256 * q + 1500].contiguous()))
        else:
            jobs.append((scn[q % 3].obj, scn[q % 3].bounce[:, 1000 * q:1000 * q + 3000].contiguous()))
    srv = dmap_amd.Server(slots=2, step_rays=4096, hit_pool_bytes=16 << 20)
    mix = _run(srv, jobs)
    for q, (obj, rays) in enumerate(jobs):
        assert _same(mix[q], _host(obj.intersect(rays))), f"request {q}"


def test_full_steps_of_tiny_requests(small_scenes):
    """300 requests of 8-40 rays in 1M-ray steps: a step fills up to its
    32-request limit (the step kernels stage the request starts in LDS and
    search them per slot), and every request's hits equal the
    single-request entry point's."""
    import dmap_amd
    scn = small_scenes
    reqs = serving.make_requests(scn, 300, seed=23, min_rays=8, max_rays=40)
    jobs = [(scn[k].obj, r) for k, r in reqs]
    srv = dmap_amd.Server(slots=1, step_rays=1 << 20, hit_pool_bytes=64 << 20)
    mix = _run(srv, jobs)
    for q, (obj, rays) in enumerate(jobs):
        assert _same(mix[q], _host(obj.intersect(rays))), f"request {q} differs"


def test_release_right_after_copy_keeps_the_copy(small_scenes):
    """dmap_server_copy_hits, then dmap_server_release, then a new request
    that can only use the released pool region, all before the copy's
    stream has run: the copy still reads the first request's hits, because
    the region returns to the pool only when the copy has completed."""
    import dmap_amd
    import torch
    scn = small_scenes[0]
    n = 8192
    rays_a = scn.primary[:, :n].contiguous()
    rays_b = scn.bounce[:, :n].contiguous()
    plain_a, plain_b = _host(scn.obj.intersect(rays_a)), _host(scn.obj.intersect(rays_b))
    assert not _same(plain_a, plain_b)
    # The pool holds one request of n rays (44 B per ray, in 256 B units).
    pool = ((n * 44 + 255) // 256) * 256 + 256
    srv = dmap_amd.Server(slots=1, step_rays=1 << 20, hit_pool_bytes=pool)
    srv.submit(scn.obj, rays_a, 0)
    while not srv.poll():
        pass
    side = torch.cuda.Stream()
    x = torch.randn((4096, 4096), device="cuda")
    with torch.cuda.stream(side):
        y = x
        for _ in range(12):  # keeps `side` busy for several milliseconds
            y = (y @ x) * 1e-3
    got_a = srv.hits(0, n, stream=side)
    srv.release(0)
    srv.submit(scn.obj, rays_b, 1)
    while not srv.poll():
        pass
    got_b = _host(srv.hits(1, n))
    side.synchronize()
    got_a = {k: v.cpu().numpy() for k, v in got_a.items()}
    assert _same(got_a, plain_a)
    assert _same(got_b, plain_b)
    srv.release(1)
    torch.cuda.synchronize()
    assert srv.idle()


def test_pool_not_a_multiple_of_256(small_scenes):
    """A hit pool of any size admits every request it accepts: sizes are
    compared in the arena's 256 B units, so no accepted request waits for
    room that can never exist."""
    import dmap_amd
    scn = small_scenes[0]
    n = 1000  # 44,000 B, 44,032 B in 256 B units
    # 44,020 B holds 44,000 B but only 171 whole units (43,776 B): refused at once.
    srv = dmap_amd.Server(slots=1, step_rays=1 << 20, hit_pool_bytes=44020)
    with pytest.raises(dmap_amd.DmapError, match="whole hit pool"):
        srv.submit(scn.obj, scn.primary[:, :n].contiguous(), 0)
    srv2 = dmap_amd.Server(slots=1, step_rays=1 << 20, hit_pool_bytes=44032 + 100)
    rays = scn.primary[:, :n].contiguous()
    srv2.submit(scn.obj, rays, 0)
    while not srv2.poll():
        pass
    assert _same(_host(srv2.hits(0, n)), _host(scn.obj.intersect(rays)))
    srv2.release(0)


def test_scattered_ray_planes_and_the_segment_cap(small_scenes):
    """Requests whose eight ray arrays are separate allocations (not one
    stride apart) go through the server's device record of their planes;
    and 120 tiny requests fill steps up to their 32-request cap. Every
    request's hits equal the single-request call's."""
    import ctypes

    import dmap_amd
    import torch
    scn = small_scenes[1]
    srv = dmap_amd.Server(slots=3, step_rays=1 << 20, hit_pool_bytes=64 << 20)
    n = 3000
    rays = scn.primary[:, 4096:4096 + n].contiguous()
    planes = [rays[q].clone() for q in range(8)]  # eight separate buffers
    for q in range(8):
        planes[q] = torch.cat([torch.zeros(q * 17 + 1, device="cuda"), planes[q]])[q * 17 + 1:]
    r = dmap_amd._Rays(*[planes[q].data_ptr() for q in range(8)])
    dmap_amd._check(dmap_amd._lib.dmap_server_submit(srv._h, scn.obj._h, ctypes.byref(r), n, ctypes.c_uint64(0)),
                    "dmap_server_submit")
    tiny = [scn.bounce[:, 100 * k:100 * k + 100].contiguous() for k in range(120)]
    for k, t in enumerate(tiny):
        srv.submit(scn.obj, t, 1 + k)
    got = {}
    while len(got) < 121:
        for tag, _ in srv.poll():
            got[tag] = _host(srv.hits(tag, n if tag == 0 else 100))
            srv.release(tag)
    torch.cuda.synchronize()
    assert _same(got[0], _host(scn.obj.intersect(rays)))
    for k, t in enumerate(tiny):
        assert _same(got[1 + k], _host(scn.obj.intersect(t))), k
    assert sum(int(np.isfinite(got[1 + k]["t"]).sum()) for k in range(120)) > 0


def test_more_objects_than_a_step_holds(small_scenes):
    """44 requests against 11 objects, more than the 8 object groups a step
    holds: quad tilings of 1 to 12 quads a side over the three tiles (the
    12 x 12 tiling has 288 triangles, a BVH too large for LDS), requests of
    500-5,000 rays, 16,384-ray steps on 2 slots. Steps hold several objects,
    each traced by its own share of the workgroups; every request's hits
    equal the single-request entry point's."""
    import dmap_amd
    scn = small_scenes
    sides = (1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12)
    objs = [dmap_amd.ObjectIntersector(scn[k % 3].tile, W.quad_plane(m, 8.0)) for k, m in enumerate(sides)]
    rng = np.random.default_rng(41)
    jobs = []
    for q in range(44):
        k = q % len(objs)
        n = int(rng.integers(500, 5001))
        pool = scn[k % 3].primary if q % 2 else scn[k % 3].bounce
        start = int(rng.integers(0, pool.shape[1] - n))
        jobs.append((objs[k], pool[:, start:start + n].contiguous()))
    srv = dmap_amd.Server(slots=2, step_rays=16384, hit_pool_bytes=64 << 20)
    mix = _run(srv, jobs)
    for q, (obj, rays) in enumerate(jobs):
        assert _same(mix[q], _host(obj.intersect(rays))), f"request {q}"
    assert np.mean([np.isfinite(m["t"]).mean() for m in mix.values()]) > 0.2
