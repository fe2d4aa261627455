"""dmap_intersect_closest_noalloc through the C ABI: the same hits (and work
counters) as dmap_intersect_closest on the same rays, bit for bit, on calls
in a row of decreasing size (the largest it takes, a smaller one, one below
the reorder threshold), and the size limit refused with an error."""
import ctypes

import numpy as np
import pytest

from dmap_amd import workload as W

pytestmark = pytest.mark.gpu

MAX_RAYS = 1 << 20


@pytest.fixture(scope="module")
def scene(torch_gpu):
    import dmap_amd
    import torch
    print("[no_alloc] C2 structure at 512^2, 2^20 + 7 rays", flush=True)
    wl = W.c2_incoherent(n=512, width=1024, height=512, n_bounce=(1 << 19) + 7)
    tile = dmap_amd.DisplacementTile(wl.texels, wl.strength, wl.offset, wl.repeat)
    obj = dmap_amd.ObjectIntersector(tile, wl.triangles)
    rays = torch.from_numpy(np.ascontiguousarray(wl.rays)).cuda()
    assert rays.shape[1] == MAX_RAYS + 7
    return dmap_amd, torch, obj, rays


def _call(dmap_amd, torch, obj, rays, entry, stats=False, stream=None):
    """One closest-hit call through the C ABI; returns (floats (8, n), ints (3, n), stats (6, n) or None)."""
    n = rays.shape[1]
    r, _ = obj._rays(rays)
    F = torch.full((8, n), -7.0, dtype=torch.float32, device="cuda")
    I = torch.full((3, n), -7, dtype=torch.int32, device="cuda")
    h = dmap_amd._Hits(*[F[q].data_ptr() for q in range(8)], *[I[q].data_ptr() for q in range(3)])
    S = None
    st = None
    if stats:
        S = torch.zeros((len(dmap_amd.STAT_NAMES) + dmap_amd.STAT_LEVELS, n), dtype=torch.int32, device="cuda")
        st = dmap_amd._Stats(*[S[q].data_ptr() for q in range(len(dmap_amd.STAT_NAMES) + 1)])
    # With the ray sort, whose buffers are most of the scratch.
    opt = dmap_amd.TraversalOptions(reorder=True)._c()
    fn = getattr(dmap_amd._lib, entry)
    rc = fn(obj._h, ctypes.byref(r), n, ctypes.byref(h), -1, ctypes.byref(opt),
            ctypes.byref(st) if st is not None else None, dmap_amd._stream_ptr(stream))
    if rc != 0:
        raise dmap_amd.DmapError(dmap_amd._lib.dmap_last_error().decode())
    return F, I, S


def test_noalloc_equals_closest_on_calls_in_a_row(scene):
    dmap_amd, torch, obj, rays = scene
    # Largest, smaller, and below the reorder threshold (no sort), one after another on one stream.
    sizes = [(0, MAX_RAYS), (123457, 300000), (450007, 4000)]
    got = []
    for k, (a, n) in enumerate(sizes):
        sub = rays[:, a:a + n].contiguous()
        got.append((sub, _call(dmap_amd, torch, obj, sub, "dmap_intersect_closest_noalloc", stats=(k == 1))))
    torch.cuda.synchronize()
    for k, (sub, (F, I, S)) in enumerate(got):
        F0, I0, S0 = _call(dmap_amd, torch, obj, sub, "dmap_intersect_closest", stats=(k == 1))
        torch.cuda.synchronize()
        assert torch.equal(F.view(torch.int32), F0.view(torch.int32)), f"call {k}: float fields differ"
        assert torch.equal(I, I0), f"call {k}: int fields differ"
        if S is not None:
            assert torch.equal(S, S0), f"call {k}: work counters differ"
        hits = int(torch.isfinite(F[0]).sum())
        print(f"[no_alloc] call {k}: {sub.shape[1]} rays, {hits} hits, identical", flush=True)
        assert hits > 1000


def test_noalloc_refuses_more_than_its_scratch(scene):
    dmap_amd, torch, obj, rays = scene
    with pytest.raises(dmap_amd.DmapError, match="at most 1048576"):
        _call(dmap_amd, torch, obj, rays, "dmap_intersect_closest_noalloc")
    # The refusal left the next call intact.
    sub = rays[:, :5000].contiguous()
    F, I, _ = _call(dmap_amd, torch, obj, sub, "dmap_intersect_closest_noalloc")
    F0, I0, _ = _call(dmap_amd, torch, obj, sub, "dmap_intersect_closest")
    torch.cuda.synchronize()
    assert torch.equal(F.view(torch.int32), F0.view(torch.int32)) and torch.equal(I, I0)
