"""The wavefront path tracer (config 3's estimator), in two forms: the torch
loop over the batched GPU intersector and the device wavefront
(dmap_pathtrace). Parity with the oracle ray by ray: the torch loop runs
once following the GPU's hits and once the oracle's, every batch traced by
both; no ray differs outside the parity bands, and every pixel whose value
differs has such a banded ray on its paths (shared counter-based random
numbers: paths part only where a hit differs). The two GPU forms trace the
same rays (equal ray counts, pixels equal up to summation order); NEE with
MIS and BSDF-only sampling agree in expectation; a white furnace (albedo 1,
environment 1, no light) returns at most 1 everywhere and about 1 on
average."""
import numpy as np
import pytest

import oracle
from dmap_amd import pathtrace as PT
from dmap_amd import workload as W
from tests.pt_oracle import assert_explained, oracle_parity

pytestmark = pytest.mark.gpu


def _small(torch, n_tex=128, res=48):
    import dmap_amd
    tex = W.c2_map(n_tex)
    tile = dmap_amd.DisplacementTile(tex, 0.1, 0.0, True, "fold")
    obj = dmap_amd.ObjectIntersector(tile, W.quad_plane(8, 8.0))
    o, d = W.pinhole_rays((4.0, -0.5, 3.5), (4.0, 4.0, 0.0), 60.0, res, res)
    cam_o = torch.from_numpy(o.T.astype(np.float64))
    cam_d = torch.from_numpy(d.T.astype(np.float64))
    sc = oracle.Scene(W.close_tile(tex).astype(np.float64), 0.1, 0.0, True, W.quad_plane(8, 8.0))
    return tile, obj, sc, cam_o, cam_d


def test_pathtrace_gpu_matches_oracle_paths(torch_gpu):
    torch = torch_gpu
    tile, obj, sc, cam_o, cam_d = _small(torch)
    tris = W.quad_plane(8, 8.0)
    st = {}
    plain = PT.render(torch, PT.Scene(PT.GpuIntersector(obj)), cam_o.cuda(), cam_d.cuda(), spp=2, depth=4, seed=7,
                      stats=st).cpu().numpy()
    assert st["closest_rays"] > 3 * cam_o.shape[1] and st["shadow_rays"] > cam_o.shape[1]
    img_g, img_o, flagged, unexplained, rays = oracle_parity(torch, obj, sc, tris, 128, cam_o, cam_d, 2, 4, 7)
    # the checking intersector answers exactly as the plain one
    assert np.array_equal(img_g, plain)
    assert unexplained == 0 and rays > 5 * cam_o.shape[1]
    assert_explained(img_g, img_o, flagged, "torch loop")


def test_pathtrace_mis_and_bsdf_agree(torch_gpu):
    torch = torch_gpu
    tile, obj, sc, cam_o, cam_d = _small(torch, res=64)
    scene = PT.Scene(PT.GpuIntersector(obj))
    o, d = cam_o.cuda(), cam_d.cuda()
    mis = PT.render(torch, scene, o, d, spp=64, depth=3, seed=1, strategy="mis").mean().item()
    bsdf = PT.render(torch, scene, o, d, spp=64, depth=3, seed=2, strategy="bsdf").mean().item()
    assert abs(mis - bsdf) <= 0.02 * mis, (mis, bsdf)


def test_pathtrace_white_furnace(torch_gpu):
    torch = torch_gpu
    tile, obj, sc, cam_o, cam_d = _small(torch, res=32)
    scene = PT.Scene(PT.GpuIntersector(obj), albedo=1.0, env=1.0, radiance=0.0, light_corner=(1e3, 1e3, 1e3))
    img = PT.render(torch, scene, cam_o.cuda(), cam_d.cuda(), spp=4, depth=24, seed=3, strategy="bsdf")
    img = img.cpu().numpy()
    assert img.max() <= 1.0 + 1e-12
    assert img.mean() > 0.99, img.mean()


def test_pathtrace_device_matches_oracle_paths(torch_gpu):
    """The device wavefront traces the torch loop's rays (same arithmetic,
    same random numbers, same intersector), so its image carries the torch
    loop's parity with the oracle: every differing pixel has a banded ray."""
    torch = torch_gpu
    tile, obj, sc, cam_o, cam_d = _small(torch)
    tris = W.quad_plane(8, 8.0)
    st = {}
    img_d = PT.render_device(PT.Scene(PT.GpuIntersector(obj)), cam_o.cuda(), cam_d.cuda(), spp=2, depth=4, seed=7,
                             stats=st).cpu().numpy()
    assert st["closest_rays"] > 3 * cam_o.shape[1] and st["shadow_rays"] > cam_o.shape[1]
    img_g, img_o, flagged, unexplained, _ = oracle_parity(torch, obj, sc, tris, 128, cam_o, cam_d, 2, 4, 7)
    assert unexplained == 0
    assert np.all(np.abs(img_d - img_g) <= 1e-12 * (1 + np.abs(img_g)))
    assert_explained(img_d, img_o, flagged, "device")


def test_pathtrace_device_matches_torch_loop(torch_gpu):
    """Same intersector, same random numbers: the device wavefront and the
    torch loop trace the same rays, so the images agree to rounding (the
    order of the per-pixel sums differs) and so do the ray counts."""
    torch = torch_gpu
    tile, obj, sc, cam_o, cam_d = _small(torch, res=64)
    scene = PT.Scene(PT.GpuIntersector(obj))
    o, d = cam_o.cuda(), cam_d.cuda()
    for strategy in ("mis", "bsdf"):
        st_d, st_t = {}, {}
        img_d = PT.render_device(scene, o, d, spp=3, depth=4, seed=11, strategy=strategy, stats=st_d).cpu().numpy()
        img_t = PT.render(torch, scene, o, d, spp=3, depth=4, seed=11, strategy=strategy, stats=st_t).cpu().numpy()
        # the same rays: equal ray counts, and pixels equal up to the order of the per-pixel sums
        assert st_d == st_t, (strategy, st_d, st_t)
        assert np.all(np.abs(img_d - img_t) <= 1e-12 * (1 + np.abs(img_t))), strategy


def test_pathtrace_device_mis_bsdf_and_furnace(torch_gpu):
    torch = torch_gpu
    tile, obj, sc, cam_o, cam_d = _small(torch, res=64)
    o, d = cam_o.cuda(), cam_d.cuda()
    scene = PT.Scene(PT.GpuIntersector(obj))
    mis = PT.render_device(scene, o, d, spp=64, depth=3, seed=1, strategy="mis").mean().item()
    bsdf = PT.render_device(scene, o, d, spp=64, depth=3, seed=2, strategy="bsdf").mean().item()
    assert abs(mis - bsdf) <= 0.02 * mis, (mis, bsdf)
    white = PT.Scene(PT.GpuIntersector(obj), albedo=1.0, env=1.0, radiance=0.0, light_corner=(1e3, 1e3, 1e3))
    img = PT.render_device(white, o[:, :1024].contiguous(), d[:, :1024].contiguous(), spp=4, depth=24, seed=3, strategy="bsdf")
    img = img.cpu().numpy()
    assert img.max() <= 1.0 + 1e-12
    assert img.mean() > 0.99, img.mean()
    # an empty image and a zero-bounce render
    empty = PT.render_device(scene, o[:, :0].contiguous(), d[:, :0].contiguous(), spp=2, depth=2)
    assert empty.numel() == 0
    direct = PT.render_device(scene, o, d, spp=1, depth=0).cpu().numpy()
    assert np.all((direct == 0.0) | (direct == scene.env) | (direct == scene.radiance))


def test_pathtrace_device_chunks(torch_gpu):
    """More paths than one wavefront holds (4,194,304): the device form runs
    them in chunks and still matches the torch loop path for path."""
    torch = torch_gpu
    tile, obj, sc, cam_o, cam_d = _small(torch, res=560)
    scene = PT.Scene(PT.GpuIntersector(obj))
    o, d = cam_o.cuda(), cam_d.cuda()
    spp = 14
    assert o.shape[1] * spp > 1 << 22
    st_d, st_t = {}, {}
    img_d = PT.render_device(scene, o, d, spp=spp, depth=2, seed=5, stats=st_d).cpu().numpy()
    img_t = PT.render(torch, scene, o, d, spp=spp, depth=2, seed=5, stats=st_t).cpu().numpy()
    assert st_d == st_t, (st_d, st_t)
    assert np.all(np.abs(img_d - img_t) <= 1e-12 * (1 + np.abs(img_t)))
