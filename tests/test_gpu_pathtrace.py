"""The wavefront path tracer (config 3's estimator), in two forms: the torch
loop over the batched GPU intersector and the device wavefront
(dmap_pathtrace). The same paths on the CPU oracle give the same image
(shared counter-based random numbers; paths part only where a hit differs),
the two GPU forms give the same image, NEE with MIS and BSDF-only sampling
agree in expectation, and a white furnace (albedo 1, environment 1, no
light) returns at most 1 everywhere and about 1 on average."""
import numpy as np
import pytest

import oracle
from dmap_amd import pathtrace as PT
from dmap_amd import workload as W
from tests.pt_oracle import CpuIntersector

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
    gpu = PT.Scene(PT.GpuIntersector(obj))
    cpu = PT.Scene(CpuIntersector(sc))
    st = {}
    img_g = PT.render(torch, gpu, cam_o.cuda(), cam_d.cuda(), spp=2, depth=4, seed=7, stats=st).cpu().numpy()
    img_c = PT.render(torch, cpu, cam_o, cam_d, spp=2, depth=4, seed=7).numpy()
    assert st["closest_rays"] > 3 * cam_o.shape[1] and st["shadow_rays"] > cam_o.shape[1]
    close = np.abs(img_g - img_c) <= 1e-4 * (1 + np.abs(img_c))
    assert close.mean() > 0.97, close.mean()
    assert abs(img_g.mean() - img_c.mean()) <= 1e-2 * img_c.mean()


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
    torch = torch_gpu
    tile, obj, sc, cam_o, cam_d = _small(torch)
    gpu = PT.Scene(PT.GpuIntersector(obj))
    cpu = PT.Scene(CpuIntersector(sc))
    st = {}
    img_d = PT.render_device(gpu, cam_o.cuda(), cam_d.cuda(), spp=2, depth=4, seed=7, stats=st).cpu().numpy()
    img_c = PT.render(torch, cpu, cam_o, cam_d, spp=2, depth=4, seed=7).numpy()
    assert st["closest_rays"] > 3 * cam_o.shape[1] and st["shadow_rays"] > cam_o.shape[1]
    close = np.abs(img_d - img_c) <= 1e-4 * (1 + np.abs(img_c))
    assert close.mean() > 0.97, close.mean()
    assert abs(img_d.mean() - img_c.mean()) <= 1e-2 * img_c.mean()


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
        close = np.abs(img_d - img_t) <= 1e-9 * (1 + np.abs(img_t))
        assert close.mean() > 0.995, (strategy, close.mean())
        assert abs(st_d["closest_rays"] - st_t["closest_rays"]) <= 1e-3 * st_t["closest_rays"], (st_d, st_t)
        assert abs(st_d["shadow_rays"] - st_t["shadow_rays"]) <= 1e-3 * max(st_t["shadow_rays"], 1), (st_d, st_t)


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
    close = np.abs(img_d - img_t) <= 1e-9 * (1 + np.abs(img_t))
    assert close.mean() > 0.995, close.mean()
