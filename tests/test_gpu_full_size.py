"""Parity at the workload configurations' own sizes.

* Incoherent closest hit: the 2048^2 fBm + ridged map on 64 quads, all
  4,194,304 rays (2,097,152 pinhole primaries, 2,097,152 cosine bounces),
  against the oracle on all host threads under tests/parity.py's bar.
* Serving: the eight resident scenes (512^2, 1024^2, 2048^2 x 3, 4096^2 x 2,
  8192^2), 19 requests of 1k to 1M rays (three of them above 512k); every
  request's hits through the server equal the single-request call bit for
  bit, and a seeded sample of each request's rays matches the oracle.
* Path tracing: the 4096^2 map, the 1024^2 camera, 4 samples per pixel,
  depth 4. The device wavefront over the whole image traces the torch loop's
  rays (equal ray counts, equal pixels up to summation order); on a seeded
  subset of 2,048 pixels, the torch loop follows the GPU's and the oracle's
  hits with every batch traced by both: no ray differs outside the parity
  bands, and every pixel that differs has a banded ray on its paths.
"""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests.parity import compare

pytestmark = pytest.mark.gpu


def report(cmp):
    return {k: v for k, v in cmp.items() if k not in ("bad_index", "ray_flag")}


def test_c2_full_size_closest_hit_parity(torch_gpu):
    import dmap_amd
    torch = torch_gpu
    print("[full] C2: 2048^2 map, 128 triangles, 4,194,304 rays", flush=True)
    wl = W.c2_incoherent()
    assert wl.rays.shape[1] == 4194304 and wl.texels.shape == (2048, 2048)
    tile = dmap_amd.DisplacementTile(wl.texels, wl.strength, wl.offset, wl.repeat)
    obj = dmap_amd.ObjectIntersector(tile, wl.triangles)
    res = obj.intersect(torch.from_numpy(wl.rays).cuda())
    torch.cuda.synchronize()
    g = {"t": res["t"].cpu().numpy(), "prim": res["prim"].cpu().numpy(), "st": res["st"].cpu().numpy().T,
         "ng": res["ng"].cpu().numpy().T}
    sc = oracle.Scene(wl.nodes, wl.strength, wl.offset, wl.repeat, wl.triangles)
    o = wl.rays[0:3].T.astype(np.float64)
    d = wl.rays[3:6].T.astype(np.float64)
    print("[full] C2: oracle on all host threads", flush=True)
    ref = sc.intersect(o, d, wl.rays[6].astype(np.float64), wl.rays[7].astype(np.float64))
    cmp = compare(g, ref, d, wl.triangles, 2048)
    print(report(cmp), flush=True)
    assert cmp["hits_ref"] > 1_000_000
    assert cmp["unexplained"] == 0, report(cmp)


def test_c5_serving_full_scenes(torch_gpu):
    import dmap_amd
    from dmap_amd import serving
    torch = torch_gpu
    print("[full] C5: building the 8 resident scenes (up to 8192^2)", flush=True)
    scenes = serving.make_scenes()
    assert [s.n for s in scenes] == [512, 1024, 2048, 2048, 2048, 4096, 4096, 8192]
    reqs = serving.make_requests(scenes, 16, seed=2027)
    rng = np.random.default_rng(5)
    for k, n in ((7, 1_000_000), (5, 900_000), (2, 600_000)):
        S = scenes[k]
        n_p = int(round(serving.PRIMARY_FRACTION * n))
        sp = int(rng.integers(0, (S.pool - n_p) // 64 + 1)) * 64
        sb = int(rng.integers(0, S.pool - (n - n_p) + 1))
        reqs.append((k, torch.cat([S.primary[:, sp:sp + n_p], S.bounce[:, sb:sb + n - n_p]], dim=1).contiguous()))
    sizes = [r.shape[1] for _, r in reqs]
    assert len(reqs) >= 16 and sum(n > 512 * 1024 for n in sizes) >= 3
    srv = dmap_amd.Server(slots=3, step_rays=1 << 20, hit_pool_bytes=4 << 30)
    for q, (k, rays) in enumerate(reqs):
        srv.submit(scenes[k].obj, rays, q)
    got = {}
    while len(got) < len(reqs):
        for tag, _ in srv.poll():
            h = srv.hits(tag, reqs[tag][1].shape[1])
            got[tag] = {key: v.cpu().numpy() for key, v in h.items()}
            srv.release(tag)
    torch.cuda.synchronize()
    assert srv.idle()
    print(f"[full] C5: {len(reqs)} requests, {sum(sizes)} rays served; checking each", flush=True)
    for q, (k, rays) in enumerate(reqs):
        plain = {key: v.cpu().numpy() for key, v in scenes[k].obj.intersect(rays).items()}
        for key in plain:
            assert np.array_equal(got[q][key], plain[key]), (q, key)
    # A seeded sample of each request's rays against the oracle, one scene at a time.
    for k in sorted({k for k, _ in reqs}):
        S = scenes[k]
        nodes = W.close_tile(S.tile._keep.cpu().numpy()).astype(np.float64)
        tris = W.quad_plane(8, 8.0)
        sc = oracle.Scene(nodes, 0.1, 0.0, True, tris)
        for q, (kk, rays) in enumerate(reqs):
            if kk != k:
                continue
            n = rays.shape[1]
            idx = np.sort(np.random.default_rng(q).choice(n, size=min(n, 4096), replace=False))
            r = rays[:, torch.from_numpy(idx).cuda()].cpu().numpy().astype(np.float64)
            ref = sc.intersect(r[0:3].T, r[3:6].T, r[6], r[7])
            g = {"t": got[q]["t"][idx], "prim": got[q]["prim"][idx], "st": got[q]["st"][:, idx].T,
                 "ng": got[q]["ng"][:, idx].T}
            cmp = compare(g, ref, r[3:6].T, tris, S.n)
            assert cmp["unexplained"] == 0, (q, k, report(cmp))
        print(f"[full] C5: scene {k} ({S.n}^2) matches the oracle", flush=True)
        del sc


def test_c3_full_size_path_tracing(torch_gpu):
    from dmap_amd import pathtrace as PT
    from tests.pt_oracle import assert_explained, oracle_parity
    torch = torch_gpu
    print("[full] C3: 4096^2 map, 1024^2 pixels x 4 spp, depth 4", flush=True)
    scene, cam_o, cam_d = PT.c3_scene(torch)
    assert scene.tile.n_leaf == 4096 and cam_o.shape[1] == 1024 * 1024
    st_d, st_t = {}, {}
    img_d = PT.render_device(scene, cam_o, cam_d, spp=4, depth=4, seed=2027, stats=st_d).cpu().numpy()
    img_t = PT.render(torch, scene, cam_o, cam_d, spp=4, depth=4, seed=2027, stats=st_t).cpu().numpy()
    print(f"[full] C3: device {st_d}, torch loop {st_t}", flush=True)
    assert st_d == st_t and st_d["closest_rays"] > 6_000_000
    assert np.all(np.abs(img_d - img_t) <= 1e-12 * (1 + np.abs(img_t)))
    # The seeded pixel subset against the oracle.
    sub = np.sort(np.random.default_rng(3).choice(cam_o.shape[1], size=2048, replace=False))
    so, sd = cam_o[:, sub].contiguous(), cam_d[:, sub].contiguous()
    nodes = W.close_tile(scene.tile._keep.cpu().numpy()).astype(np.float64)
    tris = W.quad_plane(8, 8.0)
    sc = oracle.Scene(nodes, 0.1, 0.0, True, tris)
    img_g, img_o, flagged, unexplained, rays = oracle_parity(torch, scene.obj, sc, tris, 4096, so, sd, 4, 4, 2027)
    print(f"[full] C3 subset: {rays} rays checked, {unexplained} unexplained", flush=True)
    assert unexplained == 0
    img_sub = PT.render_device(scene, so, sd, spp=4, depth=4, seed=2027).cpu().numpy()
    assert np.all(np.abs(img_sub - img_g) <= 1e-12 * (1 + np.abs(img_g)))
    assert_explained(img_g, img_o, flagged, "C3 subset")
