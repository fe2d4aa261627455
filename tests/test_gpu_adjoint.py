"""The adjoint kernel (dmap_intersect_adjoint) and the differentiable
closest hit (displaced_hits): against the oracle's fp64 implicit-function
gradient, and against central finite differences (step 1e-3 in texel units,
fp64) of the oracle's forward on the differentiation workload: 1024^2 fBm at
amplitude 0.05 on one quad, 512^2 primaries, L2 loss on depth and normal
against targets rendered from the map plus Gaussian blobs.

Central differences at the stated step carry an O(h^2) truncation error
from the normal term's curvature (a few percent at 1024^2 texels, where a
1e-3 step tilts a texel's slope by 0.05); Richardson's extrapolation from
the steps 1e-3 and 5e-4, (4 D(h/2) - D(h)) / 3, removes it, and that is what
the gradient is checked against.

Non-differentiable events are excluded: rays whose GPU and oracle hits differ
in triangle or texel (grazing and crease bands) carry no upstream gradient in
the parity comparison, and a texel whose +-step moves any nearby hit to
another texel or flips a hit is skipped by the finite-difference check."""
import numpy as np
import pytest

import oracle
from dmap_amd import workload as W
from tests import scenes as S

pytestmark = pytest.mark.gpu

FD_STEP = 1e-3
FD_RTOL = 3e-3


def fold_nodes(g_nodes, n):
    """Node gradient of the closed (n+1)^2 grid (wrapping) -> texel gradient."""
    m = n + 1
    out = np.zeros((n, n))
    idx = np.arange(m) % n
    np.add.at(out, (idx[:, None], idx[None, :]), g_nodes)
    return out


def _host_hits(res):
    return {k: res[k].cpu().numpy() for k in ("t", "ng", "prim", "texel")}


def _agree(g, ref):
    """Rays whose GPU and oracle closest hits are the same event."""
    gh, rh = np.isfinite(g["t"]), np.isfinite(ref["t"])
    same = gh == rh
    both = gh & rh
    same[both] &= (g["prim"][both] == ref["prim"][both]) & np.all(g["texel"].T[both] == ref["texel"][both], axis=1)
    return same


@pytest.mark.parametrize("kind", ["flat", "curved"])
def test_adjoint_matches_oracle_small(torch_gpu, kind):
    import dmap_amd
    torch = torch_gpu
    rng = np.random.default_rng(11)
    n = 64
    tex = rng.random((n, n)).astype(np.float32)
    tris = W.quad_plane(1, 1.0) if kind == "flat" else S.curved_quad()
    tile = dmap_amd.DisplacementTile(tex, 0.05)
    obj = dmap_amd.ObjectIntersector(tile, tris)
    o, d = W.pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 40.0, 128, 128)
    rays = torch.from_numpy(W.pack_rays(o, d)).cuda()
    nr = rays.shape[1]
    g = _host_hits(obj.intersect(rays))
    sc = oracle.Scene(W.close_tile(tex).astype(np.float64), 0.05, 0.0, True, tris)
    ref = sc.intersect(o.astype(np.float64), d.astype(np.float64))
    keep = _agree(g, ref) & np.isfinite(g["t"])
    assert keep.mean() > 0.3
    lt = (rng.normal(size=nr) * keep).astype(np.float32)
    ln = (rng.normal(size=(3, nr)) * keep).astype(np.float32)
    g_gpu = obj.adjoint(rays, torch.from_numpy(lt).cuda(), torch.from_numpy(ln).cuda()).cpu().numpy()
    g_nodes, _, _ = sc.adjoint(o.astype(np.float64), d.astype(np.float64), lt.astype(np.float64),
                               ln.T.astype(np.float64))
    g_ref = fold_nodes(g_nodes, n)
    scale = np.abs(g_ref).max()
    assert scale > 0
    assert np.max(np.abs(g_gpu - g_ref)) <= 1e-8 * scale, np.max(np.abs(g_gpu - g_ref)) / scale
    # accumulation: a second call adds into the same buffer
    acc = obj.adjoint(rays, torch.from_numpy(lt).cuda(), torch.from_numpy(ln).cuda(),
                      grad=torch.from_numpy(g_gpu).cuda()).cpu().numpy()
    assert np.allclose(acc, 2 * g_gpu, rtol=1e-12, atol=1e-12 * scale)


def test_adjoint_depth_only_and_validation(torch_gpu):
    import dmap_amd
    torch = torch_gpu
    tex = np.random.default_rng(2).random((32, 32)).astype(np.float32)
    tile = dmap_amd.DisplacementTile(tex, 0.05)
    obj = dmap_amd.ObjectIntersector(tile, W.quad_plane(1, 1.0))
    o, d = W.pinhole_rays((0.5, -0.8, 0.9), (0.5, 0.5, 0.0), 40.0, 32, 32)
    rays = torch.from_numpy(W.pack_rays(o, d)).cuda()
    lt = torch.ones(rays.shape[1], dtype=torch.float32, device="cuda")
    g = obj.adjoint(rays, lt).cpu().numpy()
    sc = oracle.Scene(W.close_tile(tex).astype(np.float64), 0.05, 0.0, True, W.quad_plane(1, 1.0))
    hits = np.isfinite(sc.intersect(o.astype(np.float64), d.astype(np.float64))["t"])
    g_ref = fold_nodes(sc.adjoint(o.astype(np.float64), d.astype(np.float64), hits.astype(np.float64))[0], 32)
    assert np.max(np.abs(g - g_ref)) <= 1e-6 * np.abs(g_ref).max()
    # Raising the surface under a camera above it shortens every hit: dt/dh < 0 in sum.
    assert g.sum() < 0
    with pytest.raises(dmap_amd.DmapError):
        obj.adjoint(rays, lt, options=dmap_amd.TraversalOptions(leaf="two_triangle"))
    with pytest.raises(dmap_amd.DmapError):
        obj.adjoint(rays, lt[:10])


@pytest.fixture(scope="module")
def c4(torch_gpu):
    import dmap_amd
    torch = torch_gpu
    print("[adjoint] differentiation workload: 1024^2 fBm, 512^2 primaries", flush=True)
    wl = W.c4_adjoint()
    tile = dmap_amd.DisplacementTile(wl.texels, wl.strength)
    obj = dmap_amd.ObjectIntersector(tile, wl.triangles)
    rays = torch.from_numpy(wl.rays).cuda()
    tgt_tile = dmap_amd.DisplacementTile(wl.target_texels, wl.strength)
    tgt_obj = dmap_amd.ObjectIntersector(tgt_tile, wl.triangles)
    tgt = tgt_obj.intersect(rays)
    return {"wl": wl, "tile": tile, "obj": obj, "rays": rays, "t_tgt": tgt["t"].double().clone(),
            "n_tgt": tgt["ng"].double().clone()}


def _oracle_forward(wl, texels, o, d):
    sc = oracle.Scene(W.close_tile(texels).astype(np.float64), wl.strength, 0.0, True, wl.triangles)
    _, t, ng = sc.adjoint(o, d, np.zeros(o.shape[0]))
    return sc, t, ng


def test_adjoint_c4_autograd_oracle_and_finite_differences(c4, torch_gpu):
    import dmap_amd
    torch = torch_gpu
    wl, tile, obj, rays = c4["wl"], c4["tile"], c4["obj"], c4["rays"]
    n = wl.texels.shape[0]
    texels = torch.from_numpy(wl.texels).cuda().requires_grad_(True)
    t, ng = dmap_amd.displaced_hits(texels, tile, obj, rays)
    mask = torch.isfinite(t) & torch.isfinite(c4["t_tgt"])
    dt = t.double()[mask] - c4["t_tgt"][mask]
    dn = ng.double()[:, mask] - c4["n_tgt"][:, mask]
    loss = (dt * dt).sum() + (dn * dn).sum()
    loss.backward()
    g_auto = texels.grad.double().cpu().numpy()
    scale = np.abs(g_auto).max()
    assert scale > 0 and np.isfinite(g_auto).all()

    # 1. The autograd path is the adjoint entry point with dL/dt, dL/dng.
    m = mask.cpu().numpy()
    lt = torch.zeros_like(t.detach())
    lt[mask] = (2 * dt.detach()).float()
    ln = torch.zeros_like(ng.detach())
    ln[:, mask] = (2 * dn.detach()).float()
    g_direct = obj.adjoint(rays, lt, ln).cpu().numpy()
    assert np.max(np.abs(g_direct - g_auto)) <= 1e-6 * scale

    # 2. The fp64 CPU gradient (oracle) with the same upstream gradients, on
    # the rays whose GPU and oracle hits are the same event.
    o = wl.rays[0:3].T.astype(np.float64)
    d = wl.rays[3:6].T.astype(np.float64)
    sc, t_o, ng_o = _oracle_forward(wl, wl.texels, o, d)
    ref = sc.intersect(o, d)
    g_hits = _host_hits(obj.intersect(rays))
    keep = _agree(g_hits, ref)
    assert keep.mean() > 0.999
    lt_h = lt.cpu().numpy() * keep
    ln_h = ln.cpu().numpy() * keep
    g_gpu = obj.adjoint(rays, torch.from_numpy(lt_h).cuda(), torch.from_numpy(ln_h).cuda()).cpu().numpy()
    g_cpu = fold_nodes(sc.adjoint(o, d, lt_h.astype(np.float64), ln_h.T.astype(np.float64))[0], n)
    assert np.max(np.abs(g_gpu - g_cpu)) <= 1e-7 * np.abs(g_cpu).max(), np.max(np.abs(g_gpu - g_cpu))

    # 3. Central differences of the fp64 loss on 64 seeded texels, over the
    # rays whose hits lie within two texels of the perturbed one.
    t_tgt = c4["t_tgt"].cpu().numpy()
    n_tgt = c4["n_tgt"].cpu().numpy().T
    tex_hit = ref["texel"]

    def window_loss(texels64, rows):
        sc1, t1, ng1 = _oracle_forward(wl, texels64, o[rows], d[rows])
        r1 = sc1.intersect(o[rows], d[rows])
        ev = np.isfinite(r1["t"]) != np.isfinite(ref["t"][rows])
        ev |= np.any(r1["texel"] != tex_hit[rows], axis=1) & np.isfinite(r1["t"])
        dtt = t1 - t_tgt[rows]
        dnn = ng1 - n_tgt[rows]
        return float(np.sum(dtt * dtt) + np.sum(dnn * dnn)), bool(ev.any())

    rng = np.random.default_rng(2029)
    cand = np.argwhere(np.abs(g_auto) > 1e-3 * scale)
    cand = cand[rng.permutation(len(cand))]
    base = wl.texels.astype(np.float64)
    checked, skipped, worst = 0, 0, 0.0
    for j, i in cand:
        if checked == 64:
            break
        di = (tex_hit[:, 0] - i + n // 2) % n - n // 2
        dj = (tex_hit[:, 1] - j + n // 2) % n - n // 2
        rows = np.flatnonzero(m & keep & (np.abs(di) <= 2) & (np.abs(dj) <= 2))
        moved = False
        D = []
        for h in (FD_STEP, FD_STEP / 2):
            vals = []
            for sgn in (1.0, -1.0):
                pert = base.copy()
                pert[j, i] += sgn * h
                v, ev = window_loss(pert, rows)
                vals.append(v)
                moved |= ev
            D.append((vals[0] - vals[1]) / (2 * h))
        if moved:
            skipped += 1
            continue
        fd = (4 * D[1] - D[0]) / 3
        # against the adjoint of the same rays (the agreeing ones)
        err = abs(fd - g_gpu[j, i])
        worst = max(worst, err / max(abs(fd), 1e-12))
        assert err <= FD_RTOL * abs(fd) + 1e-6 * scale, (j, i, D, fd, g_gpu[j, i])
        checked += 1
    print(f"[adjoint] finite differences: {checked} texels checked, {skipped} skipped (events), "
          f"worst relative error {worst:.2e}", flush=True)
    assert checked == 64
