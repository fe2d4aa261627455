"""The path tracer's intersector calls on the CPU oracle (test
infrastructure: the GPU/oracle parity of dmap_amd.pathtrace)."""
import numpy as np


class CpuIntersector:
    """The same calls on the CPU oracle, for parity tests (test infrastructure:
    takes an oracle.Scene)."""

    def __init__(self, oracle_scene, threads=0):
        self.sc = oracle_scene
        self.threads = threads

    def closest(self, o, d):
        import torch
        # the ray as the GPU receives it: fp32 components
        oo = o.to(torch.float32).to(torch.float64).T.numpy()
        dd = d.to(torch.float32).to(torch.float64).T.numpy()
        r = self.sc.intersect(oo, dd, threads=self.threads)
        return {"t": torch.from_numpy(r["t"].astype(np.float32).astype(np.float64)),
                "ng": torch.from_numpy(r["ng"].T.astype(np.float32).astype(np.float64).copy())}

    def occluded(self, o, d, tmax):
        import torch
        oo = o.to(torch.float32).to(torch.float64).T.numpy()
        dd = d.to(torch.float32).to(torch.float64).T.numpy()
        tm = tmax.to(torch.float32).to(torch.float64).numpy()
        return torch.from_numpy(self.sc.occluded(oo, dd, None, tm, threads=self.threads))


class DualIntersector:
    """Traces every batch on the GPU library and on the oracle, on the same
    fp32 rays, and answers with one of them (`primary`, "gpu" or "oracle",
    exactly as GpuIntersector / CpuIntersector would); the other checks it
    ray by ray under tests/parity.py's bar. `flags` gets one int8 array per
    call: 0 the two agree, 1 they differ inside a stated band, 2 they differ
    unexplained. Any-hit answers that differ are banded when the nearest
    crossing (the closest hit of either side, unbounded) is banded or lies
    within 1e-4 (relative) of the shadow ray's t_max."""

    def __init__(self, obj, oracle_scene, tri_rows, n_leaf, primary="gpu", threads=0):
        self.obj, self.sc = obj, oracle_scene
        self.tri, self.n_leaf = tri_rows, n_leaf
        self.primary = primary
        self.threads = threads
        self.flags = []

    def _rays(self, o, d, tmax=None):
        import torch
        n = o.shape[1]
        r = torch.empty((8, n), dtype=torch.float32)
        r[0:3] = o.cpu().to(torch.float32)
        r[3:6] = d.cpu().to(torch.float32)
        r[6] = 0.0
        r[7] = float("inf") if tmax is None else tmax.cpu().to(torch.float32)
        return r

    def _both(self, r):
        import torch
        g = self.obj.intersect(r.cuda())
        gh = {"t": g["t"].cpu().numpy(), "st": g["st"].cpu().numpy().T, "ng": g["ng"].cpu().numpy().T,
              "prim": g["prim"].cpu().numpy()}
        rn = r.numpy().astype(np.float64)
        d = rn[3:6].T
        ref = self.sc.intersect(rn[0:3].T, d, rn[6], rn[7], threads=self.threads)
        torch.cuda.synchronize()
        return gh, ref, d

    def closest(self, o, d):
        import torch
        from tests.parity import compare
        r = self._rays(o, d)
        gh, ref, dd = self._both(r)
        self.flags.append(compare(gh, ref, dd, self.tri, self.n_leaf)["ray_flag"])
        if self.primary == "gpu":
            t, ng = gh["t"].astype(np.float64), gh["ng"].T.astype(np.float64)
        else:
            t = ref["t"].astype(np.float32).astype(np.float64)
            ng = ref["ng"].T.astype(np.float32).astype(np.float64)
        return {"t": torch.from_numpy(t).to(o.device), "ng": torch.from_numpy(np.ascontiguousarray(ng)).to(o.device)}

    def occluded(self, o, d, tmax):
        import torch
        from tests.parity import compare
        r = self._rays(o, d, tmax)
        occ_g = self.obj.occluded(r.cuda()).cpu().numpy().astype(bool)
        rn = r.numpy().astype(np.float64)
        occ_o = self.sc.occluded(rn[0:3].T, rn[3:6].T, rn[6], rn[7], threads=self.threads)
        flag = np.zeros(len(occ_g), np.int8)
        diff = np.nonzero(occ_g != occ_o)[0]
        if len(diff):
            ru = r[:, diff].clone()
            ru[7] = float("inf")
            gh, ref, dd = self._both(ru)
            f = compare(gh, ref, dd, self.tri, self.n_leaf)["ray_flag"]
            t_near = np.fmin(gh["t"].astype(np.float64), ref["t"])
            tm = rn[7, diff]
            near_end = np.abs(t_near - tm) <= 1e-4 * tm
            flag[diff] = np.where((f == 1) | near_end, 1, 2)
        self.flags.append(flag)
        occ = occ_g if self.primary == "gpu" else occ_o
        return torch.from_numpy(occ).to(o.device)


def flagged_pixels(trace_log, flags, n_pix):
    """Pixels with a ray that differs between GPU and oracle on their paths
    (any sample, any bounce), and the number of unexplained rays."""
    assert len(trace_log) == len(flags)
    out = np.zeros(n_pix, bool)
    unexplained = 0
    for (_, pix), f in zip(trace_log, flags):
        out[pix.numpy()[f != 0]] = True
        unexplained += int((f == 2).sum())
    return out, unexplained


def oracle_parity(torch, obj, sc, tris, n_leaf, cam_o, cam_d, spp, depth, seed):
    """The torch loop twice over a checking intersector, once following the
    GPU's hits and once the oracle's, every batch traced by both. Returns the
    two images, the pixels with a ray on their paths where the two differ
    (under the parity bar, inside a band) and the count of rays that differ
    outside every band, which must be 0."""
    dual_g = DualIntersector(obj, sc, tris, n_leaf, primary="gpu")
    dual_o = DualIntersector(obj, sc, tris, n_leaf, primary="oracle")
    log_g, log_o = [], []
    from dmap_amd import pathtrace as PT
    img_g = PT.render(torch, PT.Scene(dual_g), cam_o.cuda(), cam_d.cuda(), spp=spp, depth=depth, seed=seed,
                      trace_log=log_g).cpu().numpy()
    img_o = PT.render(torch, PT.Scene(dual_o), cam_o.cpu(), cam_d.cpu(), spp=spp, depth=depth, seed=seed,
                      trace_log=log_o).numpy()
    n = cam_o.shape[1]
    fg, ug = flagged_pixels(log_g, dual_g.flags, n)
    fo, uo = flagged_pixels(log_o, dual_o.flags, n)
    rays = sum(len(f) for f in dual_g.flags) + sum(len(f) for f in dual_o.flags)
    return img_g, img_o, fg | fo, ug + uo, rays


def assert_explained(img, img_o, flagged, label):
    """Every pixel where img and the oracle-driven image differ beyond 1e-4
    (relative) has a banded ray on its GPU or oracle paths."""
    differ = np.abs(img - img_o) > 1e-4 * (1 + np.abs(img_o))
    print(f"[{label}] {differ.mean():.4f} of pixels differ, {flagged.mean():.4f} have a banded ray", flush=True)
    assert not (differ & ~flagged).any(), (label, np.nonzero(differ & ~flagged)[0][:10])
    assert abs(img.mean() - img_o.mean()) <= 1e-2 * img_o.mean()
