"""Run one ray set through the GPU library and through the oracle."""
import numpy as np

import oracle


def gpu_hits(obj, rays, **kw):
    import torch
    res = obj.intersect(torch.from_numpy(np.ascontiguousarray(rays)).cuda(), **kw)
    torch.cuda.synchronize()
    g = {"t": res["t"].cpu().numpy(), "prim": res["prim"].cpu().numpy(), "st": res["st"].cpu().numpy().T,
         "bary": res["bary"].cpu().numpy().T, "ng": res["ng"].cpu().numpy().T, "texel": res["texel"].cpu().numpy().T}
    if "stats" in res:
        g["stats"] = {k: v.cpu().numpy() for k, v in res["stats"].items()}
    return g


def oracle_hits(nodes, strength, offset, repeat, tris, rays, build="fold", **kw):
    sc = oracle.Scene(nodes, strength, offset, repeat, tris, build)
    o = rays[0:3].T.astype(np.float64)
    d = rays[3:6].T.astype(np.float64)
    ref = sc.intersect(o, d, rays[6].astype(np.float64), rays[7].astype(np.float64), **kw)
    return ref, d, sc
