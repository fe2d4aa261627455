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
