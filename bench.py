"""Closest-hit throughput on one MI355X for the incoherent displaced-surface
workload (64 quads, 2048^2 fBm + ridged map, 4,194,304 rays: 2048 x 1024
pinhole primaries and 2,097,152 cosine-weighted bounce rays).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

A step is one closest-hit launch over all 4,194,304 rays, rays already
resident in HBM. Prints one JSON line: Mrays/s, the traversal kernel's
roofline position (algorithmic bytes from the kernel's own work counters
over the same rays, counted in a run outside the timed region; the oracle's
counters, from the reference's walk, beside them), and the CPU oracle timed
on the host's cores on a sample of the rays.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "closest-hit Mrays/s vs CPU ref; batched serving p50/p99 latency at fixed load"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# Algorithmic bytes per ray: the ray in (8 fp32) and the hit out (8 fp32
# + 3 int32 counted as 32 B), 32 B per fp32 node test (its 8-channel node),
# 16 B per leaf test (its 4 corner heights).
BYTES_RAY_IO = 32 + 32
BYTES_NODE = 32
BYTES_LEAF = 16
# HBM bytes the traversal kernel moves per ray, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE on this workload, each scaled by the same
# counter's reading for a device copy of known size (bench.py
# --calibrate-copy); profiles/k_trace_c2_incoherent_traffic.md.
TRAFFIC_BYTES_PER_RAY = 591.8
# fp32 operations per unit of work, counted from the restated formulas
# (node test with box + slab, sub-square bound, Newton iteration).
FLOPS_NODE = 200
FLOPS_SUB = 120
FLOPS_NEWTON = 110


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=1, help="oracle sample: every k-th ray")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--order", default="near_first", choices=["near_first", "given", "sort"],
                    help="the order the rays are traced in: rays starting inside the object's bounds first "
                         "(default), as given, or sorted by direction octant and origin")
    ap.add_argument("--serve-seconds", type=float, default=2.0,
                    help="length of the fixed-rate serving run (0: no serving section)")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="serving arrival rate as a fraction of the plain kernel's Mrays/s on the same rays")
    ap.add_argument("--serve-slots", type=int, default=3)
    ap.add_argument("--serve-step", type=int, default=1 << 20, help="rays per serving step")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint timing section")
    ap.add_argument("--no-pathtrace", action="store_true", help="skip the path-tracing timing section")
    ap.add_argument("--no-fp64", action="store_true", help="skip the fp64 device run and the fp32 error report")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (the PMC byte-counter reference) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): the library has no multi-GPU path; "
              "rays would shard by index range with no exchange step", file=sys.stderr)
        sys.exit(2)

    import torch

    import dmap_amd
    from dmap_amd import workload as W

    t0 = time.time()
    wl = W.c2_incoherent()
    n = wl.rays.shape[1]
    log(f"[bench] workload {wl.description}: generated in {time.time() - t0:.1f}s")
    tile = dmap_amd.DisplacementTile(wl.texels, wl.strength, wl.offset, wl.repeat, "fold")
    obj = dmap_amd.ObjectIntersector(tile, wl.triangles)
    rays = torch.from_numpy(wl.rays).cuda()
    hits = obj.alloc_hits(n)
    opt = dmap_amd.TraversalOptions(precision=args.precision, reorder=args.order)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        obj.intersect(rays, options=opt, out=hits)
    torch.cuda.synchronize()
    log(f"[bench] warmup done ({args.warmup} steps)")

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    dmap_amd.kernel_timing_read()
    dmap_amd.kernel_timing(True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        obj.intersect(rays, options=opt, out=hits)
        ends[k].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    dmap_amd.kernel_timing(False)
    # The traversal kernel alone: HIP events the library records around it
    # on the launch's stream.
    trace_ms_sum, launches = dmap_amd.kernel_timing_read()
    avg_ms = trace_ms_sum / max(launches, 1)
    call_ms = float(np.mean([s.elapsed_time(e) for s, e in zip(starts, ends)]))
    ms_per_step = wall * 1e3 / args.steps
    value = n / (ms_per_step * 1e-3) / 1e6
    t_gpu = hits["f"][0].cpu().numpy()
    log(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step (call {call_ms:.3f} ms, traversal kernel "
        f"{avg_ms:.3f} ms over {launches} launches), {value:.1f} Mrays/s, hit fraction {np.isfinite(t_gpu).mean():.4f}")

    # Tile build (kernel #1) timing, for the record.
    b0, b1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    tile.rebuild()
    b0.record(stream)
    for _ in range(5):
        tile.rebuild()
    b1.record(stream)
    torch.cuda.synchronize()
    build_ms = b0.elapsed_time(b1) / 5
    nn = tile.n_leaf
    build_bytes = 4 * (nn + 1) ** 2 + 32 * ((4 * nn * nn - 1) // 3)

    # The kernel's own work over the same rays, outside the timed region.
    sres = obj.intersect(rays, options=opt, stats=True)
    torch.cuda.synchronize()
    gst = {k: sres["stats"][k].double().mean().item()
           for k in ("node_tests", "leaf_tests", "sub_squares", "newton_iters", "triangle_tests")}
    assert np.array_equal(sres["t"].cpu().numpy(), t_gpu), "the counting run's hits differ from the timed run's"
    del sres
    bytes_ray = BYTES_RAY_IO + BYTES_NODE * gst["node_tests"] + BYTES_LEAF * gst["leaf_tests"]
    flops_ray = FLOPS_NODE * gst["node_tests"] + FLOPS_SUB * gst["sub_squares"] + FLOPS_NEWTON * gst["newton_iters"]
    achieved = bytes_ray * n / (avg_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": int(TRAFFIC_BYTES_PER_RAY * n) if TRAFFIC_BYTES_PER_RAY else None,
            "kernel": "k_trace<float>", "kernel_ms": round(avg_ms, 4), "bytes_per_ray": round(bytes_ray, 1),
            "rays_per_launch": n, "work_counted_by": "the kernel's own counters, all rays"}
    cpu = None
    extra = {"valu": {"achieved_tflops": round(flops_ray * n / (avg_ms * 1e-3) / 1e12, 3),
                      "peak_tflops": FP32_PEAK_TFLOPS,
                      "frac": round(flops_ray * n / (avg_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, 5),
                      "flops_per_ray": round(flops_ray, 1)},
             "work_per_ray": {"gpu": {"node_tests": round(gst["node_tests"], 3),
                                      "leaf_tests": round(gst["leaf_tests"], 4),
                                      "sub_squares": round(gst["sub_squares"], 3),
                                      "newton_iters": round(gst["newton_iters"], 3),
                                      "triangle_tests": round(gst["triangle_tests"], 3)}}}
    if args.precision == "fp32" and not args.no_fp64:
        extra["fp32_vs_fp64"] = fp64_report(obj, rays, hits, wl, tile.n_leaf, log)
    if not args.no_cpu:
        import oracle
        idx = np.arange(0, n, args.cpu_sample)
        o = wl.rays[0:3, idx].T.astype(np.float64)
        d = wl.rays[3:6, idx].T.astype(np.float64)
        sc = oracle.Scene(wl.nodes, wl.strength, wl.offset, wl.repeat, wl.triangles)
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        log(f"[bench] cpu oracle on {len(idx)} rays, {threads} threads")
        c0 = time.perf_counter()
        ref = sc.intersect(o, d, wl.rays[6, idx].astype(np.float64), wl.rays[7, idx].astype(np.float64),
                           threads=threads, stats=True)
        cpu_s = time.perf_counter() - c0
        cpu_mrays = len(idx) / cpu_s / 1e6
        st = ref["stats"]
        node = float(st["node_tests"].mean())
        leaf = float(st["leaf_tests"].mean())
        sub = float(st["sub_squares"].mean())
        its = float(st["newton_iters"].mean())
        # The reference's walk (from the triangles' roots) on the sample,
        # and the algorithmic bytes it would read.
        ref_bytes = BYTES_RAY_IO + BYTES_NODE * node + BYTES_LEAF * leaf
        extra["work_per_ray"]["oracle"] = {
            "node_tests": round(node, 3), "leaf_tests": round(leaf, 4), "sub_squares": round(sub, 3),
            "newton_iters": round(its, 3), "bytes_per_ray": round(ref_bytes, 1)}
        # The same kernel time against the reference walk's work on the same
        # rays (the CPU oracle's counters): the rate at which the kernel
        # delivers the reference algorithm's results.
        ref_achieved = ref_bytes * n / (avg_ms * 1e-3) / 1e9
        roof["reference_walk"] = {
            "what": "algorithmic bytes of the reference's walk from the triangles' roots (CPU oracle counters on "
                    "the same rays) over the same kernel time",
            "bytes_per_ray": round(ref_bytes, 1), "achieved": round(ref_achieved, 2),
            "frac": round(ref_achieved / HBM_PEAK_GBS, 5)}
        # The rays that reach the pyramid: those with at least one node test
        # (by the oracle's counters); the others end at the object's boxes.
        active = st["node_tests"] > 0
        n_prim = wl.n_primary
        extra["active_rays"] = {
            "what": "rays with at least one node test (oracle counters); the rest leave the object's boxes",
            "fraction": round(float(active.mean()), 4),
            "primaries": round(float(active[idx < n_prim].mean()), 4) if (idx < n_prim).any() else None,
            "bounces": round(float(active[idx >= n_prim].mean()), 4) if (idx >= n_prim).any() else None,
            "mrays_s": round(value * float(active.mean()), 2),
            "node_tests_per_active_ray": round(float(st["node_tests"][active].mean()), 3)}
        # Parity of the timed result on the sample.
        gt = t_gpu[idx].astype(np.float64)
        both = np.isfinite(gt) & np.isfinite(ref["t"])
        extra["sample_parity"] = {"hitmiss_flips": int((np.isfinite(gt) != np.isfinite(ref["t"])).sum()),
                                  "worst_t_rel": float(np.max(np.abs(gt[both] - ref["t"][both]) / ref["t"][both]))
                                  if both.any() else 0.0}
        # One thread, best of 3 with a checksum of t (the reference's own
        # method, test_traversal_cost.cpp:117-135), on every 16th ray.
        idx1 = np.arange(0, n, 16)
        o1 = wl.rays[0:3, idx1].T.astype(np.float64)
        d1 = wl.rays[3:6, idx1].T.astype(np.float64)
        best = np.inf
        for _ in range(3):
            c0 = time.perf_counter()
            r1 = sc.intersect(o1, d1, wl.rays[6, idx1].astype(np.float64), wl.rays[7, idx1].astype(np.float64),
                              threads=1)
            best = min(best, time.perf_counter() - c0)
        checksum = float(np.sum(r1["t"][np.isfinite(r1["t"])]))
        cpu = {"value": round(cpu_mrays, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": (f"all {len(idx)} rays of the workload" if args.cpu_sample == 1 else
                          f"every {args.cpu_sample}th ray of the same 4,194,304 ({len(idx)} rays)") + ", "
                         f"C++ double oracle (g++ -O3 -march=x86-64-v3, one binary for this host and the build "
                         f"host), {threads} std::threads, {cpu_s:.2f}s; its top level is the reference's flat "
                         f"AABB loop over the 128 triangles, where the reference's own benchmark uses embree, so "
                         f"the ratio is an upper estimate",
               "gpu_over_cpu": round(value / cpu_mrays, 1),
               "single_thread": {"value": round(len(idx1) / best / 1e6, 4), "unit": "Mrays/s", "cores": 1,
                                 "sample": f"every 16th ray ({len(idx1)} rays), best of 3",
                                 "checksum_t": round(checksum, 6)}}
        log(f"[bench] cpu {cpu_mrays:.3f} Mrays/s; gpu/cpu {value / cpu_mrays:.1f}x")

    if not args.no_adjoint:
        extra["adjoint"] = adjoint(log)
    if not args.no_pathtrace:
        extra["pathtrace"] = pathtrace(log)
    if args.serve_seconds > 0:
        extra["serving"] = serve(args, log)

    out = {"metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": args.precision, "data": "synthetic",
           "config": {"workload": "C2 incoherent closest-hit: 64 quads on [0,8]^2, 2048^2 fp32 fBm(6)+ridged map, "
                                  "12 pyramid levels, 4194304 rays (2097152 pinhole primaries + 2097152 cosine "
                                  "bounces), both node test, Newton leaf",
                      "rays": n, "triangles": int(wl.triangles.shape[0]), "tile": int(tile.n_leaf)},
           "roofline": roof, "cpu_baseline": cpu,
           "pyramid_build": {"ms": round(build_ms, 4), "bytes": build_bytes,
                             "GB_s": round(build_bytes / (build_ms * 1e-3) / 1e9, 1)},
           **extra}
    print(json.dumps(out), flush=True)


def fp64_report(obj, rays, hits32, wl, n_leaf, log):
    """The same rays through the fp64 build of the kernels, outside the timed
    region: its rate, and the error of the timed fp32 results against it,
    with the bands of the parity tests (grazing |cos| < 1e-2, domain edge
    1e-5, texel-edge creases for the normal)."""
    import torch

    import dmap_amd
    from tests.parity import compare

    n = rays.shape[1]
    opt64 = dmap_amd.TraversalOptions(precision="fp64")
    h64 = obj.alloc_hits(n)
    obj.intersect(rays, options=opt64, out=h64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    obj.intersect(rays, options=opt64, out=h64)
    torch.cuda.synchronize()
    ms64 = (time.perf_counter() - t0) * 1e3

    def host(h, dtype):
        F, I = h["f"].cpu().numpy(), h["i"].cpu().numpy()
        return {"t": F[0].astype(dtype), "st": F[1:3].T.astype(dtype), "ng": F[5:8].T.astype(dtype), "prim": I[0]}

    g32, g64 = host(hits32, np.float32), host(h64, np.float64)
    d = wl.rays[3:6].T.astype(np.float64)
    c = compare(g32, g64, d, wl.triangles, n_leaf)
    # The fp32 search's own root, before the double polish both runs get:
    # the error of the fp32 Newton solve itself.
    hraw = obj.alloc_hits(n)
    obj.intersect(rays, options=dmap_amd.TraversalOptions(polish=False), out=hraw)
    torch.cuda.synchronize()
    graw = host(hraw, np.float32)
    craw = compare(graw, g64, d, wl.triangles, n_leaf)
    braw = np.isfinite(graw["t"]) & np.isfinite(g64["t"])
    t_raw = np.abs(graw["t"][braw] - g64["t"][braw]) / np.abs(g64["t"][braw])
    both = np.isfinite(g32["t"]) & np.isfinite(g64["t"])
    t_rel = np.abs(g32["t"][both] - g64["t"][both]) / np.abs(g64["t"][both])
    out = {"fp64_ms": round(ms64, 3), "fp64_mrays_s": round(n / (ms64 * 1e-3) / 1e6, 1),
           "hitmiss_flips": c["hitmiss_flips"], "flips_in_band": c["flips_in_band"],
           "unexplained": c["unexplained"], "t_rel_max": c["worst_t_rel"],
           "t_rel_p99": float(np.quantile(t_rel, 0.99)) if both.any() else 0.0,
           "st_abs_max": c["worst_st"], "ng_abs_max": c["worst_ng"], "normal_on_crease": c["normal_on_crease"],
           "pre_polish": {"what": "the fp32 Newton root as the search found it (t, st, and ng at that root), "
                                  "without the two fp64 polishing steps, against the fp64 run",
                          "unexplained": craw["unexplained"], "t_rel_max": craw["worst_t_rel"],
                          "t_rel_p99": float(np.quantile(t_raw, 0.99)) if braw.any() else 0.0,
                          "st_abs_max": craw["worst_st"], "ng_abs_max": craw["worst_ng"]}}
    log(f"[fp64] {ms64:.2f} ms ({out['fp64_mrays_s']} Mrays/s); fp32 vs fp64: {out['hitmiss_flips']} hit/miss flips "
        f"({out['flips_in_band']} in the bands), {out['unexplained']} outside the bar, t rel max {out['t_rel_max']:.2e}"
        f" p99 {out['t_rel_p99']:.2e}, st {out['st_abs_max']:.2e}, ng {out['ng_abs_max']:.2e}; before the polish: "
        f"t rel max {out['pre_polish']['t_rel_max']:.2e}, st {out['pre_polish']['st_abs_max']:.2e}, "
        f"ng {out['pre_polish']['ng_abs_max']:.2e}")
    return out


def calibrate_copy(mib, repeats=4):
    """A plain device copy of known size: what FETCH_SIZE and WRITE_SIZE read
    for it scales the counters of the traversal kernel to bytes."""
    import torch
    src = torch.ones(mib << 18, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(repeats):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": mib << 20, "repeats": repeats}), flush=True)


def pathtrace(log):
    """Config 3: wavefront path tracing (1024^2 pixels, 4 spp, 4 bounces, NEE
    with MIS on a 2 x 2 area light, environment 0.2, albedo 0.7) over a
    4096^2 displaced plane; closest-hit and any-hit batches per bounce. The
    device wavefront (dmap_pathtrace) is timed, and the torch loop over the
    same intersector beside it."""
    import torch

    import dmap_amd
    from dmap_amd import pathtrace as PT

    scene, cam_o, cam_d = PT.c3_scene(torch)
    # One untimed call at full size (device memory pools, first launches),
    # then the median of three timed calls.
    PT.render_device(scene, cam_o, cam_d, spp=4, depth=4, seed=0)
    torch.cuda.synchronize()
    st = {}
    times = []
    dmap_amd.kernel_timing_read()
    dmap_amd.kernel_timing(True)
    for _ in range(3):
        t0 = time.perf_counter()
        img = PT.render_device(scene, cam_o, cam_d, spp=4, depth=4, seed=0, stats=st)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    dmap_amd.kernel_timing(False)
    trace_ms, launches = dmap_amd.kernel_timing_read()
    trace_ms /= 3
    launches //= 3
    rays = st["closest_rays"] + st["shadow_rays"]
    t0 = time.perf_counter()
    img_t = PT.render(torch, scene, cam_o, cam_d, spp=4, depth=4, seed=0)
    torch.cuda.synchronize()
    dt_t = time.perf_counter() - t0
    same = (torch.abs(img - img_t) <= 1e-9 * (1 + torch.abs(img_t))).double().mean().item()
    log(f"[pathtrace] 1024^2 x 4 spp, 4 bounces: device wavefront {dt * 1e3:.1f} ms (median of "
        f"{', '.join(f'{t * 1e3:.1f}' for t in times)}; {trace_ms:.1f} ms in {launches} traversal launches, which "
        f"overlap), {st['closest_rays']} closest + {st['shadow_rays']} shadow rays, "
        f"{rays / dt / 1e6:.1f} Mrays/s, mean radiance {img.mean().item():.4f}; torch loop {dt_t * 1e3:.1f} ms, "
        f"same pixels {same:.4f}")
    return {"workload": "1024^2 pixels x 4 spp, 4 bounces, 4096^2 fBm+ridged on 64 quads, NEE+MIS rect light 10, "
                        "env 0.2, albedo 0.7",
            "ms": round(dt * 1e3, 2), "ms_runs": [round(t * 1e3, 2) for t in times],
            "traversal_kernel_ms": round(trace_ms, 2), "traversal_launches": launches,
            "closest_rays": st["closest_rays"], "shadow_rays": st["shadow_rays"],
            "mrays_s": round(rays / dt / 1e6, 2), "mean_radiance": round(img.mean().item(), 5),
            "torch_loop_ms": round(dt_t * 1e3, 2), "torch_loop_same_pixels": round(same, 5)}


def adjoint(log, repeats=10):
    """The differentiation workload (1024^2 fBm at amplitude 0.05 on one
    quad, 512^2 primaries, L2 loss on depth and normal against blob-perturbed
    targets): the forward closest hit and the adjoint pass (which recomputes
    the hits, then scatters dL/d(texel)), timed with HIP events."""
    import torch

    import dmap_amd
    from dmap_amd import workload as W

    wl = W.c4_adjoint()
    tile = dmap_amd.DisplacementTile(wl.texels, wl.strength)
    obj = dmap_amd.ObjectIntersector(tile, wl.triangles)
    tgt_obj = dmap_amd.ObjectIntersector(dmap_amd.DisplacementTile(wl.target_texels, wl.strength), wl.triangles)
    rays = torch.from_numpy(wl.rays).cuda()
    n = rays.shape[1]
    tgt = tgt_obj.intersect(rays)
    res = obj.intersect(rays)
    mask = torch.isfinite(res["t"]) & torch.isfinite(tgt["t"])
    lt = torch.where(mask, 2 * (res["t"] - tgt["t"]), torch.zeros_like(res["t"])).contiguous()
    ln = torch.where(mask[None], 2 * (res["ng"] - tgt["ng"]), torch.zeros_like(res["ng"])).contiguous()
    grad = torch.zeros((wl.texels.shape[0],) * 2, dtype=torch.float64, device="cuda")
    hits = obj.alloc_hits(n)
    stream = torch.cuda.current_stream()
    obj.adjoint(rays, lt, ln, grad=grad)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(stream)
    for _ in range(repeats):
        obj.intersect(rays, out=hits)
    ev[1].record(stream)
    for _ in range(repeats):
        grad.zero_()
        obj.adjoint(rays, lt, ln, grad=grad)
    ev[2].record(stream)
    torch.cuda.synchronize()
    fwd_ms = ev[0].elapsed_time(ev[1]) / repeats
    adj_ms = ev[1].elapsed_time(ev[2]) / repeats
    log(f"[adjoint] forward {fwd_ms:.3f} ms, adjoint (with recomputed hits) {adj_ms:.3f} ms over {n} rays, "
        f"{int(mask.sum())} contributing")
    return {"workload": "1 quad, 1024^2 fBm(4) amp 0.05, 512x512 primaries, L2 on depth+normal vs 16-blob targets",
            "rays": n, "contributing_rays": int(mask.sum()), "forward_ms": round(fwd_ms, 4),
            "adjoint_ms": round(adj_ms, 4), "adjoint_mrays_s": round(n / (adj_ms * 1e-3) / 1e6, 2)}


def serve(args, log):
    """Continuous batching over 8 resident scenes: the plain kernel's rate on
    the same rays, the server's saturated rate, then Poisson arrivals at a
    fixed fraction of the plain rate with p50/p99 latency."""
    import torch

    import dmap_amd
    from dmap_amd import serving as SV

    t0 = time.time()
    scenes = SV.make_scenes()
    reqs = SV.make_requests(scenes, 256, seed=17)
    mean_rays = float(np.mean([r.shape[1] for _, r in reqs]))
    log(f"[serve] 8 scenes {SV.SCENE_SIZES}, 256 distinct requests (mean {mean_rays:.0f} rays) "
        f"in {time.time() - t0:.1f}s")
    plain = SV.plain_rate(scenes, reqs)
    log(f"[serve] plain batched kernel on the same rays: {plain:.1f} Mrays/s")
    srv = dmap_amd.Server(slots=args.serve_slots, step_rays=args.serve_step, hit_pool_bytes=8 << 30)
    # Accuracy: the first 16 requests through the server against the
    # single-request entry point.
    same = 0
    for q in range(16):
        k, rays = reqs[q]
        srv.submit(scenes[k].obj, rays, q)
    got = 0
    while got < 16:
        for tag, _ in srv.poll():
            k, rays = reqs[tag]
            a = srv.hits(tag, rays.shape[1])
            b = scenes[k].obj.intersect(rays)
            same += int(all(torch.equal(a[f], b[f]) for f in ("t", "st", "bary", "ng", "prim", "texel")))
            srv.release(tag)
            got += 1
    sat = SV.summarize(SV.drive(srv, scenes, reqs, None, 0, seed=1))
    log(f"[serve] saturated: {sat}")
    rate = args.serve_load * plain * 1e6 / mean_rays
    fixed = SV.summarize(SV.drive(srv, scenes, reqs, rate, args.serve_seconds, seed=2, log=log))
    log(f"[serve] fixed rate {rate:.0f} req/s: {fixed}")
    del srv
    return {"workload": "8 resident scenes (512^2, 1024^2, 2048^2 x3, 4096^2 x2, 8192^2 fBm+ridged, 64 quads each), "
                        "requests log-uniform in [1k, 1M] rays, 70% coherent primaries / 30% bounces, "
                        "Poisson arrivals",
            "slots": args.serve_slots, "step_rays": args.serve_step,
            "plain_mrays_s": round(plain, 2),
            "saturated": sat, "saturated_over_plain": round(sat["sustained_mrays_s"] / plain, 3),
            "fixed_rate": {"arrivals_per_s": round(rate, 1), "load_vs_plain": args.serve_load, **fixed},
            "fixed_over_plain": round(fixed["sustained_mrays_s"] / plain, 3),
            "identical_to_single_request": f"{same}/16"}


if __name__ == "__main__":
    main()
