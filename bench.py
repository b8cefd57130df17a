"""Benchmark driver: track-steps/sec of the batched heterogeneous transport.

Workload (synthetic, built by vecgeom_amd.workloads): 8 solids per shape
family with log-uniform parameters, per-solid track counts log-uniform in
[1k, 4M] rescaled to --tracks (default 40M), points uniform in 1.3x each
solid's bounding box, isotropic directions, seed 42.  One bench step = one
launch advancing every track by --track-steps (default 50) track-steps.
Inputs are resident in HBM when timing starts.

Runs, in one process on identical initial states (--only picks a subset):
  fp64     plain FP64 step (safety, then boundary distance), bit-identical to
           the CPU oracle
  mixed    a guarded float safety settles the steps it covers; boundary
           distances from the mixed kernels (float selects the surface, FP64
           evaluates it); the FP64 decisions
  learned  learned safety bound on the tessellated solids (networks trained on
           the GPU first, on exact samples), exact fallback; mixed elsewhere
  tuned    each family in the mode the autotuner measured fastest among FP64,
           mixed, distance-first, culled (all with the FP64 decisions) and
           learned
`value` is the last run's whole-job track-steps/s (tuned by default, named in
config.run); the others are reported beside it (mixed_learned_value is the
learned run), with the number of tracks whose final state differs from the
FP64 run and the audited safety violations of learned bounds.  Besides the
transport: the config-1 box queries with their HBM roofline (config1), the
tube/cone FP64-vs-mixed comparison (precision) and the barrel navigation.
The CPU baseline is the FP64 oracle on the host cores over a bounded sample of
the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# The batched transport can run its shape families on four streams of their
# own next to the context stream and torch's; with HIP's default of four
# hardware queues per process two of them would share one and their kernels
# serialise.  Set before HIP initialises.
os.environ["GPU_MAX_HW_QUEUES"] = "8"

FP64_PEAK_TFLOPS = 78.6
FLOPS_FIXTURE = os.path.join(ROOT, "vecgeom_amd", "data", "track_step_flops.json")
TRAFFIC_FIXTURE = os.path.join(ROOT, "vecgeom_amd", "data", "hbm_traffic.json")
NAV_FLOPS_FIXTURE = os.path.join(ROOT, "vecgeom_amd", "data", "nav_step_flops.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def available_families():
    from vecgeom_amd import workloads
    import vecgeom_amd as vg

    L = vg.lib()
    fams = []
    for f in workloads.FAMILIES:
        if hasattr(L, f"vgm_{f}_create"):
            fams.append(f)
    return tuple(fams)


def copy_probe(torch):
    """A plain copy of known size: 2^27 doubles (1 GiB) read once and written
    once by an elementwise kernel, five times.  Under rocprofv3 --pmc its
    FETCH_SIZE and WRITE_SIZE per dispatch give the counters' scale factor."""
    n = 1 << 27
    x = torch.rand(n, dtype=torch.float64, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(5):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 5}))


RUN_NOTES = {
    "fp64": "plain FP64 step (safety, then boundary distance)",
    "mixed": "mixed precision: guarded float safety, float-selected FP64 boundary distances (FP64 decisions)",
    "learned": "mixed precision with the learned safety bound on the tessellated solids (FP64 decisions)",
    "tuned": "each family in the mode the autotuner timed fastest among fp64, mixed, distance-first, culled and "
             "learned (all take the FP64 decisions)",
}

# Algorithmic FP64 work per track-step by step order.  The mixed step takes
# the distance-first step's decisions; its FP64-equivalent work is that
# step's (the float pre-filtering is not counted).
FLOP_TABLES = {"fp64": "per_family", "distance-first": "per_family_distance_first", "culled": "per_family_culled",
               "mixed": "per_family_distance_first"}


def family_roofline(fams_ms, counts_by_family, S, modes=None, streams=1, weights=None, launch_ms=None):
    """Roofline of the dominant family kernel of a timed transport: the family
    with the largest standalone time (`weights`, from the autotuner; else the
    longest launch).  Algorithmic FP64 flops per launch (flop fixture for
    the family's step order x its tracks x track-steps) over its average
    launch time, from events on its stream.  With the families back to back
    on one stream (streams == 1) the events bracket each kernel alone; run
    concurrently, a kernel's interval also holds the time it shares the
    machine with the others, and `aggregate` prices all families' flops
    against the launch time."""
    if not os.path.exists(FLOPS_FIXTURE) or not fams_ms:
        return None
    fx = json.load(open(FLOPS_FIXTURE))
    w = {k: v for k, v in (weights or {}).items() if k in fams_ms}
    fam = max(w, key=w.get) if w else max(fams_ms, key=fams_ms.get)
    modes = modes or {}
    mode = modes.get(fam, "fp64")
    if mode not in FLOP_TABLES:
        return None  # float or matrix-core work: not priced against the FP64 peak
    per_step = float(fx[FLOP_TABLES[mode]][fam])
    flops = per_step * counts_by_family[fam] * S
    t = fams_ms[fam] * 1e-3
    achieved = flops / t / 1e12
    out = {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
           "kernel": f"transport_family_kernel<{fam}, {mode}>",
           "kernel_ms": fams_ms[fam], "flops_per_track_step": per_step, "tracks": counts_by_family[fam],
           "track_steps": S,
           "note": "FP64 vector peak; flops counted on the step the kernel runs (+ - * / sqrt); kernel time from "
                   "events on its stream" + (", the families back to back on one stream" if streams == 1 else
                                             ", while the other families run concurrently")}
    if launch_ms and all(modes.get(f, "fp64") in FLOP_TABLES for f in counts_by_family):
        total = sum(float(fx[FLOP_TABLES[modes.get(f, "fp64")]][f]) * c * S for f, c in counts_by_family.items())
        out["aggregate"] = {"achieved": total / (launch_ms * 1e-3) / 1e12, "unit": "TFLOP/s",
                            "frac": total / (launch_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "flops": total,
                            "launch_ms": launch_ms, "note": "all families' flops over the whole launch"}
    if os.path.exists(TRAFFIC_FIXTURE):
        # HBM bytes per launch of this family kernel in this mode at this
        # workload, from FETCH_SIZE and WRITE_SIZE under rocprofv3 (scaled by a
        # plain copy's reading of the same counters)
        for tr in json.load(open(TRAFFIC_FIXTURE)).get("kernels", []):
            if (tr["family"], tr["mode"], tr["tracks"], tr["track_steps"]) == (fam, mode, out["tracks"], S):
                out["traffic"] = tr["bytes_per_launch"]
                out["traffic_source"] = tr["source"]
                out["algorithmic_bytes"] = tr["algorithmic_bytes"]
    return out


HBM_PEAK_GBS = 8000.0
# algorithmic bytes per query of the SOA3D box kernels (inputs read once,
# output written once; DistanceToIn/Out read the step_max array too)
C1_BYTES = {"Contains": 25, "Inside": 28, "DistanceToIn": 64, "DistanceToOut": 64, "SafetyToIn": 32,
            "SafetyToOut": 32}


def run_config1(args, ctx, dev, torch, stream):
    """Box (10, 20, 30): points uniform in the box scaled x1.5, isotropic
    directions, seed 42, step_max = 1e30; Contains, Inside, DistanceToIn,
    DistanceToOut (on the points Inside == kInside), SafetyToIn, SafetyToOut.
    Each query's kernel time from events on the library's stream (device
    arrays, asynchronous calls, the average of --c1-reps launches), and its
    HBM roofline from the algorithmic bytes per query.  At the configuration's
    1M points a launch lasts microseconds and its fixed cost shows; the same
    queries over --c1-stream points give the streaming rate."""
    import vecgeom_amd as vg
    from vecgeom_amd import sampling

    half = (10.0, 20.0, 30.0)
    box = ctx.box(*half)
    pv = ctx.place(ctx.logical_volume(box))
    out = {}
    for label, n in (("config1", args.c1_points), ("streaming", args.c1_stream)):
        if n <= 0:
            continue
        pts = [torch.from_numpy(a).to(dev) for a in sampling.points_in_box(42, 1, n, half, 1.5)]
        dirs = [torch.from_numpy(a).to(dev) for a in sampling.isotropic_directions(42, 2, n)]
        step = torch.full((n,), 1e30, dtype=torch.float64, device=dev)
        P, D = vg.SOA3D(*pts), vg.SOA3D(*dirs)
        ins = ctx.inside(pv, P)
        sel = torch.nonzero(ins == vg.kInside).flatten()
        Pin = vg.SOA3D(*[pts[k][sel].contiguous() for k in range(3)])
        Din = vg.SOA3D(*[dirs[k][sel].contiguous() for k in range(3)])
        stin = step[sel].contiguous()
        o8 = torch.empty(n, dtype=torch.uint8, device=dev)
        o32 = torch.empty(n, dtype=torch.int32, device=dev)
        o64 = torch.empty(n, dtype=torch.float64, device=dev)
        calls = {
            "Contains": (n, lambda: ctx.contains(pv, P, out=o8, asynchronous=True)),
            "Inside": (n, lambda: ctx.inside(pv, P, out=o32, asynchronous=True)),
            "DistanceToIn": (n, lambda: ctx.distance_to_in(pv, P, D, step, out=o64, asynchronous=True)),
            "DistanceToOut": (len(sel), lambda: ctx.distance_to_out(pv, Pin, Din, stin, out=o64[:len(sel)],
                                                                    asynchronous=True)),
            "SafetyToIn": (n, lambda: ctx.safety_to_in(pv, P, out=o64, asynchronous=True)),
            "SafetyToOut": (n, lambda: ctx.safety_to_out(pv, P, out=o64, asynchronous=True)),
        }
        res = {}
        for q, (m, fn) in calls.items():
            fn()
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for _ in range(args.c1_reps):
                fn()
            b.record(stream)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b) / args.c1_reps
            gbs = C1_BYTES[q] * m / (ms * 1e-3) / 1e9
            res[q] = {"queries": m, "kernel_us": ms * 1e3, "gq_s": m / ms / 1e6,
                      "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": gbs / HBM_PEAK_GBS, "bytes_per_query": C1_BYTES[q]}}
        out[label] = {"points": n, "inside_points": len(sel), "reps": args.c1_reps, "queries": res}
        log(f"[bench] box {label} ({n} points): " + ", ".join(
            f"{q} {v['kernel_us']:.1f} us {v['roofline']['frac']:.2f}" for q, v in res.items()))
        del pts, dirs, step, P, D, Pin, Din, stin, o8, o32, o64
        torch.cuda.empty_cache()
    if not args.no_cpu:
        import oracle

        geo = oracle.Geometry()
        pvo = geo.place(geo.logical_volume(geo.box(*half)))
        n = args.c1_points
        pts = sampling.points_in_box(42, 1, n, half, 1.5)
        dirs = sampling.isotropic_directions(42, 2, n)
        nth = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        t = time.perf_counter()
        for q in ("Contains", "Inside", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"):
            geo.query(pvo, q, pts, dirs if "Distance" in q else None, nthreads=nth)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": 6 * n / dt, "unit": "queries/s", "cores": nth, "kind": "port",
                               "sample": f"the six queries over the {n} points (DistanceToOut on all of them), "
                                         "FP64 oracle, OpenMP"}
    return out


def run_precision(args, ctx, dev, torch, stream):
    """Tube and cone SOA3D queries, FP64 vs mixed precision: 8M points uniform
    in 1.2x the bounding box and 8M within +-1e-6 mm of the surface (entry or
    exit points of random rays from the FP64 kernels, jittered along the
    ray), isotropic directions.  Throughput of each query in both precisions
    and the mixed results' error against FP64."""
    import vecgeom_amd as vg
    from vecgeom_amd import sampling

    shapes = {"tube(5,20,30,0,3pi/2)": lambda: ctx.tube(5.0, 20.0, 30.0, 0.0, 1.5 * np.pi),
              "cone(0,10,5,25,40)": lambda: ctx.cone(0.0, 10.0, 5.0, 25.0, 40.0, 0.0, 2 * np.pi)}
    m = args.c2_points // 2
    out = {}
    for name, make in shapes.items():
        s = make()
        pv = ctx.place(ctx.logical_volume(s))
        mn, mx = ctx.extent(s)
        half, cen = 0.5 * (mx - mn), 0.5 * (mx + mn)
        pts = [torch.from_numpy(a).to(dev) for a in sampling.points_in_box(2, 1, m, half, 1.2, cen)]
        dirs = [torch.from_numpy(a).to(dev) for a in sampling.isotropic_directions(2, 2, m)]
        P, D = vg.SOA3D(*pts), vg.SOA3D(*dirs)
        din, dout = ctx.distance_to_in(pv, P, D), ctx.distance_to_out(pv, P, D)
        t = torch.where((din > 0) & (din < 1e30), din, torch.where((dout > 0) & (dout < 1e30), dout,
                                                                   torch.zeros_like(din)))
        jit = (2.0 * torch.from_numpy(sampling.uniform(2, 9, m)).to(dev) - 1.0) * 1e-6
        band = [pts[k] + (t + jit) * dirs[k] for k in range(3)]
        dirs2 = [torch.from_numpy(a).to(dev) for a in sampling.isotropic_directions(2, 3, m)]
        P = vg.SOA3D(*[torch.cat([pts[k], band[k]]) for k in range(3)])
        D = vg.SOA3D(*[torch.cat([dirs[k], dirs2[k]]) for k in range(3)])
        n = 2 * m
        scale = float(np.linalg.norm(mx - mn))
        res = {}
        for q in ("Inside", "Contains", "DistanceToIn", "DistanceToOut", "SafetyToIn", "SafetyToOut"):
            fn = {"Inside": lambda mx_: ctx.inside(pv, P, mixed=mx_),
                  "Contains": lambda mx_: ctx.contains(pv, P, mixed=mx_),
                  "DistanceToIn": lambda mx_: ctx.distance_to_in(pv, P, D, mixed=mx_),
                  "DistanceToOut": lambda mx_: ctx.distance_to_out(pv, P, D, mixed=mx_),
                  "SafetyToIn": lambda mx_: ctx.safety_to_in(pv, P, mixed=mx_),
                  "SafetyToOut": lambda mx_: ctx.safety_to_out(pv, P, mixed=mx_)}[q]
            r = {}
            outs = {}
            for mx_ in (False, True):
                fn(mx_)
                torch.cuda.synchronize()
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                for _ in range(3):
                    o = fn(mx_)
                b.record(stream)
                torch.cuda.synchronize()
                ms = a.elapsed_time(b) / 3
                r["mixed_gq_s" if mx_ else "fp64_gq_s"] = n / ms / 1e6
                outs[mx_] = o
            ref, got = outs[False], outs[True]
            if q in ("Inside", "Contains"):
                r["classification_mismatches"] = int((ref != got).sum().item())
            else:
                big_r, big_g = ref >= 1e300, got >= 1e300
                fin = ~big_r & ~big_g
                err = (ref[fin] - got[fin]).abs()
                r["inf_disagree"] = int((big_r != big_g).sum().item())
                r["max_abs_err"] = float(err.max().item()) if err.numel() else 0.0
                r["mean_abs_err"] = float(err.mean().item()) if err.numel() else 0.0
                r["exact_fraction"] = float((err == 0).double().mean().item()) if err.numel() else 1.0
            res[q] = r
        out[name] = {"points": n, "scale_mm": scale, "queries": res}
        log(f"[bench] precision {name}: " + ", ".join(
            f"{q} {v['fp64_gq_s']:.1f}/{v['mixed_gq_s']:.1f} Gq/s" for q, v in res.items()))
    return out


def run_navigation(args, ctx, dev, torch):
    """Barrel navigation: straight-line transport of located tracks through
    20 tube layers of box modules, nsteps of ComputeStepAndPropagatedState per
    track (tracks that leave the world stop).  Value: track-steps/s."""
    import vecgeom_amd as vg
    from vecgeom_amd import workloads

    t0 = time.perf_counter()
    world = workloads.build_barrel(ctx)
    nav = vg.Navigation(ctx, world)
    pts, dirs = workloads.barrel_tracks(args.nav_tracks, seed=42)
    n = args.nav_tracks
    P = [torch.from_numpy(a).to(dev) for a in pts]
    D = [torch.from_numpy(a).to(dev) for a in dirs]
    st0 = torch.zeros(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    nav.locate(*P, st0)
    log(f"[bench] navigation: {nav.nnodes} states, {n} tracks located ({time.perf_counter() - t0:.1f}s)")
    stream = torch.cuda.ExternalStream(ctx.stream, device=dev)

    # the same buffers for every run, so that a captured graph replays
    bx, by, bz = (a.clone() for a in P)
    bst = st0.clone()
    blast = torch.zeros(n, dtype=torch.int32, device=dev)
    bbnd = torch.zeros(n, dtype=torch.uint8, device=dev)
    bsteps = torch.zeros(n, dtype=torch.int32, device=dev)

    def one(**mode):
        for b_, a_ in zip((bx, by, bz, bst), (*P, st0)):
            b_.copy_(a_)
        for b_ in (blast, bbnd, bsteps):
            b_.zero_()
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        a.record(stream)
        nav.transport(bx, by, bz, *D, bst, blast, bbnd, args.nav_steps, bsteps, asynchronous=True, **mode)
        b.record(stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t, a.elapsed_time(b) * 1e-3, int(bsteps.sum().item())

    one()
    wall, kern, total = one()
    out = {"workload": f"barrel: 20 tube layers, {sum(l['n'] for l in workloads.barrel_layers())} box modules, "
                       f"{n} tracks x up to {args.nav_steps} steps", "value": total / kern, "unit": "track-steps/s",
           "track_steps": total, "kernel_ms": kern * 1e3, "wall_ms": wall * 1e3,
           "kernel": "nav_transport_kernel (persistent grid over a work queue of tracks)"}
    log(f"[bench] navigation: {total / kern / 1e9:.3f} G track-steps/s ({total} steps, kernel {kern * 1e3:.1f} ms)")
    if os.path.exists(NAV_FLOPS_FIXTURE):
        fx = json.load(open(NAV_FLOPS_FIXTURE))
        ach = fx["flops_per_step"] * total / kern / 1e12
        out["roofline"] = {"bound": "fp64-valu", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": ach / FP64_PEAK_TFLOPS, "kernel": "nav_transport_kernel",
                           "flops_per_step": fx["flops_per_step"], "steps": total,
                           "note": "algorithmic FP64 flops per step counted on the host by the same navigator "
                                   "code (vecgeom_amd/data/nav_step_flops.json) x steps taken / kernel time; the "
                                   "walks are latency-bound (profiles/navigation_barrel_4M.md)"}
        log(f"[bench] navigation roofline: {ach:.3f} TF/s = {ach / FP64_PEAK_TFLOPS:.4f} of the FP64 peak")
    # the step loop as one-step kernels: launched one by one, and captured in
    # a HIP graph (the first call captures it, the timed call replays it)
    one(step_kernels=True)
    w_k, k_k, t_k = one(step_kernels=True)
    t0 = time.perf_counter()
    one(graph=True)
    capture_s = time.perf_counter() - t0
    w_g, k_g, t_g = one(graph=True)
    assert t_k == total and t_g == total
    out["step_kernels"] = {"value": t_k / k_k, "kernel_ms": k_k * 1e3, "wall_ms": w_k * 1e3,
                           "launches": args.nav_steps}
    out["graph"] = {"value": t_g / k_g, "kernel_ms": k_g * 1e3, "wall_ms": w_g * 1e3,
                    "capture_and_first_run_s": capture_s, "launch_overhead_saved_ms": (w_k - w_g) * 1e3,
                    "note": f"the same {args.nav_steps} one-step kernels as step_kernels, replayed from a captured "
                            "HIP graph; overhead saved = wall time launched one by one - wall time of the replay"}
    log(f"[bench] navigation step loop: one-step kernels {t_k / k_k / 1e9:.3f} G/s (wall {w_k * 1e3:.1f} ms), "
        f"graph {t_g / k_g / 1e9:.3f} G/s (wall {w_g * 1e3:.1f} ms, capture {capture_s:.2f} s)")
    # the same loops on a few tracks, where each launch is short and its
    # fixed cost shows
    m = min(n, 20_000)
    sx, sy, sz = (a[:m].clone() for a in P)
    sd = [a[:m].contiguous() for a in D]
    sst = st0[:m].clone()
    sl = torch.zeros(m, dtype=torch.int32, device=dev)
    sb = torch.zeros(m, dtype=torch.uint8, device=dev)
    ss = torch.zeros(m, dtype=torch.int32, device=dev)

    def small(**mode):
        for b_, a_ in zip((sx, sy, sz, sst), (P[0][:m], P[1][:m], P[2][:m], st0[:m])):
            b_.copy_(a_)
        for b_ in (sl, sb, ss):
            b_.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        nav.transport(sx, sy, sz, *sd, sst, sl, sb, args.nav_steps, ss, asynchronous=True, **mode)
        torch.cuda.synchronize()
        return time.perf_counter() - t

    small(step_kernels=True)
    small(graph=True)
    wk = min(small(step_kernels=True) for _ in range(3))
    wg = min(small(graph=True) for _ in range(3))
    out["graph_small"] = {"tracks": m, "launches": args.nav_steps, "step_kernels_wall_ms": wk * 1e3,
                          "graph_wall_ms": wg * 1e3, "launch_overhead_saved_ms": (wk - wg) * 1e3}
    log(f"[bench] navigation step loop, {m} tracks: one-step kernels {wk * 1e3:.2f} ms, graph {wg * 1e3:.2f} ms")

    # the same tracks under a 10 mm physics step: the safety first, a boundary
    # distance only when the safety does not cover the step
    lim = torch.full((n,), 10.0, dtype=torch.float64, device=dev)

    def limited(reuse):
        x, y, z = (a.clone() for a in P)
        st = st0.clone()
        last = torch.zeros(n, dtype=torch.int32, device=dev)
        bnd = torch.zeros(n, dtype=torch.uint8, device=dev)
        steps = torch.zeros(n, dtype=torch.int32, device=dev)
        geo_steps = torch.zeros(n, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        nav.transport_limited(x, y, z, *D, lim, st, last, bnd, args.nav_steps, steps, geo_steps, reuse_safety=reuse)
        b.record(stream)
        torch.cuda.synchronize()
        return a.elapsed_time(b) * 1e-3, int(steps.sum().item()), int(geo_steps.sum().item())

    out["limited_10mm"] = {}
    for reuse in (False, True):
        limited(reuse)
        k2, s2, g2 = limited(reuse)
        key = "carried_safety" if reuse else "fresh_safety"
        out["limited_10mm"][key] = {"value": s2 / k2, "unit": "track-steps/s", "track_steps": s2,
                                    "geometry_steps": g2, "kernel_ms": k2 * 1e3}
        log(f"[bench] navigation, 10 mm physics step, {key}: {s2 / k2 / 1e9:.3f} G track-steps/s "
            f"({g2 / max(s2, 1):.2f} of the steps computed a boundary distance)")
    if not args.no_cpu:
        import oracle

        geo = oracle.Geometry()
        geo.nav_close(workloads.build_barrel(geo), use_bvh=True)
        m = min(n, 100_000)
        cp = [a[:m].copy() for a in pts]
        st = geo.nav_locate(cp)
        last, bnd, steps = np.zeros(m, np.uint32), np.zeros(m, np.uint8), np.zeros(m, np.int32)
        nth = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        t = time.perf_counter()
        geo.nav_transport(*cp, *[a[:m].copy() for a in dirs], st, last, bnd, args.nav_steps, steps, nthreads=nth)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": float(steps.sum()) / dt, "unit": "track-steps/s", "cores": nth,
                               "kind": "port", "sample": f"{m} tracks x up to {args.nav_steps} steps, BVH oracle"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed bench steps (launches)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tracks", type=int, default=40_000_000)
    ap.add_argument("--track-steps", type=int, default=50, help="track-steps per track per launch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--cpu-tracks", type=int, default=400_000)
    ap.add_argument("--nav-tracks", type=int, default=4_000_000, help="navigation benchmark tracks (0: skip)")
    ap.add_argument("--nav-steps", type=int, default=200)
    ap.add_argument("--c1-points", type=int, default=1_000_000, help="box points of the config-1 queries (0: skip)")
    ap.add_argument("--c1-stream", type=int, default=32_000_000,
                    help="box points of the streaming-size run of the same queries (0: skip)")
    ap.add_argument("--c1-reps", type=int, default=20)
    ap.add_argument("--c2-points", type=int, default=16_000_000,
                    help="points per shape for the tube/cone precision comparison (0: skip)")
    ap.add_argument("--launch", default=None,
                    help="block,grid,tile[,streams] launch shape instead of autotuning (streams 1 or 4)")
    ap.add_argument("--train-samples", type=int, default=50_000_000,
                    help="exact safety samples for the learned-safety networks, over all tessellated solids")
    ap.add_argument("--train-epochs", type=int, default=4)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--modes", default=None,
                    help="tuned run without the autotuner: family=mode,... (modes fp64, mixed, distance-first, "
                         "learned, culled); with --launch")
    ap.add_argument("--only", default=None,
                    help="comma list out of fp64,mixed,learned,tuned (transport runs; the headline is the last), "
                         "config1, precision, navigation (default: all)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain copy kernel of 1 GiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; runs on several GPUs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        copy_probe(torch)
        return

    import vecgeom_amd as vg
    from vecgeom_amd import workloads

    dev = torch.device("cuda:0")
    ctx = vg.Context(0)
    # Launches go to the library's own stream; events are recorded there too.
    stream = torch.cuda.ExternalStream(ctx.stream, device=dev)
    fams = available_families()
    only = set(args.only.split(",")) if args.only else {"fp64", "mixed", "learned", "tuned", "config1", "precision",
                                                        "navigation"}
    config1 = None
    if "config1" in only and args.c1_points > 0:
        config1 = run_config1(args, ctx, dev, torch, stream)
    if not only & {"fp64", "mixed", "learned", "tuned"}:
        precision = run_precision(args, ctx, dev, torch, stream) if "precision" in only and args.c2_points else None
        navigation = run_navigation(args, ctx, dev, torch) if "navigation" in only and args.nav_tracks else None
        print(json.dumps({"metric": "track-steps/sec on 1 MI355X (mixed+learned vs plain FP64), 0 safety violations",
                          "value": None, "unit": "track-steps/s", "n_gpus": 1, "steps": args.steps,
                          "warmup": args.warmup, "note": "no transport run requested (--only)",
                          "config1": config1, "precision": precision, "navigation": navigation}))
        return
    t0 = time.perf_counter()
    specs, counts = workloads.batched_solids(args.tracks, families=fams, per_family=8, seed=42)
    shapes, offsets, steps, arr = workloads.build(ctx, specs, counts, seed=42)
    n = int(offsets[-1])
    log(f"[bench] {len(specs)} solids ({', '.join(fams)}), {n} tracks, built in {time.perf_counter() - t0:.1f}s")
    init = [torch.from_numpy(a).to(dev) for a in arr]
    del arr
    side0 = torch.zeros(n, dtype=torch.int8, device=dev)
    batch = vg.Batch(ctx, shapes, offsets, steps)
    torch.cuda.synchronize()
    batch.classify(init[0], init[1], init[2], side0)
    S = args.track_steps

    def run(mixed, learned=False, tuned=False):
        state = [t.clone() for t in init]
        side = side0.clone()
        cr = torch.zeros(n, dtype=torch.int32, device=dev)
        tag = "tuned" if tuned else (("mixed" if mixed else "fp64") + ("+learned" if learned else ""))
        kw = {"mixed": mixed, "learned": learned} if not tuned else {"tuned": True}
        torch.cuda.synchronize()
        if args.launch:
            lv = [int(v) for v in args.launch.split(",")]
            batch.set_launch(*lv[:3])
            if len(lv) > 3:
                batch.set_streams(lv[3])
            if tuned and args.modes:
                batch.set_modes(dict(kv.split("=") for kv in args.modes.split(",")))
        elif not args.no_autotune:
            log(f"[bench] autotune {tag} ...")
            t = time.perf_counter()
            shape = batch.autotune(*state, side, nsteps=min(S, 10), mixed=mixed or tuned,
                                   learned=learned or tuned, distance_first=tuned, cull=tuned)
            log(f"[bench] autotune {tag}: block/grid/tile = {shape} ({time.perf_counter() - t:.1f}s)")
        step0 = 0
        for _ in range(args.warmup):
            batch.transport(*state, side, cr, S, step0=step0, seed=7, asynchronous=True, **kw)
            step0 += S
        torch.cuda.synchronize()
        evs, fms = [], []
        t = time.perf_counter()
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            batch.transport(*state, side, cr, S, step0=step0, seed=7, asynchronous=True, timing=True, **kw)
            b.record(stream)
            evs.append((a, b))
            fms.append(batch.family_ms())  # waits for this launch's family kernels
            step0 += S
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        kern = float(np.mean([a.elapsed_time(b) for a, b in evs])) * 1e-3
        fam_ms = {f: float(np.mean([m[f] for m in fms])) for f in fms[0]}
        extra = {}
        if learned and not tuned:
            # the same launches captured in a HIP graph (the first captures it,
            # the rest replay it, step0 advancing on the device)
            gs = [x.clone() for x in state]
            gside, gcr = side.clone(), cr.clone()
            g0 = step0
            batch.transport(*gs, gside, gcr, S, step0=g0, seed=7, graph=True, **kw)
            g0 += S
            torch.cuda.synchronize()
            t = time.perf_counter()
            for _ in range(args.steps):
                batch.transport(*gs, gside, gcr, S, step0=g0, seed=7, graph=True, asynchronous=True, **kw)
                g0 += S
            torch.cuda.synchronize()
            wg = time.perf_counter() - t
            extra["graph"] = {"ms_per_launch": wg / args.steps * 1e3, "ms_per_launch_uncaptured": wall / args.steps * 1e3,
                              "launch_overhead_saved_ms": (wall - wg) / args.steps * 1e3,
                              "note": "the learned families run two kernels per track-step; captured once, replayed"}
            log(f"[bench] {tag}: graph-captured launches {wg / args.steps * 1e3:.2f} ms vs {wall / args.steps * 1e3:.2f} ms")
        if mixed and not learned:
            st2 = [x.clone() for x in init]
            s2, c2 = side0.clone(), torch.zeros(n, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            extra["fp64_steps"] = batch.transport(*st2, s2, c2, S, step0=0, seed=7, mixed=True, count_fp64=True)
        if learned or tuned:
            # audit: the first launch again, every served bound checked against the exact safety
            st2 = [x.clone() for x in init]
            s2, c2 = side0.clone(), torch.zeros(n, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            batch.transport(*st2, s2, c2, S, step0=0, seed=7, audit=True, **kw)
            extra["audit"] = batch.learned_stats()
        if tuned and batch.streams() > 1:
            # each family kernel's own duration: one more launch of the same
            # modes and launch shape, the families back to back on one stream
            # (run concurrently, a kernel's interval also holds the time it
            # shares the machine, and that share changes from run to run)
            st2 = [x.clone() for x in state]
            s2, c2 = side.clone(), cr.clone()
            nstreams = batch.streams()
            batch.set_streams(1)
            torch.cuda.synchronize()
            batch.transport(*st2, s2, c2, S, step0=step0, seed=7, timing=True, **kw)
            extra["family_ms_one_stream"] = batch.family_ms()
            batch.set_streams(nstreams)
            log(f"[bench] {tag} family kernel ms, one stream: " + ", ".join(
                f"{k} {v:.1f}" for k, v in extra["family_ms_one_stream"].items()))
        if tuned or not args.launch:
            extra["modes"] = batch.modes()
        if not args.launch:  # the autotuner's per-family times (without it the weights are track counts)
            extra["weights"] = batch.weights()
        final = state + [side, cr]
        log(f"[bench] {tag}: {args.steps * n * S / wall / 1e9:.2f} G track-steps/s (kernel {kern * 1e3:.2f} ms/launch)")
        log(f"[bench] {tag} family kernel ms: " + ", ".join(f"{k} {v:.1f}" for k, v in fam_ms.items()))
        return wall, kern, extra, batch.launch() + (batch.streams(),), fam_ms, final

    counts_by_family = {}
    for (kind, _), c in zip(specs, counts):
        counts_by_family[kind] = counts_by_family.get(kind, 0) + int(c)
    res = {}
    if "fp64" in only:
        res["fp64"] = run(False)
    if "mixed" in only:
        res["mixed"] = run(True)
        log(f"[bench] mixed: fraction of steps the float safety did not settle "
            f"{res['mixed'][2]['fp64_steps'] / (n * S):.4f}")
    tess = [i for i, (kind, _) in enumerate(specs) if kind == "tessellated"]
    train = None
    if tess and not args.no_learned and only & {"learned", "tuned"}:
        per = max(8192, args.train_samples // len(tess) // 8192 * 8192)
        reports = []
        t = time.perf_counter()
        for i in tess:
            reports.append(ctx.mlp_train(int(shapes[i]), samples=per, epochs=args.train_epochs))
        train = {"networks": len(tess), "train_samples": per * len(tess), "train_epochs": args.train_epochs,
                 "train_wall_s": time.perf_counter() - t, "train_gpu_ms": sum(r["train_ms"] for r in reports),
                 "train_gflop": sum(r["gflop"] for r in reports),
                 "margin_over_scale": [r["margin"] / r["scale"] for r in reports],
                 "rms_err_over_scale": [r["rms_err"] / r["scale"] for r in reports]}
        log(f"[bench] trained {len(tess)} networks on {per * len(tess)} exact samples in {train['train_wall_s']:.1f}s "
            f"(GPU {train['train_gpu_ms'] / 1e3:.1f}s); margins/scale "
            + ", ".join(f"{v:.4f}" for v in train["margin_over_scale"]))
        if "learned" in only:
            res["learned"] = run(True, learned=True)
    if "tuned" in only:
        res["tuned"] = run(False, tuned=True)

    def rate(key):
        return args.steps * n * S / res[key][0] if key in res else None

    def ndiff(key, mask):
        """Tracks of `mask` whose final state differs from the FP64 run's."""
        if "fp64" not in res or key not in res:
            return None
        d = torch.zeros(n, dtype=torch.bool, device=dev)
        for a, b in zip(res["fp64"][5], res[key][5]):
            d |= (a != b)
        return int((d & mask).sum().item())

    def family_mask(pred):
        m = torch.zeros(n, dtype=torch.bool, device=dev)
        for i, (kind, _) in enumerate(specs):
            if pred(kind):
                m[int(offsets[i]):int(offsets[i + 1])] = True
        return m

    learned = tuned = None
    if "learned" in res:
        au = res["learned"][2]["audit"]
        learned = dict(train or {})
        learned.update({
            "queries": au["queries"], "served": au["served"], "fallbacks": au["fallbacks"],
            "served_fraction": au["served"] / max(1, au["queries"]), "violations": au["violations"],
            "audit_sample": f"first launch ({S} track-steps of all {n} tracks), every served bound checked",
            "tessellated_tracks_differing_from_fp64": ndiff("learned", family_mask(lambda k: k == "tessellated")),
            "value": rate("learned"), "kernel_ms": res["learned"][1] * 1e3, "family_kernel_ms": res["learned"][4],
            "launch": res["learned"][3], "graph": res["learned"][2].get("graph")})
        log(f"[bench] learned: audit {au}, tessellated tracks differing from fp64: "
            f"{learned['tessellated_tracks_differing_from_fp64']}")
    if "tuned" in res:
        modes = res["tuned"][2].get("modes", {})
        tuned = {"value": rate("tuned"), "modes": modes, "kernel_ms": res["tuned"][1] * 1e3,
                 "family_alone_ms": res["tuned"][2].get("weights"),
                 "family_kernel_ms": res["tuned"][4], "launch": res["tuned"][3],
                 "tracks_differing_from_fp64": ndiff("tuned", family_mask(lambda k: True)),
                 "audit": res["tuned"][2].get("audit"), "learned_networks": train}
        log(f"[bench] tuned modes {modes}; tracks differing from fp64: {tuned['tracks_differing_from_fp64']}")
        log("[bench] tuned: each family alone in its mode (autotuner), ms: " + ", ".join(
            f"{k} {v:.1f}" for k, v in (tuned["family_alone_ms"] or {}).items()))
    head = next(k for k in ("tuned", "learned", "mixed", "fp64") if k in res)
    value, wall_v, kern_v = rate(head), res[head][0], res[head][1]
    roofline = family_roofline(res[head][4], counts_by_family, S,
                               res[head][2].get("modes") if head == "tuned" else None, res[head][3][3],
                               res[head][2].get("weights") if not args.launch else None, kern_v * 1e3)
    if roofline:
        roofline["run"] = head
        alone = res[head][2].get("family_ms_one_stream")
        if alone:
            fam = roofline["kernel"].split("<")[1].split(",")[0]
            a_ach = roofline["achieved"] * roofline["kernel_ms"] / alone[fam]
            roofline["alone"] = {"kernel_ms": alone[fam], "achieved": a_ach, "frac": a_ach / roofline["peak"],
                                 "note": "the same kernel's own duration: one more launch of the same modes, the "
                                         "families back to back on one stream, after the timed launches"}
    rate64, ratemx = rate("fp64"), rate("mixed")
    used = set(res[head][2].get("modes", {}).values()) if head == "tuned" else {head}
    dtype = "f64" + ("+f32" if used & {"mixed"} else "") + ("+bf16" if used & {"learned"} else "")
    nfp64 = res["mixed"][2]["fp64_steps"] if "mixed" in res else None
    kern64 = res["fp64"][1] if "fp64" in res else None
    kernmx = res["mixed"][1] if "mixed" in res else None
    launch64 = res["fp64"][3] if "fp64" in res else None
    launchmx = res["mixed"][3] if "mixed" in res else None
    fam64 = res["fp64"][4] if "fp64" in res else None
    fammx = res["mixed"][4] if "mixed" in res else None

    cpu = None
    if not args.no_cpu:
        import oracle

        geo = oracle.Geometry()
        cspecs, ccounts = workloads.batched_solids(args.cpu_tracks, families=fams, per_family=8, seed=42)
        cs, co, cst, carr = workloads.build(geo, cspecs, ccounts, seed=42)
        cn = int(co[-1])
        side = np.zeros(cn, np.int8)
        cr = np.zeros(cn, np.int32)
        nth = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        t = time.perf_counter()
        geo.batch_transport(cs, co, cst, *carr, side, cr, S, step0=0, seed=7, classify=True, nthreads=nth)
        dt = time.perf_counter() - t
        cpu = {"value": cn * S / dt, "unit": "track-steps/s", "cores": nth, "kind": "port",
               "sample": f"{cn} tracks x {S} track-steps, same solids and generator, FP64 oracle, OpenMP"}

    precision = None
    if "precision" in only and args.c2_points > 0:
        precision = run_precision(args, ctx, dev, torch, stream)

    navigation = None
    if "navigation" in only and args.nav_tracks > 0:
        navigation = run_navigation(args, ctx, dev, torch)
    rate_ml = rate("learned")

    line = {
        "metric": "track-steps/sec on 1 MI355X (mixed+learned vs plain FP64), 0 safety violations",
        "value": value,
        "unit": "track-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_v / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"batched heterogeneous solids: {len(specs)} solids ({'/'.join(fams)}), "
                               f"{n} tracks, {S} track-steps per launch",
                   "tracks": n, "track_steps_per_launch": S, "families": list(fams),
                   "run": head, "run_note": RUN_NOTES.get(head)},
        "fp64_value": rate64,
        "mixed_value": ratemx,
        "mixed_learned_value": rate_ml,
        "mixed_learned_over_fp64": rate_ml / rate64 if (rate_ml and rate64) else None,
        "value_over_fp64": value / rate64 if rate64 else None,
        "mixed_float_unsettled_fraction": nfp64 / (n * S) if nfp64 is not None else None,
        "mixed_tracks_differing_from_fp64": ndiff("mixed", family_mask(lambda k: True)),
        "kernel_ms": {"fp64": kern64 * 1e3 if kern64 else None, "mixed": kernmx * 1e3 if kernmx else None,
                      "value": kern_v * 1e3},
        "family_kernel_ms": {"fp64": fam64, "mixed": fammx},
        "launch": {"fp64": launch64, "mixed": launchmx},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned": learned,
        "tuned": tuned,
        "precision": precision,
        "navigation": navigation,
        "config1": config1,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
