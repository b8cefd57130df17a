"""Benchmark driver: track-steps/sec of the batched heterogeneous transport.

Workload (synthetic, built by vecgeom_amd.workloads): 8 solids per shape
family with log-uniform parameters, per-solid track counts log-uniform in
[1k, 4M] rescaled to --tracks (default 40M), points uniform in 1.3x each
solid's bounding box, isotropic directions, seed 42.  One bench step = one
persistent launch advancing every track by --track-steps (default 50)
track-steps.  Inputs are resident in HBM when timing starts.

Both precision paths run in the same process on identical initial states:
plain FP64 (bit-identical to the CPU oracle) and mixed (float filter, FP64
boundary steps).  `value` is the mixed path's whole-job track-steps/s; the
FP64 figure and the ratio are reported beside it.  The CPU baseline is the
FP64 oracle on the host cores over a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
FLOPS_FIXTURE = os.path.join(ROOT, "vecgeom_amd", "data", "track_step_flops.json")


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

    def one(timed):
        x, y, z = (a.clone() for a in P)
        st = st0.clone()
        last = torch.zeros(n, dtype=torch.int32, device=dev)
        bnd = torch.zeros(n, dtype=torch.uint8, device=dev)
        steps = torch.zeros(n, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = time.perf_counter()
        a.record(stream)
        nav.transport(x, y, z, *D, st, last, bnd, args.nav_steps, steps, asynchronous=True)
        b.record(stream)
        torch.cuda.synchronize()
        return time.perf_counter() - t, a.elapsed_time(b) * 1e-3, int(steps.sum().item())

    one(False)
    wall, kern, total = one(True)
    out = {"workload": f"barrel: 20 tube layers, {sum(l['n'] for l in workloads.barrel_layers())} box modules, "
                       f"{n} tracks x up to {args.nav_steps} steps", "value": total / kern, "unit": "track-steps/s",
           "track_steps": total, "kernel_ms": kern * 1e3, "wall_ms": wall * 1e3}
    log(f"[bench] navigation: {total / kern / 1e9:.3f} G track-steps/s ({total} steps, kernel {kern * 1e3:.1f} ms)")
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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; runs on several GPUs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch

    import vecgeom_amd as vg
    from vecgeom_amd import workloads

    dev = torch.device("cuda:0")
    ctx = vg.Context(0)
    # Launches go to the library's own stream; events are recorded there too.
    stream = torch.cuda.ExternalStream(ctx.stream, device=dev)
    fams = available_families()
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

    def run(mixed):
        state = [t.clone() for t in init]
        side = side0.clone()
        cr = torch.zeros(n, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        if not args.no_autotune:
            log(f"[bench] autotune {'mixed' if mixed else 'fp64'} ...")
            t = time.perf_counter()
            shape = batch.autotune(*state, side, nsteps=2, mixed=mixed)
            log(f"[bench] autotune {'mixed' if mixed else 'fp64'}: block/grid/tile = {shape} "
                f"({time.perf_counter() - t:.1f}s)")
        step0 = 0
        for _ in range(args.warmup):
            batch.transport(*state, side, cr, S, step0=step0, seed=7, mixed=mixed, asynchronous=True)
            step0 += S
        torch.cuda.synchronize()
        evs = []
        t = time.perf_counter()
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            batch.transport(*state, side, cr, S, step0=step0, seed=7, mixed=mixed, asynchronous=True)
            b.record(stream)
            evs.append((a, b))
            step0 += S
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        kern = float(np.mean([a.elapsed_time(b) for a, b in evs])) * 1e-3
        nfp64 = 0
        if mixed:
            st2 = [x.clone() for x in init]
            s2, c2 = side0.clone(), torch.zeros(n, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            nfp64 = batch.transport(*st2, s2, c2, S,
                                    step0=0, seed=7, mixed=True, count_fp64=True)
        return wall, kern, nfp64, batch.launch()

    wall64, kern64, _, launch64 = run(False)
    rate64 = args.steps * n * S / wall64
    log(f"[bench] fp64: {rate64 / 1e9:.2f} G track-steps/s (kernel {kern64 * 1e3:.2f} ms/launch)")
    wallmx, kernmx, nfp64, launchmx = run(True)
    ratemx = args.steps * n * S / wallmx
    log(f"[bench] mixed: {ratemx / 1e9:.2f} G track-steps/s (kernel {kernmx * 1e3:.2f} ms/launch), "
        f"fp64 redo fraction {nfp64 / (n * S):.4f}")

    roofline = None
    if os.path.exists(FLOPS_FIXTURE):
        flops = json.load(open(FLOPS_FIXTURE))
        per_step = float(flops["fp64_flops_per_track_step"])
        achieved = per_step * n * S / kern64 / 1e12
        roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                    "kernel": "transport_kernel<FP64>", "flops_per_track_step": per_step}

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

    navigation = None
    if args.nav_tracks > 0:
        navigation = run_navigation(args, ctx, dev, torch)

    line = {
        "metric": "track-steps/sec on 1 MI355X (mixed+learned vs plain FP64), 0 safety violations",
        "value": ratemx,
        "unit": "track-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wallmx / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"batched heterogeneous solids: {len(specs)} solids ({'/'.join(fams)}), "
                               f"{n} tracks, {S} track-steps per launch",
                   "tracks": n, "track_steps_per_launch": S, "families": list(fams)},
        "fp64_value": rate64,
        "mixed_over_fp64": ratemx / rate64,
        "fp64_redo_fraction": nfp64 / (n * S),
        "kernel_ms": {"fp64": kern64 * 1e3, "mixed": kernmx * 1e3},
        "launch": {"fp64": launch64, "mixed": launchmx},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "navigation": navigation,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
