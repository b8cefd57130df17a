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
TRAFFIC_FIXTURE = os.path.join(ROOT, "vecgeom_amd", "data", "hbm_traffic.json")


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


def family_roofline(fams_ms, counts_by_family, S):
    """Roofline of the dominant (longest) family kernel of a timed transport:
    algorithmic FP64 flops per launch (flop fixture x the family's tracks x
    track-steps) over its average launch time."""
    if not os.path.exists(FLOPS_FIXTURE) or not fams_ms:
        return None
    fx = json.load(open(FLOPS_FIXTURE))
    fam = max(fams_ms, key=fams_ms.get)
    per_step = float(fx["per_family"][fam])
    flops = per_step * counts_by_family[fam] * S
    t = fams_ms[fam] * 1e-3
    achieved = flops / t / 1e12
    out = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
           "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
           "kernel": f"transport_family_kernel<{fam}, FP64>", "kernel_ms": fams_ms[fam],
           "flops_per_track_step": per_step, "tracks": counts_by_family[fam], "track_steps": S,
           "note": "FP64 vector peak; flops counted on the FP64 track-step path (+ - * / sqrt)"}
    if os.path.exists(TRAFFIC_FIXTURE):
        tr = json.load(open(TRAFFIC_FIXTURE))
        if tr.get("kernel") == out["kernel"] and tr.get("tracks") == out["tracks"] and tr.get("track_steps") == S:
            out["traffic"] = tr["bytes_per_launch"]
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
    ap.add_argument("--launch", default=None, help="block,grid,tile launch shape instead of autotuning")
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
        if args.launch:
            batch.set_launch(*[int(v) for v in args.launch.split(",")])
        elif not args.no_autotune:
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
        evs, fms = [], []
        t = time.perf_counter()
        for _ in range(args.steps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            batch.transport(*state, side, cr, S, step0=step0, seed=7, mixed=mixed, asynchronous=True,
                            timing=True)
            b.record(stream)
            evs.append((a, b))
            fms.append(batch.family_ms())  # waits for this launch's family kernels
            step0 += S
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        kern = float(np.mean([a.elapsed_time(b) for a, b in evs])) * 1e-3
        fam_ms = {f: float(np.mean([m[f] for m in fms])) for f in fms[0]}
        nfp64 = 0
        if mixed:
            st2 = [x.clone() for x in init]
            s2, c2 = side0.clone(), torch.zeros(n, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            nfp64 = batch.transport(*st2, s2, c2, S,
                                    step0=0, seed=7, mixed=True, count_fp64=True)
        return wall, kern, nfp64, batch.launch(), fam_ms

    counts_by_family = {}
    for (kind, _), c in zip(specs, counts):
        counts_by_family[kind] = counts_by_family.get(kind, 0) + int(c)
    wall64, kern64, _, launch64, fam64 = run(False)
    rate64 = args.steps * n * S / wall64
    log(f"[bench] fp64: {rate64 / 1e9:.2f} G track-steps/s (kernel {kern64 * 1e3:.2f} ms/launch)")
    log("[bench] fp64 family kernel ms: " + ", ".join(f"{k} {v:.1f}" for k, v in fam64.items()))
    wallmx, kernmx, nfp64, launchmx, fammx = run(True)
    ratemx = args.steps * n * S / wallmx
    log(f"[bench] mixed: {ratemx / 1e9:.2f} G track-steps/s (kernel {kernmx * 1e3:.2f} ms/launch), "
        f"fp64 redo fraction {nfp64 / (n * S):.4f}")

    log("[bench] mixed family kernel ms: " + ", ".join(f"{k} {v:.1f}" for k, v in fammx.items()))
    roofline = family_roofline(fam64, counts_by_family, S)

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
        "family_kernel_ms": {"fp64": fam64, "mixed": fammx},
        "launch": {"fp64": launch64, "mixed": launchmx},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "navigation": navigation,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
