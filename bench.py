"""Benchmark driver (one JSON line on stdout).

Workload: box SOA3D queries (half-lengths 10,20,30 mm; 1,000,000 points uniform
in the box scaled x1.5; isotropic directions, seed 42; stepMax 1e30).  One step
= Inside + DistanceToIn + DistanceToOut (inside points) + SafetyToIn +
SafetyToOut over the batch, inputs resident in HBM.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import torch

    import vecgeom_amd as vg
    from vecgeom_amd import sampling

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    ctx = vg.Context(0, stream=stream.cuda_stream)
    n = 1_000_000
    half = (10.0, 20.0, 30.0)
    pts = sampling.points_in_box(42, 1, n, half, 1.5)
    dirs = sampling.isotropic_directions(42, 2, n)
    shape = ctx.box(*half)
    pv = ctx.place(ctx.logical_volume(shape))
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    P = vg.SOA3D(*(T(a) for a in pts))
    D = vg.SOA3D(*(T(a) for a in dirs))
    step = T(np.full(n, 1e30))
    ins = ctx.inside(pv, P)
    sel = (ins == vg.kInside)
    Pin = vg.SOA3D(P.x[sel].contiguous(), P.y[sel].contiguous(), P.z[sel].contiguous())
    Din = vg.SOA3D(D.x[sel].contiguous(), D.y[sel].contiguous(), D.z[sel].contiguous())
    stepin = step[sel].contiguous()
    n_in = Pin.n
    o_i = torch.empty(n, dtype=torch.int32, device=dev)
    o_d = torch.empty(n, dtype=torch.float64, device=dev)
    o_do = torch.empty(n_in, dtype=torch.float64, device=dev)
    o_s = torch.empty(n, dtype=torch.float64, device=dev)

    ev = {}

    def one_step(record=False):
        def rec(name, fn):
            if record:
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                fn()
                b.record(stream)
                ev.setdefault(name, []).append((a, b))
            else:
                fn()
        rec("Inside", lambda: ctx.inside(pv, P, out=o_i, asynchronous=True))
        rec("DistanceToIn", lambda: ctx.distance_to_in(pv, P, D, step, out=o_d, asynchronous=True))
        rec("DistanceToOut", lambda: ctx.distance_to_out(pv, Pin, Din, stepin, out=o_do, asynchronous=True))
        rec("SafetyToIn", lambda: ctx.safety_to_in(pv, P, out=o_s, asynchronous=True))
        rec("SafetyToOut", lambda: ctx.safety_to_out(pv, P, out=o_s, asynchronous=True))

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_step(record=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    queries = args.steps * (4 * n + n_in)
    kern_ms = {k: float(np.mean([a.elapsed_time(b) for a, b in v])) for k, v in ev.items()}
    # dominant kernel: DistanceToIn, 24 B point + 24 B dir + 8 B step_max in, 8 B out
    bytes_q = 64.0
    t_dti = kern_ms["DistanceToIn"] * 1e-3
    achieved = bytes_q * n / t_dti / 1e9

    cpu = None
    if not args.no_cpu:
        import oracle
        geo = oracle.Geometry()
        pvo = geo.place(geo.logical_volume(geo.box(*half)))
        m = 200_000
        sp = tuple(a[:m] for a in pts)
        sd = tuple(a[:m] for a in dirs)
        nth = os.cpu_count() or 1
        best = 1e30
        for _ in range(3):
            c0 = time.perf_counter()
            geo.query(pvo, "DistanceToIn", sp, sd, np.full(m, 1e30), nthreads=nth)
            best = min(best, time.perf_counter() - c0)
        cpu = {"value": m / best, "unit": "DistanceToIn queries/s", "cores": nth, "kind": "port",
               "sample": f"{m} box DistanceToIn queries, best of 3"}

    line = {
        "metric": "box SOA3D queries/s (Inside, DistanceToIn, DistanceToOut, SafetyToIn, SafetyToOut)",
        "value": queries / wall,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "box 10x20x30 mm, 1M points in 1.5x box, seed 42", "points": n},
        "kernel_ms": kern_ms,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
