"""Per-family counters of the batched transport, from rocprofv3 CSV output.

Reads the counter-collection CSVs of `rocprofv3 --pmc` runs over
`bench.py --only tuned --launch 256,0,256,1 --modes ...` (families back to
back on one stream) and of `bench.py --copy-probe` (a plain 1 GiB read + 1 GiB
write), and writes

  * hbm_traffic.json: per family and step mode, FETCH_SIZE and WRITE_SIZE per
    dispatch, scaled to bytes by the copy probe, against the algorithmic bytes
    of a launch (a track's state read once and written once), which bench.py
    reports as `roofline.traffic`;
  * Markdown tables (stdout) of the same, and of the SQ/TCC counters when
    given, for the profile summaries.

Usage:
  python3 make_traffic.py --probe-fetch P/run_counter_collection.csv
      --probe-write Q/run_counter_collection.csv
      --set LABEL:FETCH.csv:WRITE.csv [--set ...]
      [--trace T/run_kernel_stats.csv] [--sq S.csv] [--tcc C.csv]
      [--tracks N] [--track-steps K] [--json hbm_traffic.json]
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import re
import sys
from collections import defaultdict

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

# a track's state: six doubles, the int8 side and the int32 crossing count
STATE_BYTES = 6 * 8 + 1 + 4
KERNEL_RE = re.compile(r"transport_family_kernel<vgm::(\w+)Shape, (\d), (true|false), (\d)>")
FAMILY_NAMES = {"Box": "box", "Tube": "tube", "Cone": "cone", "Trd": "trd", "Trapezoid": "trapezoid",
                "Polycone": "polycone", "Polyhedron": "polyhedron", "Tessellated": "tessellated", "Orb": "orb",
                "CutTube": "cut_tube", "Boolean": "boolean"}


def kernel_key(name):
    """(family, mode) of a transport family kernel, or None."""
    m = KERNEL_RE.search(name)
    if not m:
        return None
    fam = FAMILY_NAMES[m.group(1)]
    mode, order = int(m.group(2)), int(m.group(4))
    if mode == 1:
        return fam, "mixed"
    return fam, {0: "fp64", 1: "distance-first", 2: "culled"}[order]


def read_counters(path):
    """{(kernel name, counter): [value per dispatch]}, plus VGPR/scratch per kernel."""
    vals = defaultdict(list)
    res = {}
    with open(path) as f:
        for row in csv.DictReader(f):
            vals[(row["Kernel_Name"], row["Counter_Name"])].append(float(row["Counter_Value"]))
            # the CSV counts VGPRs in granules of two registers
            res[row["Kernel_Name"]] = (2 * int(row["VGPR_Count"]), int(row["Scratch_Size"]))
    return vals, res


def mean(v):
    return sum(v) / len(v)


def probe_kib(path, counter):
    """The copy probe's counter per 1 GiB dispatch (mean of its dispatches of >= 2^27 work-items)."""
    with open(path) as f:
        rows = [r for r in csv.DictReader(f) if r["Counter_Name"] == counter and int(r["Grid_Size"]) >= (1 << 24)
                and "elementwise" in r["Kernel_Name"]]
    if not rows:
        raise SystemExit(f"no copy-probe dispatch with {counter} in {path}")
    return mean([float(r["Counter_Value"]) for r in rows])


def kernel_ms(path):
    out = {}
    if not path:
        return out
    with open(path) as f:
        for row in csv.DictReader(f):
            k = kernel_key(row["Name"])
            if k:
                out[k] = float(row["AverageNs"]) * 1e-6
    return out


def trace_table(stats_path, tracks):
    """Markdown table of the transport kernels of a kernel-trace run."""
    trace_path = os.path.join(os.path.dirname(stats_path), os.path.basename(stats_path).replace(
        "kernel_stats", "kernel_trace"))
    res = {}
    with open(trace_path) as f:
        for row in csv.DictReader(f):
            res[row["Kernel_Name"]] = (2 * int(row["VGPR_Count"]), int(row["Scratch_Size"]), int(row["LDS_Block_Size"]))
    rows = []
    with open(stats_path) as f:
        for row in csv.DictReader(f):
            k = kernel_key(row["Name"])
            if k:
                rows.append((k, int(row["Calls"]), float(row["AverageNs"]) * 1e-6, res.get(row["Name"], (0, 0, 0))))
    print("| kernel | mode | tracks | dispatches | mean ms | VGPRs | scratch B/lane | LDS B |")
    print("|---|---|---|---|---|---|---|---|")
    total = 0.0
    for (fam, mode), calls, t, (vg, scr, lds) in sorted(rows, key=lambda r: -r[2]):
        total += t
        print(f"| transport_family_kernel<{fam}> | {mode} | {tracks[fam]} | {calls} | {t:.2f} | {vg} | {scr} | {lds} |")
    print(f"| total | | {sum(tracks.values())} | | {total:.1f} | | | |\n")


def family_tracks(total, fams):
    from vecgeom_amd import workloads

    specs, counts = workloads.batched_solids(total, families=fams, per_family=8, seed=42)
    per = defaultdict(int)
    for (kind, _), c in zip(specs, counts):
        per[kind] += int(c)
    return per


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--probe-fetch", required=True)
    ap.add_argument("--probe-write", required=True)
    ap.add_argument("--set", action="append", default=[], help="LABEL:FETCH.csv:WRITE.csv")
    ap.add_argument("--trace", default=None, help="run_kernel_stats.csv of a kernel-trace run (kernel ms)")
    ap.add_argument("--sq", default=None, help="counter CSV of the SQ counters")
    ap.add_argument("--tcc", default=None, help="counter CSV of TCC_HIT_sum / TCC_MISS_sum")
    ap.add_argument("--tracks", type=int, default=40_000_000)
    ap.add_argument("--track-steps", type=int, default=50)
    ap.add_argument("--json", default=None)
    args = ap.parse_args()

    gib_kib = float(1 << 30) / 1024.0
    fetch_factor = gib_kib / probe_kib(args.probe_fetch, "FETCH_SIZE")
    write_factor = gib_kib / probe_kib(args.probe_write, "WRITE_SIZE")
    ms = kernel_ms(args.trace)
    tracks = family_tracks(args.tracks, ("box", "tube", "cone", "trd", "trapezoid", "polycone", "polyhedron",
                                         "tessellated"))
    kernels = []
    if args.trace:
        trace_table(args.trace, tracks)
    print(f"Scale factors from the copy probe: FETCH_SIZE x{fetch_factor:.3f}, WRITE_SIZE x{write_factor:.3f}.\n")
    print("| kernel | mode | VGPRs | scratch B/lane | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM GB per launch (scaled) "
          "| algorithmic GB | ratio |")
    print("|---|---|---|---|---|---|---|---|---|")
    for spec in args.set:
        label, fpath, wpath = spec.split(":")
        fv, res = read_counters(fpath)
        wv, _ = read_counters(wpath)
        for (name, counter), v in fv.items():
            k = kernel_key(name)
            if not k or counter != "FETCH_SIZE":
                continue
            w = wv.get((name, "WRITE_SIZE"))
            if not w:
                continue
            fam, mode = k
            fk, wk = mean(v), mean(w)
            hbm = (fk * fetch_factor + wk * write_factor) * 1024.0
            alg = 2 * STATE_BYTES * tracks[fam]
            vg, scr = res[name]
            print(f"| transport_family_kernel<{fam}> | {mode} | {vg} | {scr} | {fk:.0f} | {wk:.0f} | {hbm / 1e9:.3f} "
                  f"| {alg / 1e9:.3f} | {hbm / alg:.1f} |")
            kernels.append({"family": fam, "mode": mode, "tracks": tracks[fam], "track_steps": args.track_steps,
                            "fetch_kib": fk, "write_kib": wk, "bytes_per_launch": hbm, "algorithmic_bytes": alg,
                            "kernel_ms": ms.get(k), "vgprs": vg, "scratch_bytes_per_lane": scr,
                            "source": f"profiles/transport_40M_hbm_counters.md ({label})"})
    if args.sq:
        sq, _ = read_counters(args.sq)
        tcc, _ = read_counters(args.tcc) if args.tcc else ({}, {})
        names = sorted({n for (n, c) in sq if kernel_key(n)}, key=lambda n: -(ms.get(kernel_key(n)) or 0))
        print("\n| kernel | mode | ms | waves | VALU insts | FP64 flops | executed FP64 TF/s | lane util "
              "| VALU issue share | L2 hit |")
        print("|---|---|---|---|---|---|---|---|---|---|")
        for n in names:
            fam, mode = kernel_key(n)
            g = lambda c: mean(sq[(n, c)]) if (n, c) in sq else float("nan")
            t = ms.get((fam, mode))
            f64 = g("SQ_INSTS_VALU_FLOPS_FP64")
            util = g("SQ_THREAD_CYCLES_VALU") / (64.0 * g("SQ_ACTIVE_INST_VALU"))
            issue = g("SQ_ACTIVE_INST_VALU") / g("SQ_WAVE_CYCLES")
            hit = tcc.get((n, "TCC_HIT_sum"))
            miss = tcc.get((n, "TCC_MISS_sum"))
            l2 = mean(hit) / (mean(hit) + mean(miss)) if hit and miss else float("nan")
            rate = f64 / (t * 1e-3) / 1e12 if t else float("nan")
            print(f"| transport_family_kernel<{fam}> | {mode} | {t:.2f} | {g('SQ_WAVES'):.0f} | "
                  f"{g('SQ_INSTS_VALU'):.3g} | {f64:.3g} | {rate:.2f} | {util:.2f} | {issue:.3f} | {l2:.3f} |")
    if args.json:
        with open(args.json, "w") as f:
            json.dump({"fetch_factor": fetch_factor, "write_factor": write_factor, "kernels": kernels}, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
