"""Benchmark of the PIC time step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--precision fp64|fp32]
                    [--workload laser-foil|two-stream] [--no-cpu]

A step is one full main-loop iteration (H half step, push + charge-conserving
deposit, H step, E step, J fold, laser source, Mur/periodic walls) over the
whole synthetic state, resident in HBM before the timed region.  Prints one
JSON line: marker-pushes/s (live markers x steps / wall), the roofline of the
push+deposit kernel from HIP events on the run's stream, and the CPU oracle
timed on a bounded sample of the same workload on this host.
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


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--precision", default="fp64", choices=["fp64", "fp32"])
    p.add_argument("--workload", default="laser-foil", choices=["laser-foil", "two-stream"])
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--untiled", action="store_true",
                   help="unsorted one-thread-per-marker push (J atomics straight to HBM)")
    p.add_argument("--cpu-seconds", type=float, default=12.0)
    return p.parse_args()


def push_bytes(n_markers, plasma_cells, precision):
    """Algorithmic HBM bytes of one push+deposit launch: SoA marker read
    (x y z ux uy uz as FP64 + w qm as FP32 = 56 B) and write (48 B), plus,
    per cell holding markers, E/H read (48 B) and J read-modify-write (48 B)."""
    if precision == "fp64":
        return n_markers * (56 + 48) + plasma_cells * (48 + 48)
    return n_markers * (32 + 28) + plasma_cells * (24 + 48)


def cpu_baseline(wl, seconds):
    """The CPU oracle (serial FP64 restatement of the reference path) on the
    same state, timed for a bounded number of steps."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    n = wl["n"]
    o = Oracle(n, wl["h"], wl["L"], wl["tau"], wl["bc"], 0.0)
    z = np.zeros(3 * int(np.prod([v + 1 for v in n])))
    o.set_fields(z, z)
    o.set_markers(wl["markers"])
    if wl["laser"] is not None:
        o.add_laser(wl["laser"][0], wl["laser"][1])
    o.init()
    nm = o.marker_count()
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= seconds or steps >= 50:
            break
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    o.close()
    return {"value": nm * steps / el, "unit": "marker-pushes/s", "cores": 1, "kind": "port",
            "sample": "%d full steps of the same workload (%d markers) on the serial FP64 "
                      "oracle, %.1f s, %s" % (steps, nm, el, cpu)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); "
              "multi-GPU sharding is design-only" % a.gpus, file=sys.stderr)
        sys.exit(2)
    if a.untiled:
        os.environ["MPIC_UNTILED"] = "1"
    import mandor3_amd as ma
    from mandor3_amd import workloads
    if ma.device_count() < 1:
        raise SystemExit("no HIP device visible")
    wl = workloads.laser_foil() if a.workload == "laser-foil" else workloads.two_stream_box()
    prec = ma.FP64 if a.precision == "fp64" else ma.FP32_SPLIT
    M = wl["markers"]
    r = ma.Run(wl["n"], wl["h"], wl["L"], wl["tau"], wl["bc"], 0.0, precision=prec,
               max_markers=M.size)
    z = np.zeros(3 * int(np.prod([v + 1 for v in wl["n"]])))
    r.set_fields(z, z)
    r.load_markers(M)
    if wl["laser"] is not None:
        r.add_laser(*wl["laser"])
    r.init()
    live = r.marker_count()
    r.advance(a.warmup)
    r.sync()
    t0 = time.perf_counter()
    r.advance(a.steps)
    r.sync()
    wall = time.perf_counter() - t0
    push_ms, field_ms = r.timing()
    live_end = r.marker_count()
    pushes = 0.5 * (live + live_end) * a.steps
    ncells = int(np.prod(wl["n"]))
    avg_push_ms = push_ms / a.steps
    alg = push_bytes(live, wl["plasma_cells"], a.precision)
    achieved = alg / (avg_push_ms * 1e-3) / 1e9
    out = {
        "metric": "marker-pushes/s incl. deposit+FDTD, FP32 vs FP64 on 1 GPU; p50/p99 run latency",
        "value": pushes / wall,
        "unit": "marker-pushes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": wl["name"], "grid": list(wl["n"]), "markers": int(live),
                   "precision": a.precision, "tiled": not a.untiled,
                   "cell_updates_per_s": ncells * a.steps / wall,
                   "push_ms_per_step": avg_push_ms, "field_ms_per_step": field_ms / a.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
    }
    r.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, a.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
