"""Benchmark of the PIC time step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--precision fp32|fp64|both]
                    [--workload laser-foil|two-stream|large|ensemble|serve] [--no-cpu]

A step is one full main-loop iteration (H half step, push + charge-conserving
deposit, per-step sort by cell, H step, E step, J fold, laser source, Mur /
periodic walls) over the whole synthetic state, resident in HBM before the
timed region; the K steps are enqueued without host synchronisation.  Prints
one JSON line: marker-pushes/s (live markers x steps / wall) of the FP32
split form with the FP64 plain form beside it, the energy diagnostics of
every step computed inside the timed region (em_energy and
plasma_temperature of the reference loop, kept in a device ring), the
roofline of the push kernel from HIP events on the run's stream, p50/p99
time-to-result of a bounded serving stream in both forms, and the CPU oracle
timed on a bounded sample of the same workload on this host's cores (one
replicated copy per core, run as separate processes).
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
    p.add_argument("--precision", default="both", choices=["fp32", "fp64", "both"],
                   help="both: the FP32 split form gives 'value', the FP64 plain form "
                        "is reported beside it")
    p.add_argument("--workload", default="laser-foil",
                   choices=["laser-foil", "two-stream", "large", "ensemble", "serve"],
                   help="large: the 512x256x256 full-box fill (1.05e9 markers, FP32 split "
                        "form, streamed in by mpic_markers_append); "
                        "ensemble: 64 laser-foil scan runs advanced together (graph replay); "
                        "serve: Poisson stream of scan requests through the serving loop")
    p.add_argument("--runs", type=int, default=64, help="ensemble size")
    p.add_argument("--requests", type=int, default=48, help="serving: number of requests")
    p.add_argument("--rate", type=float, default=4.0, help="serving: mean arrivals per second")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--autotune", action="store_true",
                   help="pick the FP32 push's launch shape with mpic_autotune before the timed "
                        "steps (the default shape is one wavefront per tile, 16 workgroups per "
                        "CU)")
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="CPU baseline: seconds of oracle steps per core")
    p.add_argument("--cpu-cores", type=int, default=0,
                   help="CPU baseline: copies run at once (0: the cores this process may use, "
                        "at most 16)")
    p.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    p.add_argument("--serve-requests", type=int, default=24,
                   help="default workload: requests of the bounded serving stream giving "
                        "p50/p99 (0: skip)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="first run a plain device copy of MIB MiB three times: the reference "
                        "point that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    return p.parse_args()


def copy_calibration(mib):
    """A plain device-to-device copy of known size (read mib MiB, write mib
    MiB; the runtime's blit kernel) as the reference dispatch of a counter
    run.  It goes through the HIP runtime that libmpic.so itself is linked
    against, so that one runtime serves the whole process."""
    import ctypes
    import mandor3_amd as ma
    ma.lib()                                   # libamdhip64 is loaded with libmpic.so
    hip = ctypes.CDLL("libamdhip64.so.7")
    n = ctypes.c_size_t(mib << 20)
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), n) or hip.hipMalloc(ctypes.byref(dst), n):
        raise SystemExit("copy calibration: hipMalloc failed")
    ok = hip.hipMemset(src, 0, n) == 0
    for _ in range(3):
        ok = ok and hip.hipMemcpy(dst, src, n, 3) == 0      # hipMemcpyDeviceToDevice
    ok = ok and hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    if not ok:
        raise SystemExit("copy calibration: HIP call failed")
    print("copy calibration: %d MiB read + %d MiB written per dispatch" % (mib, mib),
          file=sys.stderr, flush=True)


def push_bytes(n_markers, plasma_cells, precision):
    """Algorithmic HBM bytes of one launch of the push kernel.

    FP64 plain form: marker read 64 B (its index in the order's
    permutation, x y z ux uy uz FP64, w qm FP32, id) and write 64 B (the
    state and id at its position in the order, and its new key); per cell
    holding markers E/H read 48 B and the J read-modify-write 48 B (the
    kernel flushes J itself).
    FP32 split form: marker read 40 B (its index in the order's permutation,
    two 16-byte records: cell word + 3 offsets, 3 momenta + w, and the id)
    and write 40 B (the records and id at its position in the order, and
    its new key); per cell holding markers E/H read 24 B.  Its J (48 B per
    node, FP64 read-modify-write) is summed by the separate gather kernel."""
    if precision == "fp64":
        return n_markers * (64 + 64) + plasma_cells * (48 + 48)
    return n_markers * (40 + 40) + plasma_cells * 24


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    counter summary (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled
    by the same counter's reading for a plain copy of known size), or None
    when no summary covers this kernel and workload."""
    path = os.path.join(ROOT, "profiles", "push_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        recs = json.load(f)
    for rec in recs:
        if rec.get("kernel") == kernel and rec.get("workload") == workload:
            return rec.get("bytes_per_launch")
    return None


PUSH32_KERNEL = "k_one_w<Push32ArgsW<%d> >"     # the FP32 push as the kernel trace names it
PUSH64_KERNEL = "k_one_w<Push64TArgs>"


def run_gpu(ma, wl, precision, steps, warmup, autotune=False):
    prec = ma.FP64 if precision == "fp64" else ma.FP32_SPLIT
    M = wl["markers"]
    r = ma.Run(wl["n"], wl["h"], wl["L"], wl["tau"], wl["bc"], 0.0, precision=prec,
               max_markers=M.size)
    z = np.zeros(3 * int(np.prod([v + 1 for v in wl["n"]])))
    r.set_fields(z, z)
    r.load_markers(M)
    if wl["laser"] is not None:
        r.add_laser(*wl["laser"])
    r.init()
    tuned = None
    if precision == "fp32" and autotune:
        # the launch shape of the push (markers per work item, workgroups),
        # picked by timing real steps before the timed region
        tuned = r.autotune(steps=2)
    live = r.marker_count()
    r.advance(warmup)
    r.sync()
    t0 = time.perf_counter()
    r.advance(steps)
    r.sync()
    wall = time.perf_counter() - t0
    push_ms, dep_ms, rest_ms = r.timing()
    live_end = r.marker_count()
    hist = r.diag_history(steps)
    r.close()
    if hist.shape[0] != steps or not np.all(np.isfinite(hist)):
        raise SystemExit("energy diagnostics of the timed steps missing or not finite")
    avg_push = push_ms / steps
    alg = push_bytes(live, wl["plasma_cells"], precision)
    return {"pushes_per_s": 0.5 * (live + live_end) * steps / wall,
            "ms_per_step": 1e3 * wall / steps,
            "cell_updates_per_s": int(np.prod(wl["n"])) * steps / wall,
            "markers": int(live),
            "push_kernel_ms": avg_push, "deposit_rest_ms": dep_ms / steps,
            "rest_of_step_ms": rest_ms / steps,
            "push_kernel_bytes": alg,
            "push_kernel_GBps": alg / (avg_push * 1e-3) / 1e9,
            "launch_shape": ({"workgroups": tuned[0], "wavefronts_per_workgroup": tuned[1],
                              "markers_per_item": tuned[2], "picked_by": "mpic_autotune"}
                             if tuned is not None else
                             {"workgroups": "16 per CU", "wavefronts_per_workgroup": 1,
                              "markers_per_item": 0, "picked_by": "default"}
                             if precision == "fp32" else None),
            "last_diag": [float(v) for v in hist[-1]]}


def cpu_worker(spec):
    """One CPU-baseline process: the oracle (serial FP64 restatement of the
    reference path) on a fresh copy of the workload, full steps for the given
    seconds; prints one JSON line {markers, steps, seconds}."""
    workload, seconds = spec.split(":")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    from mandor3_amd import workloads
    wl = workloads.laser_foil() if workload == "laser-foil" else workloads.two_stream_box()
    n = wl["n"]
    o = Oracle(n, wl["h"], wl["L"], wl["tau"], wl["bc"], 0.0)
    z = np.zeros(3 * int(np.prod([v + 1 for v in n])))
    o.set_fields(z, z)
    o.set_markers(wl["markers"])
    if wl["laser"] is not None:
        o.add_laser(wl["laser"][0], wl["laser"][1])
    o.init()
    nm = o.marker_count()
    del wl, z
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step()
        steps += 1
        el = time.perf_counter() - t0
        if el >= float(seconds) or steps >= 200:
            break
    o.close()
    print(json.dumps({"markers": nm, "steps": steps, "seconds": el}))


def host_cores():
    try:
        c = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        c = os.cpu_count() or 1
    return max(1, min(c, 16))


def cpu_baseline(workload, seconds, cores):
    """The CPU oracle on this host's cores: `cores` processes, each running
    full steps of its own copy of the same workload for `seconds` (the
    reference's ensemble decomposition, independent members, one per core,
    with no communication: an upper bound of what its MPI path reaches on
    these cores).  The rate is the sum of the copies' marker-pushes/s."""
    import subprocess
    cores = cores if cores > 0 else host_cores()
    env = dict(os.environ, OMP_NUM_THREADS="1")
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                               "%s:%g" % (workload, seconds)], stdout=subprocess.PIPE, env=env)
             for _ in range(cores)]
    outs = []
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise SystemExit("CPU baseline worker failed (%d)" % p.returncode)
        outs.append(json.loads(out.decode().strip().splitlines()[-1]))
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    rate = sum(o["markers"] * o["steps"] / o["seconds"] for o in outs)
    per = [o["markers"] * o["steps"] / o["seconds"] for o in outs]
    return {"value": rate, "unit": "marker-pushes/s", "cores": cores, "kind": "port",
            "sample": "%d processes, one per core, each %d-%d full steps of its own copy of the "
                      "same workload (%d markers) on the serial FP64 oracle, %.1f-%.1f s; "
                      "per-core %.3e-%.3e marker-pushes/s; %s"
                      % (cores, min(o["steps"] for o in outs), max(o["steps"] for o in outs),
                         outs[0]["markers"], min(o["seconds"] for o in outs),
                         max(o["seconds"] for o in outs), min(per), max(per), cpu)}


def serve_latency(requests, rate, steps=(500, 3000), seed=2024):
    """A bounded serving stream (Poisson arrivals at `rate`, scan requests
    with 10 % 256x128x128 requests at 16 markers per cell, one of them
    forced into the stream) through the continuous-batching loop in both
    forms: p50/p99 time-to-result and aggregate marker-pushes/s."""
    from mandor3_amd import serve
    import mandor3_amd as ma
    out = {}
    for p in ("fp32", "fp64"):
        reqs = serve.poisson_requests(requests, rate, seed, 0.1, steps)
        if not any(tuple(r.params["n"]) == (256, 128, 128) for r in reqs):
            from mandor3_amd import workloads
            rng = np.random.default_rng(seed + 1)
            reqs[len(reqs) // 2].params = workloads.scan_request(rng, 1.0, steps)
        srv = serve.Server(ma.FP32_SPLIT if p == "fp32" else ma.FP64, 240.0, 25)
        for req in reqs:
            srv.submit(req)
        out[p] = srv.serve()
        print("serve %s: %d requests, p50 %.2f s p99 %.2f s" % (p, out[p]["requests"],
              out[p]["p50_s"], out[p]["p99_s"]), file=sys.stderr, flush=True)
    return out


def run_ensemble(ma, workloads, precision, runs, steps, warmup, seed=64):
    """The ensemble configuration: `runs` scan points drawn from `seed`
    (grids 32^3 .. 128x64x64, 4-32 markers per cell, a0 0.5-5, 5-50 n_c,
    0.2-1 lambda), each on its own stream, advanced together by replaying
    their captured steps."""
    rng = np.random.default_rng(seed)
    prec = ma.FP64 if precision == "fp64" else ma.FP32_SPLIT
    from mandor3_amd import serve
    rs, markers = [], 0
    for _ in range(runs):
        st = workloads.foil_request(**workloads.scan_request(rng, 0.0, (steps, steps)))
        r = serve.start_run(st, prec, graphs=True)
        rs.append(r)
        markers += st["markers"].size
    print("ensemble %s: %d runs, %d markers set up" % (precision, runs, markers), file=sys.stderr,
          flush=True)
    ma.batch_advance(rs, warmup)
    for r in rs:
        r.sync()
    t0 = time.perf_counter()
    ma.batch_advance(rs, steps)
    for r in rs:
        r.sync()
    wall = time.perf_counter() - t0
    for r in rs:
        r.close()
    return {"pushes_per_s": markers * steps / wall, "ms_per_step": 1e3 * wall / steps,
            "markers": int(markers), "runs": runs}


def main_ensemble(a):
    import mandor3_amd as ma
    from mandor3_amd import workloads
    res = {p: run_ensemble(ma, workloads, p, a.runs, a.steps, a.warmup) for p in ("fp32", "fp64")}
    m = res["fp32"]
    print(json.dumps({
        "metric": "marker-pushes/s incl. deposit+FDTD, FP32 vs FP64 on 1 GPU; p50/p99 run latency",
        "value": m["pushes_per_s"], "unit": "marker-pushes/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": m["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "ensemble of %d laser-foil scan runs (grids 32^3..128x64x64, "
                               "4-32 ppc, a0 0.5-5, 5-50 n_c, 0.2-1 lambda), graph replay, "
                               "one stream per run" % a.runs,
                   "markers": m["markers"]},
        "fp64_plain": {"value": res["fp64"]["pushes_per_s"], "ms_per_step": res["fp64"]["ms_per_step"]},
        "fp32_over_fp64": m["pushes_per_s"] / res["fp64"]["pushes_per_s"]}))


def main_serve(a):
    from mandor3_amd import serve
    import mandor3_amd as ma
    out = {}
    for p in ("fp32", "fp64"):
        srv = serve.Server(ma.FP32_SPLIT if p == "fp32" else ma.FP64, 240.0, 25)
        for req in serve.poisson_requests(a.requests, a.rate, 2024, 0.1, (a.steps, 6 * a.steps)):
            srv.submit(req)
        out[p] = srv.serve()
        print("serve %s: %d requests done" % (p, out[p]["requests"]), file=sys.stderr, flush=True)
    m = out["fp32"]
    print(json.dumps({
        "metric": "marker-pushes/s incl. deposit+FDTD, FP32 vs FP64 on 1 GPU; p50/p99 run latency",
        "value": m["marker_pushes_per_s"], "unit": "marker-pushes/s", "n_gpus": 1,
        "steps": a.steps, "warmup": 0, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "serving: Poisson %.1f req/s, %d laser-foil scan requests "
                               "(10%% at 256x128x128 16 ppc), 240 GB budget" % (a.rate, a.requests)},
        "p50_s": m["p50_s"], "p99_s": m["p99_s"], "fp64_plain": out["fp64"]}))


def main_large(a):
    """The 512x256x256 single-GPU configuration at its full-box fill, FP32
    split form only (the FP64 plain form would need 2 x 1.05e9 x 60 B of
    double-buffered markers).  The records are built a few planes at a time
    on a second thread while the previous piece is appended.  Besides the
    rate it checks what holds at any size: the per-cell counts of the final
    sort add up to the live markers, and markers are only ever removed (by
    the open walls)."""
    from concurrent.futures import ThreadPoolExecutor
    import mandor3_amd as ma
    from mandor3_amd import workloads
    if ma.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if a.copy_calibration > 0:
        copy_calibration(a.copy_calibration)
    wl = workloads.large_foil()
    r = ma.Run(wl["n"], wl["h"], wl["L"], wl["tau"], wl["bc"], 0.0, precision=ma.FP32_SPLIT,
               max_markers=wl["n_markers"])
    z = np.zeros(3 * int(np.prod([v + 1 for v in wl["n"]])))
    r.set_fields(z, z)
    del z
    t0 = time.perf_counter()
    loaded, pieces = 0, 0
    gen = wl["chunks"]()
    with ThreadPoolExecutor(1) as ex:
        fut = ex.submit(next, gen, None)
        while True:
            M = fut.result()
            if M is None:
                break
            fut = ex.submit(next, gen, None)       # the other buffer
            r.append_markers(M)
            loaded += M.size
            pieces += 1
            if pieces % 8 == 0:
                print("large: %d of %d markers appended (%.0f s)" % (loaded, wl["n_markers"],
                      time.perf_counter() - t0), file=sys.stderr, flush=True)
    r.add_laser(*wl["laser"])
    r.init()
    live = r.marker_count()
    print("large: %d live markers after init (%.0f s)" % (live, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    r.advance(a.warmup)
    r.sync()
    t1 = time.perf_counter()
    r.advance(a.steps)
    r.sync()
    wall = time.perf_counter() - t1
    push_ms, dep_ms, rest_ms = r.timing()
    live_end = r.marker_count()
    counts = r.sort(with_counts=True)
    counts_ok = int(counts.sum()) == live_end == r.marker_count()
    r.close()
    avg_push = push_ms / a.steps
    alg = push_bytes(live, wl["plasma_cells"], "fp32")
    gbs = alg / (avg_push * 1e-3) / 1e9
    print(json.dumps({
        "metric": "marker-pushes/s incl. deposit+FDTD, FP32 vs FP64 on 1 GPU; p50/p99 run latency",
        "value": 0.5 * (live + live_end) * a.steps / wall, "unit": "marker-pushes/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl["name"], "grid": list(wl["n"]), "markers": int(live),
                   "markers_end": int(live_end), "loaded": int(loaded), "pieces": pieces,
                   "precision": "fp32 split (int32 cell + FP32 offset; FP64 J sum)",
                   "cell_updates_per_s": int(np.prod(wl["n"])) * a.steps / wall,
                   "push_kernel_ms": avg_push, "deposit_gather_ms": dep_ms / a.steps,
                   "rest_of_step_ms": rest_ms / a.steps, "setup_s": t1 - t0},
        "roofline": {"bound": "hbm", "kernel": PUSH32_KERNEL % 1, "achieved": gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                     "traffic": measured_traffic(PUSH32_KERNEL % 1, wl["name"])},
        "checks": {"sort_counts_sum_to_live": counts_ok,
                   "markers_only_removed": bool(live_end <= live)}}))


def main():
    a = parse()
    if a.cpu_worker:
        return cpu_worker(a.cpu_worker)
    if a.gpus == 1 and a.workload == "ensemble":
        return main_ensemble(a)
    if a.gpus == 1 and a.workload == "serve":
        return main_serve(a)
    if a.gpus == 1 and a.workload == "large":
        return main_large(a)
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only (--gpus %d requested); "
              "multi-GPU sharding is design-only" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import mandor3_amd as ma
    from mandor3_amd import workloads
    if ma.device_count() < 1:
        raise SystemExit("no HIP device visible")
    if a.copy_calibration > 0:
        copy_calibration(a.copy_calibration)
    wl = workloads.laser_foil() if a.workload == "laser-foil" else workloads.two_stream_box()
    main_prec = "fp64" if a.precision == "fp64" else "fp32"
    res = {main_prec: run_gpu(ma, wl, main_prec, a.steps, a.warmup, a.autotune)}
    if a.precision == "both":
        res["fp64"] = run_gpu(ma, wl, "fp64", a.steps, a.warmup)
    m = res[main_prec]
    kernel = (PUSH32_KERNEL % (m["launch_shape"]["wavefronts_per_workgroup"])
              if main_prec == "fp32" else PUSH64_KERNEL)
    out = {
        "metric": "marker-pushes/s incl. deposit+FDTD, FP32 vs FP64 on 1 GPU; p50/p99 run latency",
        "value": m["pushes_per_s"],
        "unit": "marker-pushes/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_prec == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": wl["name"], "grid": list(wl["n"]), "markers": m["markers"],
                   "precision": "fp32 split (int32 cell + FP32 offset; FP64 J sum)"
                   if main_prec == "fp32" else "fp64 plain",
                   "cell_updates_per_s": m["cell_updates_per_s"],
                   "push_kernel_ms": m["push_kernel_ms"],
                   "deposit_gather_ms": m["deposit_rest_ms"],
                   "rest_of_step_ms": m["rest_of_step_ms"],
                   "push_launch_shape": m["launch_shape"]},
        "roofline": {"bound": "hbm", "kernel": kernel, "achieved": m["push_kernel_GBps"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": m["push_kernel_GBps"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(kernel, wl["name"])},
    }
    if "fp64" in res and main_prec != "fp64":
        f = res["fp64"]
        out["fp64_plain"] = {"value": f["pushes_per_s"], "ms_per_step": f["ms_per_step"],
                             "push_kernel_ms": f["push_kernel_ms"],
                             "push_kernel_GBps": f["push_kernel_GBps"]}
        out["fp32_over_fp64"] = m["pushes_per_s"] / f["pushes_per_s"]
    out["config"]["energy_diagnostics"] = "every step, on the device (last row {Time, W_E, W_M, " \
                                          "WTx, WTy, WTz}: %s)" % (
        ", ".join("%.6g" % v for v in m["last_diag"]))
    if a.serve_requests > 0:
        sv = serve_latency(a.serve_requests, a.rate)
        out["p50_s"] = sv["fp32"]["p50_s"]
        out["p99_s"] = sv["fp32"]["p99_s"]
        out["serving"] = {
            "stream": "Poisson %.1f req/s, %d laser-foil scan requests (grids 32^3..128x64x64, "
                      "4-32 ppc, 500-3000 steps; 10%% at 256x128x128 16 ppc), 240 GB budget"
                      % (a.rate, a.serve_requests),
            "fp32": sv["fp32"], "fp64_plain": sv["fp64"]}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(a.workload, a.cpu_seconds, a.cpu_cores)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
