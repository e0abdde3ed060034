"""Continuous-batching service of laser-foil scan requests on one GPU.

Requests (a grid, markers per cell, a0, density, thickness, steps) arrive
over time.  The loop keeps every admitted run resident in HBM (fields,
markers, tiles, per-run stream and captured step graphs), admits waiting
requests whenever their device footprint (mpic_footprint) fits the HBM
budget, and advances all resident runs together in slices of steps with
mpic_batch_advance: the resident runs' steps are fused into one launch per
kernel kind for all of them (one captured graph per step), so the runs'
work shares every launch.  Between slices it retires finished runs (their
result is read back) and makes prepared arrivals resident: a request is
prepared on an admission thread (host set-up, upload, initial sort) while
the resident runs advance, and joins at the next slice.

Throughput counts the live markers of each run (the mean of its first and
last count) times its steps, over the whole wall time (the offered load)
and over the time some run was resident (the busy rate).

    python -m mandor3_amd.serve --requests 40 --rate 8 --precision fp32

prints one JSON line: p50/p99 time-to-result, aggregate marker-pushes/s,
and the admission statistics.
"""
import argparse
import heapq
import json
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import FP32_SPLIT, FP64, MpicError, Run, batch_advance, lib
from . import workloads

GB = 1 << 30


class Request:
    def __init__(self, rid, params, arrival):
        self.rid = rid
        self.params = params
        self.arrival = arrival
        self.start = None
        self.done = None
        self.run = None
        self.left = int(params["steps"])
        self.markers = 0
        self.footprint = 0
        self.result = None


def footprint(state, precision):
    """Device bytes of a run (the library's own figure)."""
    import ctypes
    from . import Config
    c = Config()
    for a in range(3):
        c.n[a] = state["n"][a]
        c.h[a] = state["h"][a]
        c.L[a] = state["L"][a]
    c.tau = state["tau"]
    for q in range(6):
        c.bc[q] = state["bc"][q]
    c.precision = precision
    c.max_markers = max(int(state["markers"].size), 1)
    out = ctypes.c_size_t(0)
    if lib().mpic_footprint(ctypes.byref(c), ctypes.byref(out)) != 0:
        raise MpicError("mpic_footprint failed")
    return int(out.value)


def start_run(state, precision, graphs=True):
    r = Run(state["n"], state["h"], state["L"], state["tau"], state["bc"], 0.0,
            precision=precision, max_markers=max(int(state["markers"].size), 1))
    z = np.zeros(3 * int(np.prod([v + 1 for v in state["n"]])))
    r.set_fields(z, z)
    r.load_markers(state["markers"])
    if state["laser"] is not None:
        r.add_laser(*state["laser"])
    r.init()
    if graphs:
        r.set_graphs(True)
    return r


def result_of(run, full):
    """What a client gets back: field and kinetic energies and the live
    marker count, plus the fields themselves when asked."""
    run.sync()
    out = {"time": run.time, "markers": run.marker_count()}
    if full:
        out["E"], out["H"] = run.fields()
    return out


class Server:
    """The serving loop.  One admission thread prepares the arrived requests
    in arrival order -- the request's markers and laser map on the host, its
    device footprint, then a reservation of that footprint against the HBM
    budget (it waits while the resident runs leave too little) and the run
    itself (allocation, upload, initial sort, captured step graphs) -- while
    the main thread advances the resident runs slice by slice; a prepared
    run joins at the next slice boundary.  Run set-up and the slices of
    other runs overlap on the device: each run has its own stream."""

    def __init__(self, precision=FP32_SPLIT, budget_gb=240.0, slice_steps=25, full_results=False):
        self.precision = precision
        self.budget = int(budget_gb * GB)
        self.slice = slice_steps
        self.full = full_results
        self.waiting = []          # heap of (arrival, rid, request)
        self.active = []
        self.finished = []
        self.resident = 0          # reserved device bytes (prepared and resident runs)
        self.peak_resident = 0
        self.pushes = 0.0
        self.t0 = None
        self.idle = 0.0
        self._cv = threading.Condition()
        self._ready = queue.Queue()
        self._pending = 0          # requests handed to the admission thread
        self._pool = None
        self._closing = False

    def submit(self, req):
        heapq.heappush(self.waiting, (req.arrival, req.rid, req))

    def _now(self):
        return time.perf_counter() - self.t0

    def _prepare(self, req):
        """Admission thread: set-up, footprint, reservation, run start."""
        try:
            state = workloads.foil_request(**req.params)
            need = footprint(state, self.precision)
            if need > self.budget:
                req.result = {"error": "footprint %.1f GB exceeds the budget" % (need / GB)}
                self._ready.put(req)
                return
            with self._cv:
                while self.resident + need > self.budget and not self._closing:
                    self._cv.wait()                    # until a resident run retires
                if self._closing:
                    return
                self.resident += need
                self.peak_resident = max(self.peak_resident, self.resident)
            req.footprint = need
            req.markers = int(state["markers"].size)
            req.run = start_run(state, self.precision)
            self._ready.put(req)
        except BaseException as e:                     # re-raised by the main thread
            self._ready.put((req, e))

    def _release(self, need):
        with self._cv:
            self.resident -= need
            self._cv.notify_all()

    def _admit(self, block_s=0.0):
        """Hands arrived requests to the admission thread and makes the
        prepared runs resident; waits up to block_s for one to be ready."""
        while self.waiting and self.waiting[0][0] <= self._now():
            req = heapq.heappop(self.waiting)[2]
            self._pending += 1
            self._pool.submit(self._prepare, req)
        while self._pending:
            try:
                item = self._ready.get(timeout=block_s) if block_s > 0 else self._ready.get_nowait()
            except queue.Empty:
                return
            block_s = 0.0
            self._pending -= 1
            if isinstance(item, tuple):
                raise item[1]
            req = item
            if req.result is not None:                 # refused: larger than the budget
                req.done = self._now()
                self.finished.append(req)
                continue
            req.start = self._now()
            self.active.append(req)

    def _retire(self):
        keep = []
        for req in self.active:
            if req.left > 0:
                keep.append(req)
                continue
            req.result = result_of(req.run, self.full)
            req.done = self._now()
            steps = int(req.params["steps"])
            self.pushes += 0.5 * (req.markers + req.result["markers"]) * steps
            req.run.close()
            req.run = None
            self._release(req.footprint)
            self.finished.append(req)
        self.active = keep

    def serve(self):
        self.t0 = time.perf_counter()
        self.idle = 0.0
        self._pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="admission")
        try:
            while self.waiting or self._pending or self.active:
                self._admit()
                if not self.active:
                    t = self._now()
                    nxt = self.waiting[0][0] - t if self.waiting else 0.05
                    if self._pending:
                        self._admit(block_s=min(max(nxt, 0.001), 0.05))
                    elif self.waiting:
                        time.sleep(max(0.0, min(nxt, 0.05)))
                    self.idle += self._now() - t
                    continue
                # one slice for every resident run: the shortest remaining
                # work bounds it so that finished runs leave at the slice
                # boundary; the slice is waited for, so that a run prepared
                # meanwhile joins at the next boundary
                k = min(self.slice, min(r.left for r in self.active))
                batch_advance([r.run for r in self.active], k)
                self.active[0].run.sync()
                for r in self.active:
                    r.left -= k
                if any(r.left == 0 for r in self.active):
                    self._retire()
        finally:
            with self._cv:                             # an admission waiting for memory ends
                self._closing = True
                self._cv.notify_all()
            self._pool.shutdown(wait=True)
            while not self._ready.empty():             # prepared runs never made resident
                item = self._ready.get_nowait()
                if not isinstance(item, tuple) and item.run is not None:
                    item.run.close()
                    item.run = None
        return self.report()

    def report(self):
        lat = np.array([r.done - r.arrival for r in self.finished if "error" not in r.result])
        wall = self._now()
        return {"requests": len(self.finished),
                "rejected": sum(1 for r in self.finished if "error" in r.result),
                "p50_s": float(np.percentile(lat, 50)) if lat.size else None,
                "p99_s": float(np.percentile(lat, 99)) if lat.size else None,
                "marker_pushes_per_s": self.pushes / wall if wall > 0 else 0.0,
                "marker_pushes_per_s_busy": (self.pushes / (wall - self.idle)
                                             if wall > self.idle else 0.0),
                "wall_s": wall,
                "busy_s": wall - self.idle,
                "peak_resident_gb": self.peak_resident / GB}


def poisson_requests(n, rate, seed, large_fraction=0.1, steps=(500, 3000)):
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for rid in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(Request(rid, workloads.scan_request(rng, large_fraction, steps), t))
    return out


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--requests", type=int, default=40)
    p.add_argument("--rate", type=float, default=4.0, help="mean arrivals per second")
    p.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    p.add_argument("--budget-gb", type=float, default=240.0)
    p.add_argument("--slice", type=int, default=25)
    p.add_argument("--steps", type=int, nargs=2, default=[500, 3000])
    p.add_argument("--large-fraction", type=float, default=0.1)
    p.add_argument("--seed", type=int, default=2024)
    a = p.parse_args()
    srv = Server(FP32_SPLIT if a.precision == "fp32" else FP64, a.budget_gb, a.slice)
    for req in poisson_requests(a.requests, a.rate, a.seed, a.large_fraction, tuple(a.steps)):
        srv.submit(req)
    rep = srv.serve()
    rep.update(precision=a.precision, rate=a.rate, budget_gb=a.budget_gb)
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
