"""Serving loop for a stream of LAS tiles of uneven size.

Each request is one tile: int32 coordinates on the host.  A request is
admitted when one of the `slots` workers is free and its estimated device
footprint fits in the memory budget; the worker builds the tile's octree in
HBM (it stays resident until the request completes), runs the
classification pass and copies the class bytes back.  Every worker owns a
wb_ctx and so a HIP stream of its own: requests that arrive while others are
on the GPU start as soon as a slot and memory are free, and their kernels
overlap on the device (continuous batching at stream level; small tiles are
not fused into one launch).

`python -m wolkenbase_amd.serve` runs the load test of the serving workload:
a pool of tiles drawn like the serving stream (lognormal point count, median
2M, density 5-60 pts/m^2), each timed alone, the saturated capacity with all
slots busy, then Poisson arrivals at a fraction of that capacity.  It prints
one JSON line with p50/p99 latency, throughput, the ratio to single-tile
throughput, and the number of requests whose class bytes differ from the
same tile served alone (must be 0).
"""
import argparse
import collections
import json
import math
import sys
import threading
import time

import numpy as np

import wolkenbase_amd as W
from wolkenbase_amd import synth

# device bytes per point at the peak of build + classify: sort buffers,
# coordinates, per-point tile membership, classify scratch (an upper estimate)
BYTES_PER_POINT = 240


class Request:
    def __init__(self, tile, tag=None):
        self.tile = tile
        self.tag = tag
        self.n = len(tile["xi"])
        self.need = BYTES_PER_POINT * self.n + (64 << 20)
        self.done = threading.Event()
        self.t_arrive = self.t_start = self.t_end = None
        self.cls = None
        self.error = None


class Server:
    """policy "fifo": requests start in arrival order; "sjf": the waiting
    request with the fewest points starts first."""

    def __init__(self, slots=4, device=0, mem_fraction=0.8, params=None, policy="fifo"):
        _, total = W.device_memory(device)
        self.budget = int(total * mem_fraction)
        self.used = 0
        self.device = device
        self.params = params or W.default_params()
        self.policy = policy
        self.q = collections.deque()
        self.cv = threading.Condition()
        self.stopping = False
        self.failed = None
        # the contexts are opened here, so a device that cannot be opened
        # fails the constructor instead of a worker thread
        ctxs = []
        try:
            for _ in range(slots):
                ctxs.append(W.Context(device))
        except BaseException:
            for c in ctxs:
                c.close()
            raise
        self.workers = [threading.Thread(target=self._run, args=(c,), daemon=True) for c in ctxs]
        for w in self.workers:
            w.start()

    def submit(self, req):
        req.t_arrive = time.perf_counter()
        with self.cv:
            if self.failed is None:
                self.q.append(req)
                self.cv.notify_all()
                return req
            err = self.failed
        req.error = RuntimeError(f"serving loop failed: {err!r}")
        req.done.set()
        return req

    def _fail(self, err):
        """A worker died: no request may wait for it.  Queued requests and
        every later submit fail with the error."""
        with self.cv:
            if self.failed is None:
                self.failed = err
            self.stopping = True
            pending = list(self.q)
            self.q.clear()
            self.cv.notify_all()
        for r in pending:
            r.error = RuntimeError(f"serving loop failed: {err!r}")
            r.done.set()

    def close(self):
        with self.cv:
            self.stopping = True
            self.cv.notify_all()
        for w in self.workers:
            w.join()

    def _next(self):
        # the chosen request waits until its memory fits (a request larger
        # than the budget runs alone)
        with self.cv:
            while True:
                if self.q:
                    k = 0 if self.policy == "fifo" else min(range(len(self.q)),
                                                            key=lambda j: self.q[j].n)
                    need = self.q[k].need
                    if self.used == 0 or self.used + need <= self.budget:
                        req = self.q[k]
                        del self.q[k]
                        self.used += need
                        return req
                elif self.stopping:
                    return None
                self.cv.wait()

    def _run(self, ctx):
        req = None
        try:
            while True:
                req = self._next()
                if req is None:
                    return
                req.t_start = time.perf_counter()
                try:
                    req.cls = classify_tile(ctx, req.tile, self.params)
                except Exception as e:  # reported with the request
                    req.error = e
                req.t_end = time.perf_counter()
                with self.cv:
                    self.used -= req.need
                    self.cv.notify_all()
                req.done.set()
                req = None
        except BaseException as e:
            if req is not None and not req.done.is_set():
                req.error = e
                req.done.set()
            self._fail(e)
        finally:
            ctx.close()


def classify_tile(ctx, tile, params):
    """Build the tile's octree and classify it; returns class bytes (input order)."""
    center, side = W.Octree.size_fit(tile["bbmin"], tile["bbmax"])
    tree = W.Octree.build(ctx, tile["xi"], tile["yi"], tile["zi"], tile["scale"],
                          tile["offset"], center, side)
    try:
        cls, _, _ = W.classify(tree, tile["bbmin"], tile["bbmax"], params, want_tiles=False)
    finally:
        tree.free()
    return cls


def make_pool(ntiles, seed, max_points):
    """Tiles of the serving stream: tile k is synth.serving_tile(k) (its own
    seed), with the point count capped at max_points for the load test."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(ntiles):
        p = synth.serving_tile(k, rng)
        n = p["extent"] ** 2 * p["density"]
        if n > max_points:
            p["extent"] = math.sqrt(max_points / p["density"])
        pool.append(synth.generate(**p))
    return pool


def load_test(pool, slots, load, nreq, seed, device=0, policy="fifo"):
    ctx = W.Context(device)
    params = W.default_params()
    # each tile alone: warm-up, then timed
    classify_tile(ctx, pool[0], params)
    iso_t, iso_cls = [], []
    for tile in pool:
        t0 = time.perf_counter()
        iso_cls.append(classify_tile(ctx, tile, params))
        iso_t.append(time.perf_counter() - t0)
    ctx.close()
    npts = np.array([len(t["xi"]) for t in pool], dtype=np.float64)
    single = npts.sum() / sum(iso_t)
    srv = Server(slots=slots, device=device, params=params, policy=policy)
    try:
        # capacity: every pool tile submitted at once
        t0 = time.perf_counter()
        reqs = [srv.submit(Request(t, k)) for k, t in enumerate(pool)]
        for r in reqs:
            r.done.wait()
        capacity = npts.sum() / (time.perf_counter() - t0)
        # open loop: Poisson arrivals at `load` x capacity, tiles drawn from the pool
        rng = np.random.default_rng(seed)
        picks = rng.integers(0, len(pool), nreq)
        rate = load * capacity / npts.mean()  # requests per second
        gaps = rng.exponential(1.0 / rate, nreq)
        reqs = []
        t_next = time.perf_counter()
        for k, g in zip(picks, gaps):
            t_next += g
            while True:
                d = t_next - time.perf_counter()
                if d <= 0:
                    break
                time.sleep(min(d, 0.002))
            reqs.append(srv.submit(Request(pool[k], int(k))))
        for r in reqs:
            r.done.wait()
    finally:
        srv.close()
    errors = [r for r in reqs if r.error is not None]
    if errors:
        raise errors[0].error
    lat = np.array([r.t_end - r.t_arrive for r in reqs])
    ratio = np.array([(r.t_end - r.t_arrive) / iso_t[r.tag] for r in reqs])
    span = max(r.t_end for r in reqs) - min(r.t_arrive for r in reqs)
    served = sum(r.n for r in reqs) / span
    mismatched = sum(int(not np.array_equal(r.cls, iso_cls[r.tag])) for r in reqs)
    return {
        "metric": "serving: classified points/s and p50/p99 tile latency",
        "slots": slots, "policy": policy, "load": load, "requests": nreq,
        "pool_tiles": len(pool),
        "pool_points": {"min": int(npts.min()), "median": int(np.median(npts)),
                        "max": int(npts.max())},
        "single_tile_points_per_s": round(single, 1),
        "capacity_points_per_s": round(capacity, 1),
        "served_points_per_s": round(served, 1),
        "served_over_single": round(served / single, 3),
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "p50_over_isolated": round(float(np.percentile(ratio, 50)), 3),
        "p99_over_isolated": round(float(np.percentile(ratio, 99)), 3),
        "isolated_ms": {"min": round(1e3 * min(iso_t), 3), "max": round(1e3 * max(iso_t), 3)},
        "mismatched_results": mismatched,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description="serving load test")
    ap.add_argument("--pool", type=int, default=24, help="distinct tiles")
    ap.add_argument("--max-points", type=float, default=12e6, help="cap on a pool tile")
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--load", type=float, default=0.7, help="arrival rate / capacity")
    ap.add_argument("--requests", type=int, default=200)
    ap.add_argument("--seed", type=int, default=3000)
    ap.add_argument("--policy", choices=("fifo", "sjf"), default="fifo")
    args = ap.parse_args(argv)
    pool = make_pool(args.pool, args.seed, args.max_points)
    res = load_test(pool, args.slots, args.load, args.requests, args.seed, policy=args.policy)
    print(json.dumps(res), flush=True)
    return 0 if res["mismatched_results"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
