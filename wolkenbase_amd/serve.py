"""Serving loop for a stream of LAS tiles of uneven size.

Each request is one tile: int32 coordinates in host memory (numpy arrays,
pinned receive buffers from wolkenbase_amd.pinned_empty, which the copies to
the device read by DMA without staging, or host tensors).  A request is admitted when a worker
is free and its estimated device footprint fits in the memory budget; the
worker builds the tile's octree in HBM (it stays resident until the request
completes), runs the classification pass and copies the class bytes back.

Continuous batching, at two levels:
  * every worker owns a wb_ctx and so a HIP stream: requests that arrive
    while others are on the GPU start as soon as a worker and memory are
    free, and their kernels overlap on the device;
  * small tiles are batched: a worker that takes a small request also takes
    the small requests queued behind it (up to `batch_points` points and
    `max_batch` tiles), builds their octrees and runs their ground tests in
    one launch of the ground kernel (wb_classify_batch).  A batch that fails
    is retried tile by tile, so a bad tile fails only its own request.

Scheduling: the next request is the one with the highest response ratio
(waiting time + estimated service time) / estimated service time, so a small
tile does not wait long behind large ones and a large one is not starved;
`small_slots` workers take only small tiles, so one is free for a small tile
while large ones run.  `policy="fifo"` takes requests in arrival order.

`python -m wolkenbase_amd.serve` runs the load test of the serving workload
(see load_test) and prints one JSON line.
"""
import argparse
import collections
import json
import math
import os
import statistics
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import wolkenbase_amd as W
from wolkenbase_amd import synth

# device bytes per point at the peak of build + classify: sort buffers,
# coordinates, per-point tile membership, classify scratch (an upper estimate)
BYTES_PER_POINT = 240
# service-time model for the response ratio: a fixed cost per tile (launches,
# host waits, the slowest wave of the ground kernel) plus a cost per point
SERVICE_FIXED_S = 2.0e-3
SERVICE_PER_POINT_S = 2.2e-9


def service_estimate(n):
    return SERVICE_FIXED_S + SERVICE_PER_POINT_S * n


class Request:
    def __init__(self, tile, tag=None):
        self.tile = tile
        self.tag = tag
        self.n = len(tile["xi"])
        self.need = BYTES_PER_POINT * self.n + (64 << 20)
        self.est = service_estimate(self.n)
        self.done = threading.Event()
        self.t_arrive = self.t_start = self.t_end = None
        self.batch = 1
        self.times = {}
        self.cls = None
        self.error = None


class Server:
    """slots: workers (each a context and stream); small_slots of them take
    only requests of at most small_points points.  policy "ratio" (highest
    response ratio next), "fifo" (arrival order) or "sjf" (fewest points
    first).  sharing: the workers' contexts in sharing mode
    (wb_ctx_set_sharing)."""

    def __init__(self, slots=4, device=0, mem_fraction=0.8, params=None, policy="ratio",
                 small_slots=1, small_points=1_500_000, batch_points=4_000_000, max_batch=16,
                 sharing=False):
        if policy not in ("ratio", "fifo", "sjf"):
            raise ValueError(f"unknown policy {policy!r}")
        _, total = W.device_memory(device)
        self.budget = int(total * mem_fraction)
        self.used = 0
        self.device = device
        self.params = params or W.default_params()
        self.policy = policy
        self.small_points = small_points
        self.batch_points = batch_points
        self.max_batch = max(1, int(max_batch))
        # at least one worker takes requests of any size
        self.small_slots = max(0, min(small_slots, slots - 1))
        self.q = collections.deque()
        self.cv = threading.Condition()
        self.stopping = False
        self.failed = None
        # the contexts are opened here, so a device that cannot be opened
        # fails the constructor instead of a worker thread
        ctxs = []
        try:
            for _ in range(slots):
                # sharing: the workers' long kernels leave room on every CU,
                # so a small tile's kernels are not held behind one large
                # tile's (with several large tiles at once the device fills
                # again)
                ctxs.append(W.Context(device, sharing=True) if sharing else W.Context(device))
        except BaseException:
            for c in ctxs:
                c.close()
            raise
        self.workers = [threading.Thread(target=self._run, args=(c, k < self.small_slots),
                                         daemon=True)
                        for k, c in enumerate(ctxs)]
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

    def _order(self, idx, now):
        q = self.q
        if self.policy == "fifo":
            return idx
        if self.policy == "sjf":
            return sorted(idx, key=lambda j: q[j].n)
        return sorted(idx, key=lambda j: -((now - q[j].t_arrive) + q[j].est) / q[j].est)

    def _fits(self, need):
        return self.used == 0 or self.used + need <= self.budget

    def _next(self, small_only):
        """The next batch of requests for a worker (a list; empty when the
        server stops).  The first request waits until its memory fits (a
        request larger than the budget runs alone); small requests queued
        behind a small first one join it while they fit."""
        with self.cv:
            while True:
                now = time.perf_counter()
                idx = [j for j in range(len(self.q))
                       if not small_only or self.q[j].n <= self.small_points]
                order = self._order(idx, now)
                if order:
                    first = self.q[order[0]]
                    if self._fits(first.need):
                        take = [order[0]]
                        used, pts = first.need, first.n
                        if first.n <= self.small_points:
                            for j in order[1:]:
                                r = self.q[j]
                                if len(take) >= self.max_batch:
                                    break
                                if (r.n > self.small_points or pts + r.n > self.batch_points
                                        or self.used + used + r.need > self.budget):
                                    continue
                                take.append(j)
                                used += r.need
                                pts += r.n
                        reqs = [self.q[j] for j in take]
                        for j in sorted(take, reverse=True):
                            del self.q[j]
                        self.used += used
                        return reqs
                if self.stopping and not self.q:
                    return []
                self.cv.wait(0.05 if order else None)

    def _run(self, ctx, small_only):
        reqs = []
        try:
            while True:
                reqs = self._next(small_only)
                if not reqs:
                    return
                t0 = time.perf_counter()
                for r in reqs:
                    r.t_start = t0
                    r.batch = len(reqs)
                if len(reqs) == 1:
                    try:
                        reqs[0].cls = classify_tile(ctx, reqs[0].tile, self.params, reqs[0].times)
                    except Exception as e:  # reported with the request
                        reqs[0].error = e
                else:
                    try:
                        out = classify_tiles(ctx, [r.tile for r in reqs], self.params)
                        for r, c in zip(reqs, out):
                            r.cls = c
                    except Exception:
                        # a bad tile fails only its own request
                        for r in reqs:
                            try:
                                r.cls = classify_tile(ctx, r.tile, self.params, r.times)
                            except Exception as e:
                                r.error = e
                t1 = time.perf_counter()
                with self.cv:
                    self.used -= sum(r.need for r in reqs)
                    self.cv.notify_all()
                for r in reqs:
                    r.t_end = t1
                    r.done.set()
                reqs = []
        except BaseException as e:
            for r in reqs:
                if not r.done.is_set():
                    r.error = e
                    r.done.set()
            self._fail(e)
        finally:
            ctx.close()


def _build(ctx, tile):
    center, side = W.Octree.size_fit(tile["bbmin"], tile["bbmax"])
    return W.Octree.build(ctx, tile["xi"], tile["yi"], tile["zi"], tile["scale"],
                          tile["offset"], center, side)


def classify_tile(ctx, tile, params, times=None):
    """Build the tile's octree and classify it; returns class bytes (input
    order).  times: a dict that receives the wall time of the build call and
    of the classification call (s)."""
    t0 = time.perf_counter()
    tree = _build(ctx, tile)
    t1 = time.perf_counter()
    try:
        cls, _, _ = W.classify(tree, tile["bbmin"], tile["bbmax"], params, want_tiles=False,
                               out=np.empty(tree.n_input, dtype=np.uint8))
    finally:
        tree.free()
    if times is not None:
        times["build"] = t1 - t0
        times["classify"] = time.perf_counter() - t1
    return cls


def classify_tiles(ctx, tiles, params):
    """Several tiles on one context: their octrees built one after another,
    their classification passes run as one batch (the ground tests of all
    of them in one kernel launch).  Returns class bytes per tile."""
    trees = []
    try:
        for t in tiles:
            trees.append(_build(ctx, t))
        outs = [np.empty(tr.n_input, dtype=np.uint8) for tr in trees]
        cls, _ = W.classify_batch(trees, [t["bbmin"] for t in tiles],
                                  [t["bbmax"] for t in tiles], params, outs=outs)
    finally:
        for tr in trees:
            tr.free()
    return cls


# ---------------------------------------------------------------------------
# the serving workload and its load test
# ---------------------------------------------------------------------------
def stream_specs(ntiles, seed=3000, max_points=4e7):
    """Generator parameters of a pool of tiles that stands for the serving
    stream: tile k has the point count of the lognormal's quantile
    (k + 1/2)/ntiles (median 2M, sigma 1, clipped to [50k, max_points]), a
    density drawn from U[5, 60] pts/m^2 and its own seed 3,000,000 + k (the
    generator's buildings and trees, no embankments or outliers).  A stream
    that draws tiles uniformly from the pool has the lognormal's point-count
    distribution, in ntiles strata."""
    rng = np.random.default_rng(seed)
    nd = statistics.NormalDist(math.log(2e6), 1.0)
    specs = []
    for k in range(ntiles):
        n = float(np.clip(math.exp(nd.inv_cdf((k + 0.5) / ntiles)), 5e4, max_points))
        dens = float(rng.uniform(5, 60))
        specs.append(dict(seed=3000000 + k, extent=math.sqrt(n / dens), density=dens,
                          buildings=40, trees=300, embankments=0, outliers=0.0, linear=0.0))
    return specs


def stream_pool(ntiles, seed=3000, max_points=4e7, threads=16):
    """The tiles of stream_specs, generated on `threads` host threads."""
    specs = stream_specs(ntiles, seed, max_points)
    with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
        return list(ex.map(lambda p: synth.generate(**p), specs))


def make_pool(ntiles, seed, max_points):
    """Tiles of the serving stream drawn at random: tile k is
    synth.serving_tile(k) (its own seed), with the point count capped at
    max_points."""
    rng = np.random.default_rng(seed)
    pool = []
    for k in range(ntiles):
        p = synth.serving_tile(k, rng)
        n = p["extent"] ** 2 * p["density"]
        if n > max_points:
            p["extent"] = math.sqrt(max_points / p["density"])
        pool.append(synth.generate(**p))
    return pool


def pin_tiles(pool):
    """The pool's coordinates copied into pinned host memory (the receive
    buffers of a server: copies to the device read them by DMA)."""
    out = []
    for t in pool:
        p = dict(t)
        for k in ("xi", "yi", "zi"):
            a = W.pinned_empty(len(t[k]), np.int32)
            a[:] = t[k]
            p[k] = a
        out.append(p)
    return out


def load_test(pool, slots, load, nreq, seed, device=0, policy="ratio", small_slots=1,
              small_points=1_500_000, batch_points=4_000_000, ref_tile=None, sharing=False):
    """Each pool tile timed alone (one context, nothing else on the device),
    the saturated capacity (every pool tile submitted at once), then `nreq`
    requests drawn uniformly from the pool arriving as a Poisson process at
    `load` x capacity.  ref_tile: a tile whose isolated rate is the
    single-tile throughput the served rate is compared with (by default the
    pool's largest tile).  Every served result is compared with the same
    tile's isolated result."""
    ctx = W.Context(device)
    params = W.default_params()
    # each tile alone: warm-up, then timed
    classify_tile(ctx, pool[0], params)
    iso_t, iso_cls = [], []
    for tile in pool:
        t0 = time.perf_counter()
        iso_cls.append(classify_tile(ctx, tile, params))
        iso_t.append(time.perf_counter() - t0)
    npts = np.array([len(t["xi"]) for t in pool], dtype=np.float64)
    if ref_tile is None:
        ref_tile = pool[int(np.argmax(npts))]
    classify_tile(ctx, ref_tile, params)
    ref_t = []
    for _ in range(3):
        t0 = time.perf_counter()
        classify_tile(ctx, ref_tile, params)
        ref_t.append(time.perf_counter() - t0)
    ref_n = len(ref_tile["xi"])
    ref_s = float(np.median(ref_t))
    ctx.close()
    single_pool = npts.sum() / sum(iso_t)
    single_ref = ref_n / ref_s
    srv = Server(slots=slots, device=device, params=params, policy=policy,
                 small_slots=small_slots, small_points=small_points, batch_points=batch_points,
                 sharing=sharing)
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
    batches = np.array([r.batch for r in reqs])
    # the slowest 1% of requests by slowdown: their size, queueing and run time
    worst = sorted(reqs, key=lambda r: -(r.t_end - r.t_arrive) / iso_t[r.tag])
    worst = worst[:max(1, len(reqs) // 100)]
    return {
        "metric": "serving: classified points/s and p50/p99 tile latency",
        "slots": slots, "small_slots": srv.small_slots, "policy": policy, "load": load,
        "sharing": bool(sharing),
        "requests": nreq, "pool_tiles": len(pool),
        "pool_points": {"min": int(npts.min()), "median": int(np.median(npts)),
                        "max": int(npts.max()), "mean": int(npts.mean())},
        "requested_points": int(sum(r.n for r in reqs)),
        "single_tile": {"points": ref_n, "ms": round(1e3 * ref_s, 3),
                        "points_per_s": round(single_ref, 1)},
        "single_pool_points_per_s": round(single_pool, 1),
        "capacity_points_per_s": round(capacity, 1),
        "served_points_per_s": round(served, 1),
        "served_over_single_tile": round(served / single_ref, 3),
        "served_over_single_pool": round(served / single_pool, 3),
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "p99_over_single_tile_ms": round(float(np.percentile(lat, 99)) / ref_s, 3),
        "p50_over_isolated": round(float(np.percentile(ratio, 50)), 3),
        "p99_over_isolated": round(float(np.percentile(ratio, 99)), 3),
        "isolated_ms": {"min": round(1e3 * min(iso_t), 3), "max": round(1e3 * max(iso_t), 3)},
        "slowest_1pct": [{"points": r.n, "isolated_ms": round(1e3 * iso_t[r.tag], 2),
                          "wait_ms": round(1e3 * (r.t_start - r.t_arrive), 2),
                          "run_ms": round(1e3 * (r.t_end - r.t_start), 2), "batch": r.batch,
                          **{k + "_ms": round(1e3 * v, 2) for k, v in r.times.items()}}
                         for r in worst],
        # the requests at or above the 99th latency percentile: their tile
        # sizes, waits and run times
        "p99_requests": [{"points": r.n, "isolated_ms": round(1e3 * iso_t[r.tag], 2),
                          "latency_ms": round(1e3 * (r.t_end - r.t_arrive), 2),
                          "wait_ms": round(1e3 * (r.t_start - r.t_arrive), 2)}
                         for r in sorted(reqs, key=lambda r: -(r.t_end - r.t_arrive))
                         [:max(1, len(reqs) // 100)]],
        "batched_requests": int((batches > 1).sum()),
        "mean_batch": round(float(batches.mean()), 2),
        "mismatched_results": mismatched,
    }


# the tile whose isolated rate is the single-tile throughput served
# throughput is compared with: 20M pulses of the stream's generator at
# 25 pts/m^2
REF_TILE = dict(seed=3999999, extent=math.sqrt(20e6 / 25.0), density=25.0, buildings=40,
                trees=300, embankments=0, outliers=0.0, linear=0.0)


def main(argv=None):
    ap = argparse.ArgumentParser(description="serving load test")
    ap.add_argument("--pool", type=int, default=128, help="distinct tiles (lognormal strata)")
    ap.add_argument("--max-points", type=float, default=4e7, help="cap on a pool tile")
    # 6 workers, 2 of them for small tiles: the shortest p99 measured (4
    # workers: the same capacity, a longer p99 from mid-size tiles waiting
    # for a worker; 8: longer again)
    ap.add_argument("--slots", type=int, default=6)
    ap.add_argument("--small-slots", type=int, default=2)
    ap.add_argument("--small-points", type=float, default=1.5e6)
    ap.add_argument("--batch-points", type=float, default=4e6)
    ap.add_argument("--load", type=float, default=0.7, help="arrival rate / capacity")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=3000)
    ap.add_argument("--policy", choices=("ratio", "fifo", "sjf"), default="ratio")
    ap.add_argument("--pageable", action="store_true",
                    help="keep the tiles in pageable host memory instead of pinned buffers")
    ap.add_argument("--sharing", action="store_true",
                    help="workers' contexts in sharing mode (resident, capped long kernels)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="hardware queues per process (GPU_MAX_HW_QUEUES; 0 keeps the "
                         "environment's)")
    args = ap.parse_args(argv)
    if args.hw_queues > 0:
        # read when the HIP runtime starts, so before anything touches the device
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    pool = stream_pool(args.pool, args.seed, args.max_points)
    ref = synth.generate(**REF_TILE)
    if not args.pageable:
        pool = pin_tiles(pool)
        ref = pin_tiles([ref])[0]
    res = load_test(pool, args.slots, args.load, args.requests, args.seed, policy=args.policy,
                    small_slots=args.small_slots, small_points=int(args.small_points),
                    batch_points=int(args.batch_points), ref_tile=ref, sharing=args.sharing)
    res["host_buffers"] = "pageable" if args.pageable else "pinned"
    res["hw_queues"] = os.environ.get("GPU_MAX_HW_QUEUES", "runtime default")
    print(json.dumps(res), flush=True)
    return 0 if res["mismatched_results"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
