#!/usr/bin/env python3
"""Benchmark: classified points/s on the 20M-point tile, one MI355X.

A step is one pass of the hot path over the tile: octree build from the
int32 LAS coordinates already resident in HBM (Morton keys, radix sort,
canonical split, SoA permute) followed by SCAN, POSTSCAN and CLASSIFY of
every flowsnake tile, ending with the class bytes in HBM.

Prints one JSON line (rank 0).  --gpus above 1 is refused: the build runs on
one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("classified points/s on 20M-pt tile (1 GPU, %HBM roofline); p50/p99 tile latency")
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "tile20m": ("tile20m: 900 m x 900 m, 25 pulses/m^2 (~20.3M points), seed 2000, "
                "6-sinusoid ground + 5 embankments (20-45 deg) + 2% outliers to +-50 m, "
                "box buildings and cone trees (1-3 returns), int32 at 0.001 m"),
    "tile20m_high": ("tile20m_high: the tile20m pulses, terrain, buildings and trees with its "
                     "2% outliers all 1-50 m above the surface instead of half of them below "
                     "(about 55% ground instead of 0.3%), int32 at 0.001 m"),
    "tile200m": ("tile200m: 3 km x 3 km, 17 pulses/m^2 (~205M points), seed 4000, an urban "
                 "half (60 buildings per 200 m x 200 m) beside a forested half (1200 trees per "
                 "200 m x 200 m, 44% multi-return pulses), int32 at 0.001 m"),
    "tile1m": ("tile1m: 200 m x 200 m, 25 pulses/m^2 (~1.2M points), seed 1000, "
               "6-sinusoid ground, 40 box buildings, 300 cone trees, int32 at 0.001 m"),
    "las20m": ("las20m: the tile20m points as one LAS 1.4 point-format-1 file (28-byte "
               "records) resident in HBM; a step decodes the records (READ-phase zero rule), "
               "builds the octree with the store's duplicate merge (full-key sort), runs "
               "SCAN/POSTSCAN/CLASSIFY and plans and encodes the class-separated output "
               "files (10M points per file) in HBM"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def percentile(v, q):
    v = sorted(v)
    if not v:
        return None
    k = (len(v) - 1) * q / 100.0
    f = math.floor(k)
    c = min(f + 1, len(v) - 1)
    return v[f] + (v[c] - v[f]) * (k - f)


def cpu_baseline(cloud, window):
    """The oracle (single-threaded C restatement of the reference pass) on the
    points of a window x window m square at the tile centre."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    x = cloud["offset"][0] + cloud["scale"][0] * cloud["xi"].astype(np.float64)
    y = cloud["offset"][1] + cloud["scale"][1] * cloud["yi"].astype(np.float64)
    cx = 0.5 * (cloud["bbmin"][0] + cloud["bbmax"][0])
    cy = 0.5 * (cloud["bbmin"][1] + cloud["bbmax"][1])
    m = (np.abs(x - cx) < window / 2) & (np.abs(y - cy) < window / 2)
    sub = {k: cloud[k][m] for k in ("xi", "yi", "zi")}
    for k in ("scale", "offset"):
        sub[k] = cloud[k]
    z = cloud["offset"][2] + cloud["scale"][2] * sub["zi"].astype(np.float64)
    sub["bbmin"] = np.array([x[m].min(), y[m].min(), z.min()])
    sub["bbmax"] = np.array([x[m].max(), y[m].max(), z.max()])
    n = int(m.sum())
    t0 = time.perf_counter()
    O.run(sub)
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="points/s", cores=1, kind="port",
                sample=f"oracle full pass (build, scan, postscan, classify) over the {n} points "
                       f"of a {window:g} m x {window:g} m window at the tile centre, "
                       f"{dt:.1f} s single-threaded")


def cpu_baseline_las(cloud, window):
    """The oracle's LAS pipeline (decode, duplicate merge, build, scan,
    postscan, classify, class-separated encode) on the points of a window x
    window m square at the tile centre, written as a LAS file first."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_las as OL
    from wolkenbase_amd import synth
    x = cloud["offset"][0] + cloud["scale"][0] * cloud["xi"].astype(np.float64)
    y = cloud["offset"][1] + cloud["scale"][1] * cloud["yi"].astype(np.float64)
    cx = 0.5 * (cloud["bbmin"][0] + cloud["bbmax"][0])
    cy = 0.5 * (cloud["bbmin"][1] + cloud["bbmax"][1])
    m = (np.abs(x - cx) < window / 2) & (np.abs(y - cy) < window / 2)
    sub = {k: cloud[k][m] for k in ("xi", "yi", "zi", "intensity", "returns")}
    for k in ("scale", "offset"):
        sub[k] = cloud[k]
    z = cloud["offset"][2] + cloud["scale"][2] * sub["zi"].astype(np.float64)
    sub["bbmin"] = np.array([x[m].min(), y[m].min(), z.min()])
    sub["bbmax"] = np.array([x[m].max(), y[m].max(), z.max()])
    img = synth.las_image(sub, 1)
    n = int(m.sum())
    t0 = time.perf_counter()
    OL.pipeline([img])
    dt = time.perf_counter() - t0
    return dict(value=n / dt, unit="points/s", cores=1, kind="port",
                sample=f"oracle LAS pipeline (record decode, duplicate merge, build, scan, "
                       f"postscan, classify, class-separated encode) over the {n} points of a "
                       f"{window:g} m x {window:g} m window at the tile centre, {dt:.1f} s "
                       f"single-threaded")


def main_las(args):
    """The las20m workload: the LAS path around the classification."""
    import numpy as np
    import torch
    import wolkenbase_amd as W
    from wolkenbase_amd import las as WL
    from wolkenbase_amd import synth

    t0 = time.perf_counter()
    cloud = synth.preset("tile20m")
    img = synth.las_image(cloud, 1)
    log(f"[bench] las20m image: {len(img) / 2**20:.0f} MiB in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    torch.cuda.synchronize()
    ctx = W.Context(0)
    las = WL.LasFile(ctx, image=img)
    n = las.number_points()
    rlen = las.header["point_length"]
    cls = torch.empty(n, dtype=torch.uint8, device=dev)

    def step(dry=True, name="/tmp/wb_las20m"):
        cl = WL.Cloud(ctx, [las], zeros="read")
        tree = cl.build()
        _, _, st = cl.classify(tree, want_tiles=False, out=cls)
        ws, names = cl.write(tree, cls, name, separate_classes=True, points_per_file=10000000,
                             dry=dry)
        d = dict(ms_decode=cl.ms_decode, ms_build=st["ms_build"], ms_sort=st["ms_sort"],
                 ms_tiles=st["ms_tiles"], ms_scan=st["ms_scan"], ms_classify=st["ms_classify"],
                 ms_classify_kernel=st["ms_classify_kernel"],
                 ms_heavy_kernel=st["ms_heavy_kernel"], ms_out_classes=ws["ms_classes"],
                 ms_plan=ws["ms_plan"], ms_encode=ws["ms_encode"], ms_copy=ws["ms_copy"],
                 ms_disk=ws["ms_disk"], ndup=tree.ndup, nfiles=ws["nfiles"],
                 bytes_out=ws["bytes"], classes=ws["class_totals"], names=names)
        tree.free()
        cl.close()
        return d

    for w in range(args.warmup):
        d = step()
        log(f"[bench] warmup {w + 1}/{args.warmup}: decode {d['ms_decode']:.2f} ms, "
            f"encode {d['ms_encode']:.2f} ms")
    torch.cuda.synchronize()
    lat, stats = [], []
    t_start = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        stats.append(step())
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
        log(f"[bench] step {k + 1}/{args.steps}: {lat[-1]:.2f} ms")
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    value = n * args.steps / elapsed

    def avg(f):
        return sum(s[f] for s in stats) / len(stats)

    # outside the timed region: the file read from disk (disk + PCIe), and the
    # class-separated files copied back and written (PCIe + disk)
    path = "/tmp/wb_las20m_in.las"
    with open(path, "wb") as f:
        f.write(img)
    t1 = time.perf_counter()
    las_disk = WL.LasFile(ctx, path=path)
    ms_open_disk = (time.perf_counter() - t1) * 1e3
    las_disk.close()
    os.remove(path)
    t1 = time.perf_counter()
    las_mem = WL.LasFile(ctx, image=img)
    ms_open_mem = (time.perf_counter() - t1) * 1e3
    las_mem.close()
    full = step(dry=False)
    for nm in full["names"]:
        os.remove(nm)
    st = stats[-1]
    store = None
    if args.store:
        # the store's blocks of the classified tree (OctBuffer::write of every
        # leaf) into pinned host memory: the spill tier's write rate
        cl = WL.Cloud(ctx, [las], zeros="read")
        tree = cl.build()
        cl.classify(tree, want_tiles=False, out=cls)
        nl = tree.info()["nleaves"]
        host = torch.empty(nl * W.STORE_BLOCK, dtype=torch.uint8, pin_memory=True)
        tree.store_blocks(cloud=cl, cls=cls, out=host)  # warm
        reps = 3
        t1 = time.perf_counter()
        for _ in range(reps):
            tree.store_blocks(cloud=cl, cls=cls, out=host)
        ms = (time.perf_counter() - t1) * 1e3 / reps
        store = {"blocks": nl, "bytes": nl * W.STORE_BLOCK, "points": tree.n,
                 "ms_pack_and_copy_to_pinned_host": round(ms, 2),
                 "gbs_to_host": round(nl * W.STORE_BLOCK / (ms * 1e-3) / 1e9, 1)}
        log(f"[bench] store blocks: {nl} blocks in {ms:.2f} ms")
        del host
        tree.free()
        cl.close()
    kern = {"k_ground": avg("ms_classify_kernel"), "k_ground_heavy": avg("ms_heavy_kernel"),
            "radix_sort": avg("ms_sort"), "k_las_decode": avg("ms_decode"),
            "k_las_encode": avg("ms_encode")}
    dominant = max(kern, key=kern.get)
    # algorithmic bytes: decode reads each record and writes x, y, z and the
    # record index; encode reads the stored point's record and class and
    # writes its output record; the ground test 13 B per point; the sort 24 B
    # per point and pass plus the 8-B key read of each histogram (8 passes)
    nd = n - st["ndup"]
    alg = {"k_ground": 13.0 * nd, "k_ground_heavy": 13.0 * nd, "radix_sort": 32.0 * 8 * n,
           "k_las_decode": (rlen + 28.0) * n, "k_las_encode": (rlen + 1.0 + 28.0) * nd}
    achieved = alg[dominant] / (kern[dominant] * 1e-3) / 1e9

    def kgbs(k):
        return dict(ms=round(kern[k], 3), algorithmic_bytes=alg[k],
                    gbs=round(alg[k] / (kern[k] * 1e-3) / 1e9, 1),
                    frac=round(alg[k] / (kern[k] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS["las20m"], "n_points": n, "record_bytes": rlen,
                   "input": "LAS records resident in HBM; output file images left in HBM"},
        "p50_ms": round(percentile(lat, 50), 3),
        "p99_ms": round(percentile(lat, 99), 3),
        "roofline": dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                         frac=round(achieved / HBM_PEAK_GBS, 6), traffic=None, kernel=dominant,
                         kernel_ms=round(kern[dominant], 3),
                         algorithmic_bytes_per_launch=alg[dominant]),
        "las_kernels": {"k_las_decode": kgbs("k_las_decode"),
                        "k_las_encode": kgbs("k_las_encode")},
        "breakdown_ms": {k: round(avg(k), 3) for k in
                         ("ms_decode", "ms_build", "ms_sort", "ms_tiles", "ms_scan",
                          "ms_classify", "ms_classify_kernel", "ms_heavy_kernel",
                          "ms_out_classes", "ms_plan", "ms_encode")},
        "io_outside_timed_ms": {"open_from_disk": round(ms_open_disk, 1),
                                "open_from_host_memory": round(ms_open_mem, 1),
                                "write_copy_d2h": round(full["ms_copy"], 1),
                                "write_disk": round(full["ms_disk"], 1),
                                "bytes_in": len(img), "bytes_out": full["bytes_out"]},
        "duplicates": st["ndup"],
        "files": st["nfiles"],
        "store_outside_timed": store,
        "classes": {str(k): v for k, v in st["classes"].items()},
    }
    if args.workload == "tile20m" and not args.no_serving:
        out["serving"] = serving_summary(args.serving_requests)
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline on a {args.cpu_window:g} m window ...")
        out["cpu_baseline"] = cpu_baseline_las(cloud, args.cpu_window)
    else:
        out["cpu_baseline"] = None
    las.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def run_tile(name, steps, warmup, ctx, pmc_calibrate=False):
    """Generates workload `name`, puts its int32 x/y/z in HBM and times
    `steps` build + classify steps after `warmup` untimed ones."""
    import torch

    import wolkenbase_amd as W
    from wolkenbase_amd import synth

    t0 = time.perf_counter()
    cloud = synth.preset(name)
    n = len(cloud["xi"])
    log(f"[bench] generated {name}: {n} points in {time.perf_counter() - t0:.1f} s")
    dev = torch.device("cuda:0")
    xi = torch.from_numpy(cloud["xi"]).to(dev)
    yi = torch.from_numpy(cloud["yi"]).to(dev)
    zi = torch.from_numpy(cloud["zi"]).to(dev)
    cls = torch.empty(n, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    if pmc_calibrate:
        a = torch.empty(1 << 28, dtype=torch.uint8, device=dev).fill_(1)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    params = W.default_params()

    def step():
        tree = W.Octree.build(ctx, xi, yi, zi, cloud["scale"], cloud["offset"], center, side)
        _, _, st = W.classify(tree, cloud["bbmin"], cloud["bbmax"], params, want_tiles=False,
                              out=cls)
        tree.free()
        return st

    for w in range(warmup):
        st = step()
        log(f"[bench] {name} warmup {w + 1}/{warmup}: build {st['ms_build']:.2f} ms, "
            f"classify kernel {st['ms_classify_kernel']:.2f} ms")
    torch.cuda.synchronize()
    ctx.sync()
    lat, stats = [], []
    t_start = time.perf_counter()
    for k in range(steps):
        ts = time.perf_counter()
        st = step()
        torch.cuda.synchronize()
        lat.append((time.perf_counter() - ts) * 1e3)
        stats.append(st)
        log(f"[bench] {name} step {k + 1}/{steps}: {lat[-1]:.2f} ms")
    ctx.sync()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    del xi, yi, zi, cls
    return dict(cloud=cloud, n=n, elapsed=elapsed, lat=lat, stats=stats,
                value=n * steps / elapsed)


def serving_summary(nreq):
    """A short run of the serving loop on the serving stream's shape (48
    lognormal strata of tile size, median 2M points; Poisson arrivals at 0.7
    of the measured capacity; 4 workers): served points/s, p50/p99 latency
    and the comparison with a 20M-pulse tile of the same generator alone.
    The full run (128 strata, 2000 requests) is `python -m
    wolkenbase_amd.serve`.  A failure is reported, not raised."""
    try:
        from wolkenbase_amd import serve, synth

        t0 = time.perf_counter()
        pool = serve.pin_tiles(serve.stream_pool(48, 3000, 4e7))
        ref = serve.pin_tiles([synth.generate(**serve.REF_TILE)])[0]
        log(f"[bench] serving pool generated in {time.perf_counter() - t0:.1f} s; "
            f"{nreq} requests ...")
        r = serve.load_test(pool, slots=4, load=0.7, nreq=nreq, seed=3000, ref_tile=ref)
        keep = ("requests", "pool_tiles", "pool_points", "single_tile", "capacity_points_per_s",
                "served_points_per_s", "served_over_single_tile", "p50_ms", "p99_ms",
                "p99_over_single_tile_ms", "p50_over_isolated", "p99_over_isolated",
                "mismatched_results")
        return {k: r[k] for k in keep}
    except Exception as e:  # reported in the line, the headline stands
        return {"error": repr(e)}


def _points_in_flat(tree, kind, P):
    """wb_points_in through the C ABI: (offsets, flat input indices)."""
    import ctypes as C

    import numpy as np

    import wolkenbase_amd as W

    L = W.lib()
    P = np.ascontiguousarray(P, dtype=np.float64)
    pp = P.ctypes.data_as(C.POINTER(C.c_double))
    offs = np.zeros(len(P) + 1, dtype=np.uint64)
    op = offs.ctypes.data_as(C.POINTER(C.c_uint64))
    tree.ctx.check(L.wb_points_in(tree.h, kind, pp, len(P), op, None, 0), "wb_points_in")
    idx = np.empty(max(1, int(offs[-1])), dtype=np.uint32)
    tree.ctx.check(L.wb_points_in(tree.h, kind, pp, len(P), op,
                                  idx.ctypes.data_as(C.POINTER(C.c_uint32)), int(offs[-1])),
                   "wb_points_in")
    return offs, idx[:int(offs[-1])]


def query_kernels(ctx, cloud, nq, seed=7):
    """Batched OctStore::countPointsIn / pointsIn (octree.cpp:1214-1264) over
    the tile's octree: nq spheres of radius 1 m and nq cylinders of radius
    0.877 m (the tile cylinder of a 1 m tiling of a 900 m cube) centred on
    input points drawn at random (seeded); pointsIn for the spheres.  Kernel
    times from the library's HIP events: the counting pass includes ordering
    the shapes (k_query_keys and the radix sort of their Morton keys).
    Algorithmic bytes of a batch: the 40-B shapes in, the 8-B counts out and
    the 24-B FP64 coordinates of every point some shape of the batch holds,
    read once (the union, from the batch's pointsIn); pointsIn adds 4 B per
    index written."""
    import numpy as np

    import wolkenbase_amd as W

    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    tree = W.Octree.build(ctx, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                          cloud["offset"], center, side)
    rng = np.random.default_rng(seed)
    pick = rng.integers(0, len(cloud["xi"]), nq)
    c = [cloud["offset"][k] + cloud["scale"][k] * cloud[a][pick].astype(np.float64)
         for k, a in enumerate(("xi", "yi", "zi"))]
    zero = np.zeros(nq)
    batches = {
        "sphere_r1m": (W.WB_SHAPE_SPHERE, np.stack([c[0], c[1], c[2], zero + 1.0, zero], 1)),
        "cylinder_r0.877m": (W.WB_SHAPE_CYLINDER,
                             np.stack([c[0], c[1], zero + 0.877, zero, zero], 1)),
    }
    out = {}
    n = tree.n_input
    try:
        for name, (kind, prm) in batches.items():
            tree.count_points_in((kind, prm))  # warm-up
            ms, wall = [], []
            for _ in range(3):
                t0 = time.perf_counter()
                cnt = tree.count_points_in((kind, prm))
                wall.append(time.perf_counter() - t0)
                ms.append(tree.query_ms()[0])
            kms = sorted(ms)[1]
            offs, idx = _points_in_flat(tree, kind, prm)
            pms = tree.query_ms()[1]
            assert np.array_equal(np.diff(offs), cnt)
            seen = np.zeros(n, dtype=bool)
            seen[idx] = True
            union = int(seen.sum())
            del seen
            alg = 48.0 * nq + 24.0 * union
            gbs = alg / (kms * 1e-3) / 1e9
            out[name] = {"queries": nq, "mean_points_in": round(float(cnt.sum()) / nq, 2),
                         "distinct_points": union,
                         "kernel": "k_query_count", "kernel_ms": round(kms, 3),
                         "queries_per_s": round(nq / (kms * 1e-3), 1),
                         "call_ms_host_in_out": round(1e3 * sorted(wall)[1], 3),
                         "algorithmic_bytes": alg, "gbs": round(gbs, 1),
                         "frac_of_8TBs": round(gbs / HBM_PEAK_GBS, 5)}
            if kind == W.WB_SHAPE_SPHERE:
                palg = 48.0 * nq + 24.0 * union + 4.0 * len(idx)
                pgbs = palg / (pms * 1e-3) / 1e9
                out["points_in_" + name] = {
                    "queries": nq, "indices": int(len(idx)), "kernel": "k_query_points",
                    "kernel_ms": round(pms, 3), "queries_per_s": round(nq / (pms * 1e-3), 1),
                    "algorithmic_bytes": palg, "gbs": round(pgbs, 1),
                    "frac_of_8TBs": round(pgbs / HBM_PEAK_GBS, 5)}
            del idx, offs
    finally:
        tree.free()
    return out


def kernel_times(stats):
    def avg(f):
        return sum(s[f] for s in stats) / len(stats)

    return {"k_ground": avg("ms_classify_kernel"), "k_ground_heavy": avg("ms_heavy_kernel"),
            "radix_sort": avg("ms_sort")}


def breakdown(stats):
    return {k: round(sum(s[k] for s in stats) / len(stats), 3) for k in
            ("ms_build", "ms_sort", "ms_tiles", "ms_scan", "ms_classify",
             "ms_classify_kernel", "ms_heavy_kernel")}


def classes(st):
    n = int(st["n_class1"] + st["n_class2"] + st["n_unclassified"])
    return {"ground": int(st["n_class2"]), "not_ground": int(st["n_class1"]),
            "unclassified": int(st["n_unclassified"]),
            "ground_share": round(int(st["n_class2"]) / max(1, n), 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="tile20m", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mix", action="store_true",
                    help="tile20m: skip the second measurement on tile20m_high")
    ap.add_argument("--store", action="store_true",
                    help="las20m: also time the store's blocks into pinned host memory")
    ap.add_argument("--cpu-window", type=float, default=60.0)
    ap.add_argument("--no-serving", action="store_true",
                    help="tile20m: skip the short serving-stream measurement")
    ap.add_argument("--serving-requests", type=int, default=400)
    ap.add_argument("--no-queries", action="store_true",
                    help="tile20m: skip the batched shape-query measurement")
    ap.add_argument("--queries", type=int, default=1 << 20,
                    help="tile20m: shapes per batched query measurement")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="run one 256 MiB device copy first, so a --pmc profile of this "
                         "command can scale FETCH_SIZE/WRITE_SIZE against known bytes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        if int(os.environ.get("RANK", "0")) == 0:
            print("bench.py runs on one GPU only: the classification of one tile is not "
                  "sharded across GPUs.", flush=True)
        sys.exit(2)
    if args.workload == "las20m":
        return main_las(args)

    import torch

    import wolkenbase_amd as W

    # torch ships its own HIP runtime: it has to open the device before the
    # library's runtime does, or it finds no GPU
    torch.zeros(1, device="cuda:0")
    torch.cuda.synchronize()
    ctx = W.Context(0)
    r = run_tile(args.workload, args.steps, args.warmup, ctx, args.pmc_calibrate)
    cloud, n, stats, lat = r["cloud"], r["n"], r["stats"], r["lat"]
    value = r["value"]
    st = stats[-1]
    kern = kernel_times(stats)
    dominant = max(kern, key=kern.get)
    passes = int(st["sort_passes"])
    # algorithmic bytes per launch (compulsory traffic per point): the ground
    # test reads each point once (12 B as int32) and writes its class byte;
    # the sort moves its (key, u32 index) records, read and written once per
    # pass, plus the key read of each digit histogram: a build of at most 4
    # passes without the duplicate merge sorts u32 keys (16 + 4 B per point
    # and pass), a deeper one u64 keys (24 + 8 B)
    nh = int(st["heavy_points"])
    kb = 4 if passes <= 4 else 8
    sort_bpp = 2.0 * (kb + 4) + kb
    alg = {"k_ground": 13.0 * n, "k_ground_heavy": 13.0 * nh,
           "radix_sort": sort_bpp * passes * n}
    achieved = alg[dominant] / (kern[dominant] * 1e-3) / 1e9
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dominant}_{args.workload}.txt")
    if os.path.exists(tfile):
        for line in open(tfile):
            if line.startswith("hbm_bytes_per_launch"):
                traffic = float(line.split("=")[1])
    roof = dict(bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 6), traffic=traffic, kernel=dominant,
                kernel_ms=round(kern[dominant], 3),
                algorithmic_bytes_per_launch=alg[dominant])
    D = int(st["maxdepth"])
    P = math.ceil(3 * D / 8)
    B = 12 + 12 + 8 + 24 * P + 24 + 12 * 1.2097 + 12 + 1
    pipeline_frac = B * value / (HBM_PEAK_GBS * 1e9)
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["elapsed"] * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload], "n_points": n,
                   "tile_size_m": 1.0, "max_slope": 1.0, "thickness": 0.0, "min_hs_m": 0.1,
                   "input": "int32 x/y/z resident in HBM; class bytes left in HBM"},
        "p50_ms": round(percentile(lat, 50), 3),
        "p99_ms": round(percentile(lat, 99), 3),
        "roofline": roof,
        "pipeline_roofline": {"bytes_per_point": B, "maxdepth": D, "passes_formula": P,
                              "frac_of_8TBs": round(pipeline_frac, 6)},
        "breakdown_ms": breakdown(stats),
        "classes": classes(st),
        "exact_paths": {"gap_points": int(st["exact_gap_points"]),
                        "heavy_points": int(st["heavy_points"]),
                        "atan_x87_steps": int(st["atan_exact_calls"]),
                        "sector_borderline": int(st["sector_borderline"])},
    }
    # the radix sort (K2) against the same peak
    sort_alg = sort_bpp * passes * n
    sort_gbs = sort_alg / (kern["radix_sort"] * 1e-3) / 1e9
    out["kernels"] = {"radix_sort": {"passes": passes, "key_bytes": kb,
                                     "ms": round(kern["radix_sort"], 3),
                                     "algorithmic_bytes": sort_alg, "gbs": round(sort_gbs, 1),
                                     "frac_of_8TBs": round(sort_gbs / HBM_PEAK_GBS, 5)}}
    if args.workload == "tile20m" and not args.no_queries:
        log("[bench] batched shape queries ...")
        out["kernels"].update(query_kernels(ctx, cloud, args.queries))
    if args.workload == "tile20m" and not args.no_mix:
        # the same tile with every outlier above the surface: a realistic
        # ground share, where most points need the exhaustive search
        m = run_tile("tile20m_high", max(3, args.steps // 2), max(1, args.warmup // 2), ctx)
        mk = kernel_times(m["stats"])
        mst = m["stats"][-1]
        mtraffic = None
        tfile = os.path.join(ROOT, "profiles", "traffic_k_ground_tile20m_high.txt")
        if os.path.exists(tfile):
            for line in open(tfile):
                if line.startswith("hbm_bytes_per_launch"):
                    mtraffic = float(line.split("=")[1])
        out["ground_mix"] = {
            "workload": WORKLOADS["tile20m_high"], "n_points": m["n"],
            "value": round(m["value"], 1), "unit": "points/s",
            "ms_per_step": round(m["elapsed"] * 1e3 / len(m["stats"]), 3),
            "steps": len(m["stats"]),
            "p50_ms": round(percentile(m["lat"], 50), 3),
            "p99_ms": round(percentile(m["lat"], 99), 3),
            "k_ground_ms": round(mk["k_ground"], 3),
            "k_ground_gbs": round(13.0 * m["n"] / (mk["k_ground"] * 1e-3) / 1e9, 2),
            "k_ground_traffic": mtraffic,
            "k_ground_traffic_gbs": (round(mtraffic / (mk["k_ground"] * 1e-3) / 1e9, 1)
                                     if mtraffic else None),
            "breakdown_ms": breakdown(m["stats"]),
            "classes": classes(mst),
            "exact_paths": {"gap_points": int(mst["exact_gap_points"]),
                            "heavy_points": int(mst["heavy_points"])},
        }
    if args.workload == "tile20m" and not args.no_serving:
        out["serving"] = serving_summary(args.serving_requests)
    if not args.no_cpu_baseline:
        log(f"[bench] cpu baseline on a {args.cpu_window:g} m window ...")
        out["cpu_baseline"] = cpu_baseline(cloud, args.cpu_window)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
