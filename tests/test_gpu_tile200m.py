"""The 200M-point tile (3 km x 3 km, an urban and a forested half) at full
size on one GPU.  The oracle cannot run 200M points, so these check
properties that do not depend on size: the octree is the canonical one
(every leaf holds 1..537 points, the cube above every leaf more than 537,
every point lies in its leaf's cube, leaves in DFS order with points in
input order), batched queries fall inside brute-force bounds, SCAN's tile
membership equals the count query on the same cylinders, and the
classification covers every point and repeats bit for bit."""
import math
import sys
import time

import numpy as np
import pytest

import oracle_lib as O
import wolkenbase_amd as W
from wolkenbase_amd import synth

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(1200)]


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def big():
    t0 = time.time()
    cloud = synth.preset("tile200m")
    n = len(cloud["xi"])
    log(f"[tile200m] {n} points generated in {time.time() - t0:.0f} s")
    ctx = W.Context(0)
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    t0 = time.time()
    tree = W.Octree.build(ctx, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                          cloud["offset"], center, side)
    log(f"[tile200m] built in {time.time() - t0:.1f} s (host copies included): {tree.info()}")
    xyz = [cloud["offset"][a] + cloud["scale"][a] * cloud[k].astype(np.float64)
           for a, k in enumerate(("xi", "yi", "zi"))]
    yield dict(cloud=cloud, ctx=ctx, tree=tree, center=np.asarray(center), side=side, xyz=xyz)
    tree.free()
    ctx.close()


def test_octree_is_canonical(big):
    tree, n = big["tree"], len(big["cloud"]["xi"])
    assert n > 190_000_000
    L = tree.leaves()
    cnt = L["count"].astype(np.int64)
    depth = L["depth"].astype(np.int64)
    log(f"[tile200m] {len(cnt)} leaves, depth {depth.min()}..{depth.max()}")
    assert cnt.min() >= 1 and cnt.max() <= 537 and cnt.sum() == n
    assert np.array_equal(L["first"].astype(np.int64), np.concatenate([[0], np.cumsum(cnt)[:-1]]))
    # each leaf cube is the cell of its depth under the root cube (exact dyadics)
    lo = big["center"] - big["side"] / 2
    for d in np.unique(depth):
        s = big["side"] / 2.0 ** d
        m = depth == d
        assert (L["side"][m] == s).all()
        k = np.floor((L["center"][m] - lo) / s)
        assert np.array_equal(L["center"][m], lo + (k + 0.5) * s)
    # the cube above every leaf holds more than 537 points: the leaves under
    # the parent cell at depth d-1, summed
    for d in np.unique(depth):
        if d <= 1:
            continue
        s = big["side"] / 2.0 ** (d - 1)
        under = depth >= d
        keys = np.floor((L["center"][under] - lo) / s).astype(np.int64)
        flat = (keys[:, 0] * 4096 + keys[:, 1]) * 4096 + keys[:, 2]
        u, inv = np.unique(flat, return_inverse=True)
        tot = np.bincount(inv, weights=cnt[under])
        child = depth[under] == d
        assert tot[inv[child]].min() > 537, d


def test_points_in_their_leaves_in_input_order(big):
    tree = big["tree"]
    L = tree.leaves()
    order = tree.order().astype(np.int64)  # canonical position -> input index
    n = len(order)
    assert np.bincount(order, minlength=n).max() == 1  # a permutation
    leaf = np.repeat(np.arange(len(L["count"])), L["count"].astype(np.int64))
    half = L["side"][leaf] / 2
    for a in range(3):
        c = L["center"][leaf, a]
        v = big["xyz"][a][order]
        assert ((v >= c - half) & (v <= c + half)).all(), "axis %d" % a
    same = leaf[1:] == leaf[:-1]
    assert (np.diff(order)[same] > 0).all()


def test_queries_within_brute_force_bounds(big):
    tree = big["tree"]
    x, y, z = big["xyz"]
    rng = np.random.default_rng(200)
    eps = 1e-9
    for kind in (W.WB_SHAPE_CYLINDER, W.WB_SHAPE_SPHERE, W.WB_SHAPE_HYPERBOLOID):
        P = []
        for i in rng.integers(0, len(x), 3):
            if kind == W.WB_SHAPE_CYLINDER:
                P.append([x[i] + 0.1, y[i], 2.5, 0, 0])
            elif kind == W.WB_SHAPE_SPHERE:
                P.append([x[i], y[i], z[i] + 0.5, 4.0, 0])
            else:
                P.append([x[i], y[i], z[i], 0.4, 1.0])
        P = np.array(P)
        got = tree.count_points_in((kind, P))
        for s, p in enumerate(P):
            box = 60.0
            m = np.flatnonzero((np.abs(x - p[0]) < box) & (np.abs(y - p[1]) < box))
            dx, dy, dz = x[m] - p[0], y[m] - p[1], z[m] - p[2]
            if kind == W.WB_SHAPE_CYLINDER:
                q = np.hypot(dx, dy) - p[2]
                lo, hi = (q < -eps).sum(), (q <= eps).sum()
            elif kind == W.WB_SHAPE_SPHERE:
                q = np.sqrt(dx * dx + dy * dy + dz * dz) - p[3]
                lo, hi = (q < -eps).sum(), (q <= eps).sum()
            else:
                por = p[3] * p[4] ** 2
                zd = (p[2] + por) - z[m]
                f = zd * zd - (p[4] * np.hypot(dx, dy)) ** 2 - por * por
                sc = eps * (zd * zd + por * por + 1)
                lo = ((zd > 0) & (f > sc)).sum()
                hi = ((zd > 0) & (f >= -sc)).sum()
            log(f"[tile200m] kind {kind} shape {s}: {got[s]} in [{lo}, {hi}]")
            assert lo <= got[s] <= hi


def test_scan_membership_equals_count_query(big):
    cloud, tree = big["cloud"], big["tree"]
    t0 = time.time()
    tiles, _ = W.scan_tiles(tree, cloud["bbmin"], cloud["bbmax"])
    log(f"[tile200m] {len(tiles)} tiles scanned in {time.time() - t0:.1f} s")
    sn = O.snake_setup(cloud["bbmin"], cloud["bbmax"], 1.0)
    assert len(tiles) == sn.hi - sn.lo + 1
    rng = np.random.default_rng(7)
    pick = rng.choice(np.flatnonzero(tiles["npoints"] > 0), 4000, replace=False)
    P = np.zeros((len(pick), 5))
    for j, k in enumerate(pick):
        ex, ey = O.to_flowsnake(sn.lo + int(k))
        assert (ex, ey) == (tiles["flowsnake_x"][k], tiles["flowsnake_y"][k])
        P[j, :3] = O.snake_cyl(sn, ex, ey)
    got = tree.count_points_in((W.WB_SHAPE_CYLINDER, P))
    assert np.array_equal(got.astype(np.int64), tiles["npoints"][pick].astype(np.int64))


def test_classification_covers_every_point_and_repeats(big):
    cloud, tree = big["cloud"], big["tree"]
    times = []
    out = []
    for _ in range(2):
        t0 = time.time()
        cls, _, st = W.classify(tree, cloud["bbmin"], cloud["bbmax"], want_tiles=False)
        times.append(time.time() - t0)
        out.append(cls)
    cls = out[0]
    log(f"[tile200m] classified in {times[1]:.2f} s (host copy included): ground "
        f"{st['n_class2']}, not ground {st['n_class1']}, none {st['n_unclassified']}, "
        f"heavy {st['heavy_points']}, classify kernel {st['ms_classify_kernel']:.1f} ms")
    assert np.array_equal(out[0], out[1])
    assert st["n_class1"] + st["n_class2"] + st["n_unclassified"] == len(cls)
    assert set(np.unique(cls)) <= {0, 1, 2}
    # a point outside every tile cylinder can only sit at the tile's edge
    x = big["xyz"][0]
    none = cls == 0
    assert none.sum() < 1e-3 * len(cls)
    # the generator's ground points: mostly ground in both halves
    truth = cloud["truth"]
    split = cloud["offset"][0] + 1500.0
    for half, m in (("urban", x < split), ("forest", x >= split)):
        g = (truth == 2) & m
        share = (cls[g] == 2).mean()
        log(f"[tile200m] {half} half: {share:.3f} of ground points classified ground")
        assert share > 0.8
