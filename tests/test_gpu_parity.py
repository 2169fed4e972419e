"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  Octree contents, point order, neighbour counts and class
bytes must be bit-identical."""
import numpy as np
import pytest

import oracle_lib as O
import wolkenbase_amd as W
from wolkenbase_amd import synth

pytestmark = pytest.mark.gpu


def decode(c):
    x = c["offset"][0] + c["scale"][0] * c["xi"].astype(np.float64)
    y = c["offset"][1] + c["scale"][1] * c["yi"].astype(np.float64)
    z = c["offset"][2] + c["scale"][2] * c["zi"].astype(np.float64)
    return x, y, z


def build_both(ctx, cloud):
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    oc, os_ = O.sizefit(cloud["bbmin"], cloud["bbmax"])
    assert np.array_equal(center, oc) and side == os_
    tree = W.Octree.build(ctx, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                          cloud["offset"], center, side)
    x, y, z = decode(cloud)
    ref = O.Tree(x, y, z, center, side)
    return tree, ref, (x, y, z)


CLOUDS = [
    dict(seed=1, extent=30.0, density=25.0),
    dict(seed=2, extent=60.0, density=10.0, embankments=2, outliers=0.02),
    dict(seed=3, extent=20.0, density=60.0, trees=900),
]


@pytest.mark.parametrize("cfg", CLOUDS)
def test_tree_matches_oracle(ctx, cfg):
    cloud = synth.generate(**cfg)
    tree, ref, _ = build_both(ctx, cloud)
    g = tree.leaves()
    r = ref.leaves()
    print("n", tree.n, "leaves", len(g["side"]), "depth", tree.info())
    assert len(g["side"]) == ref.nleaves
    assert np.array_equal(g["center"], r["center"])
    assert np.array_equal(g["side"], r["side"])
    assert np.array_equal(g["count"], r["count"])
    assert np.array_equal(g["first"], r["first"])
    assert np.array_equal(g["depth"], r["depth"])
    assert np.array_equal(tree.order().astype(np.int64), ref.order())


def test_tree_small_sizes(ctx):
    for n in (1, 2, 536, 537, 538, 4296, 4297):
        cloud = synth.generate(seed=100 + n, extent=max(2.0, np.sqrt(n / 25.0)), density=25.0,
                               buildings=0, trees=0)
        k = min(n, len(cloud["xi"]))
        for key in ("xi", "yi", "zi"):
            cloud[key] = cloud[key][:k]
        x, y, z = decode(cloud)
        pad = 1.0 if k == 1 else 0.0   # a single point gives sizeFit side 0
        cloud["bbmin"] = np.array([x.min(), y.min(), z.min()]) - pad
        cloud["bbmax"] = np.array([x.max(), y.max(), z.max()]) + pad
        tree, ref, _ = build_both(ctx, cloud)
        g, r = tree.leaves(), ref.leaves()
        assert np.array_equal(g["count"], r["count"]), n
        assert np.array_equal(g["center"], r["center"]), n
        assert np.array_equal(tree.order().astype(np.int64), ref.order()), n


def test_tree_deep_cluster_more_sort_digits(ctx):
    # 1000 distinct points in a 1 cm cube inside a 30 m tile need leaves far
    # below the depth the default sort covers: the build sorts again with
    # more digits (and, with duplicates, merges again) until it reaches them
    cloud = synth.generate(seed=61, extent=30.0, density=20.0, buildings=0, trees=10)
    g = np.arange(10)
    gx, gy, gz = np.meshgrid(g, g, g, indexing="ij")
    c0 = (int(np.median(cloud["xi"])), int(np.median(cloud["yi"])), int(np.median(cloud["zi"])))
    for key, gk, c in (("xi", gx, c0[0]), ("yi", gy, c0[1]), ("zi", gz, c0[2])):
        cloud[key] = np.concatenate([cloud[key], (gk.ravel() + c).astype(np.int32)])
    tree, ref, (x, y, z) = build_both(ctx, cloud)
    assert tree.info()["maxdepth"] == ref.maxdepth == 12
    assert tree.info()["sort_passes"] == 5  # the default for a 32 m root is 3
    assert np.array_equal(tree.leaves()["count"], ref.leaves()["count"])
    assert np.array_equal(tree.order().astype(np.int64), ref.order())
    # the same with every cluster point given twice, merged
    xi, yi, zi = (np.concatenate([cloud[k], cloud[k][-1000:]]) for k in ("xi", "yi", "zi"))
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    t2 = W.Octree.build(ctx, xi, yi, zi, cloud["scale"], cloud["offset"], center, side,
                        dedupe=True)
    assert t2.ndup == 1000
    assert np.array_equal(t2.order().astype(np.int64), ref.order())
    pairs = t2.duplicates()
    n0 = len(cloud["xi"])
    assert sorted(pairs[:, 0].tolist()) == list(range(n0, n0 + 1000))
    assert np.array_equal(pairs[np.argsort(pairs[:, 0]), 1], np.arange(n0 - 1000, n0))


def test_degenerate_root_rejected(ctx):
    # Octree::sizeFit gives side 0 for a single location; the reference then
    # asserts side>0 on the first findBlock (octree.cpp:208).  The ABI
    # returns WB_EINVAL instead.
    c, s = W.Octree.size_fit([1.0, 2.0, 3.0], [1.0, 2.0, 3.0])
    assert s == 0.0
    with pytest.raises(W.WbError, match="root side"):
        W.Octree.build(ctx, [1000], [2000], [3000], [0.001] * 3, [0.0] * 3, c, s)


def random_shapes(rng, x, y, z, kind, m):
    i = rng.integers(0, len(x), m)
    if kind == W.WB_SHAPE_CYLINDER:
        return np.stack([x[i] + rng.normal(0, 0.3, m), y[i] + rng.normal(0, 0.3, m),
                         rng.uniform(0.05, 3.0, m), np.zeros(m), np.zeros(m)], 1)
    if kind == W.WB_SHAPE_SPHERE:
        return np.stack([x[i], y[i], z[i] + rng.normal(0, 0.5, m), rng.uniform(0.05, 4.0, m),
                         np.zeros(m)], 1)
    return np.stack([x[i], y[i], z[i], rng.uniform(0.05, 1.0, m), rng.choice([1.0, 0.7, -1.0], m)], 1)


@pytest.mark.parametrize("kind", [W.WB_SHAPE_CYLINDER, W.WB_SHAPE_SPHERE, W.WB_SHAPE_HYPERBOLOID])
def test_count_and_points_in(ctx, kind):
    cloud = synth.generate(seed=5, extent=30.0, density=25.0)
    tree, ref, (x, y, z) = build_both(ctx, cloud)
    rng = np.random.default_rng(kind)
    P = random_shapes(rng, x, y, z, kind, 300)
    counts = tree.count_points_in((kind, P))
    lists = tree.points_in((kind, P[:60]))
    for s in range(len(P)):
        want = ref.count_points_in(kind, P[s])
        assert counts[s] == want, (s, P[s])
    for s in range(60):
        assert np.array_equal(lists[s].astype(np.int64), ref.points_in(kind, P[s])), s


def _noalloc_counts(tree, kind, P):
    """The wave-per-shape kernel (the no-allocation entry point, which takes
    the shapes in the caller's order), in batches of its maximum."""
    import ctypes as C

    out = np.zeros(len(P), dtype=np.uint64)
    step = 65536
    for a in range(0, len(P), step):
        blk = np.ascontiguousarray(P[a:a + step])
        o = np.zeros(len(blk), dtype=np.uint64)
        rc = W.lib().wb_count_points_in_noalloc(
            tree.h, kind, blk.ctypes.data_as(C.POINTER(C.c_double)), len(blk),
            o.ctypes.data_as(C.POINTER(C.c_uint64)))
        assert rc == 0
        out[a:a + step] = o
    return out


@pytest.mark.parametrize("kind", [W.WB_SHAPE_CYLINDER, W.WB_SHAPE_SPHERE, W.WB_SHAPE_HYPERBOLOID])
def test_batched_queries_large_batch(ctx, kind):
    """Many shapes in random order (the lane-per-shape kernels visit them in
    Morton order of their centres): counts equal the wave-per-shape kernel's
    for all of them and the oracle's for a sample; index lists equal the
    oracle's, and their lengths the counts.  Edge shapes: far outside the
    cloud, radius zero, and one that holds every leaf whole."""
    cloud = synth.generate(seed=31, extent=80.0, density=25.0, outliers=0.01)
    tree, ref, (x, y, z) = build_both(ctx, cloud)
    rng = np.random.default_rng(100 + kind)
    P = random_shapes(rng, x, y, z, kind, 20000)
    far = P[:3].copy()
    far[:, 0] += 1e4
    zero = P[3:6].copy()
    if kind == W.WB_SHAPE_HYPERBOLOID:
        zero[:, 3] = 0.0
        big = np.array([[x.mean(), y.mean(), z.max() + 1e4, 1.0, 1.0]])
    elif kind == W.WB_SHAPE_SPHERE:
        zero[:, 3] = 0.0
        big = np.array([[x.mean(), y.mean(), z.mean(), 1e5, 0.0]])
    else:
        zero[:, 2] = 0.0
        big = np.array([[x.mean(), y.mean(), 1e5, 0.0, 0.0]])
    P = np.concatenate([P, far, zero, big])
    counts = tree.count_points_in((kind, P))
    assert np.array_equal(counts, _noalloc_counts(tree, kind, P))
    # batches of up to 64 shapes take a wave per shape, larger ones a lane
    for k in (1, 64, 65):
        assert np.array_equal(tree.count_points_in((kind, P[:k])), counts[:k])
    assert counts[-1] == len(x)
    assert (counts[-7:-4] == 0).all()
    for s in list(rng.integers(0, len(P), 150)) + list(range(len(P) - 7, len(P))):
        assert counts[s] == ref.count_points_in(kind, P[s]), (s, P[s])
    sub = np.concatenate([P[:3000], P[-7:]])
    lists = tree.points_in((kind, sub))
    assert [len(a) for a in lists] == [int(c) for c in np.concatenate([counts[:3000],
                                                                       counts[-7:]])]
    for s in list(rng.integers(0, 3000, 80)) + list(range(3000, 3007)):
        assert np.array_equal(lists[s].astype(np.int64), ref.points_in(kind, sub[s])), s


# the third cloud is dense enough for tiles of 65-128, 129-256 and more than
# 256 points (each size class of the scan has its own kernel)
@pytest.mark.parametrize("cfg", [dict(seed=21, extent=24.0, density=25.0),
                                 dict(seed=22, extent=30.0, density=20.0, embankments=2,
                                      outliers=0.02),
                                 dict(seed=26, extent=12.0, density=500.0, buildings=1, trees=20)])
def test_classify_matches_oracle(ctx, cfg):
    cloud = synth.generate(**cfg)
    tree, _, _ = build_both(ctx, cloud)
    cls, tiles, st = W.classify(tree, cloud["bbmin"], cloud["bbmax"])
    rcls, rtiles, rst = O.run(cloud)
    print("stats", st, "oracle", rst.n_class1, rst.n_class2, rst.n_unclassified)
    for f in ("npoints", "nground", "tree_flags"):
        assert np.array_equal(tiles[f], rtiles[f]), f
    for f in ("density", "hs_scan", "hs", "height", "tilt_x", "tilt_y"):
        assert np.array_equal(tiles[f], rtiles[f]), f
    assert not tiles["fit_fp64"].any() and st["fit_fallbacks"] == 0
    assert int((cls != rcls).sum()) == 0
    if cfg["seed"] == 26:
        m = tiles["npoints"]
        assert ((m > 64) & (m <= 128)).any() and ((m > 128) & (m <= 256)).any() and (m > 256).any()


def test_classify_sharing_context_matches_oracle():
    """A context in sharing mode (the resident ground kernel, capped grids)
    gives the oracle's class bytes and tile records."""
    cloud = synth.generate(seed=22, extent=30.0, density=20.0, embankments=2, outliers=0.02)
    sctx = W.Context(0, sharing=True)
    try:
        tree, _, _ = build_both(sctx, cloud)
        cls, tiles, _ = W.classify(tree, cloud["bbmin"], cloud["bbmax"])
        rcls, rtiles, _ = O.run(cloud)
        for f in ("npoints", "nground", "tree_flags", "hs"):
            assert np.array_equal(tiles[f], rtiles[f]), f
        assert int((cls != rcls).sum()) == 0
    finally:
        sctx.close()


def test_classify_sharing_context_equals_default(ctx):
    """On a cloud with more 64-point units (6.9k) than the capped resident
    grid has waves (4 blocks per CU), so waves go on to take units from the
    counter, sharing mode gives the default mode's class bytes, tile records
    and handed-over point counts."""
    cloud = synth.generate(seed=28, extent=100.0, density=40.0, buildings=60, trees=500,
                           outliers=0.01)
    assert len(cloud["xi"]) > 4 * 256 * 4 * 64
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    res = []
    for c in (ctx, W.Context(0, sharing=True)):
        tree = W.Octree.build(c, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                              cloud["offset"], center, side)
        res.append(W.classify(tree, cloud["bbmin"], cloud["bbmax"]))
        tree.free()
        if c is not ctx:
            c.close()
    (c0, t0, s0), (c1, t1, s1) = res
    assert np.array_equal(c0, c1)
    for f in t0.dtype.names:
        assert np.array_equal(t0[f], t1[f]), f
    assert s0["heavy_points"] == s1["heavy_points"] > 0
    assert s0["exact_gap_points"] == s1["exact_gap_points"]


# classification parameters other than the defaults (the hyperboloid's
# vertex drop and slope, the tile spacing, the floor of the hyperboloid size)
PARAMS = [dict(thickness=0.05), dict(max_slope=0.7), dict(max_slope=1.5), dict(tile_size=2.0),
          dict(min_hs=0.3),
          dict(tile_size=0.7, max_slope=1.2, thickness=0.02, min_hs=0.2)]


@pytest.mark.parametrize("prm", PARAMS, ids=lambda p: ",".join(f"{k}={v}" for k, v in p.items()))
def test_classify_params_match_oracle(ctx, prm):
    cloud = synth.generate(seed=23, extent=30.0, density=20.0, embankments=2, outliers=0.02,
                           outlier_low=0.2)
    tree, _, _ = build_both(ctx, cloud)
    gp = W.default_params(**prm)
    op = O.default_params()
    for k, v in prm.items():
        setattr(op, k, v)
    cls, tiles, st = W.classify(tree, cloud["bbmin"], cloud["bbmax"], gp)
    rcls, rtiles, rst = O.run(cloud, op)
    print(prm, "stats", st["n_class1"], st["n_class2"], "oracle", rst.n_class1, rst.n_class2)
    assert len(tiles) == len(rtiles)
    for f in ("npoints", "tree_flags"):
        assert np.array_equal(tiles[f], rtiles[f]), f
    for f in ("density", "hs_scan", "hs", "height", "tilt_x", "tilt_y"):
        assert np.array_equal(tiles[f], rtiles[f]), f
    assert int((cls != rcls).sum()) == 0
    assert 0 < st["n_class2"] < len(cls)
    assert np.array_equal(tiles["nground"], rtiles["nground"])


def header_box(cloud, lo_frac, hi_frac):
    """A header box over the middle of the cloud in x and y and the 20th to
    90th z percentiles: a LAS header whose bounds do not cover its records."""
    lo, hi = np.array(cloud["bbmin"], float), np.array(cloud["bbmax"], float)
    ext = hi - lo
    _, _, z = decode(cloud)
    c = dict(cloud)
    c["bbmin"] = [lo[0] + lo_frac * ext[0], lo[1] + lo_frac * ext[1], float(np.percentile(z, 20))]
    c["bbmax"] = [lo[0] + hi_frac * ext[0], lo[1] + hi_frac * ext[1], float(np.percentile(z, 90))]
    return c


@pytest.mark.parametrize("seed", [31, 32])
def test_classify_points_outside_header_box(ctx, seed):
    # the root cube and the tiling come from the box; the points outside the
    # root cube sit in its extreme cells, far outside their leaves' cubes,
    # and still count as neighbours of the points inside
    cloud = header_box(synth.generate(seed=seed, extent=40.0, density=20.0, embankments=1,
                                      outliers=0.01), 0.2, 0.75)
    tree, _, (x, y, z) = build_both(ctx, cloud)
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    out = ((np.abs(x - center[0]) > side / 2) | (np.abs(y - center[1]) > side / 2) |
           (np.abs(z - center[2]) > side / 2))
    assert out.sum() > 10000
    cls, tiles, st = W.classify(tree, cloud["bbmin"], cloud["bbmax"])
    rcls, rtiles, rst = O.run(cloud)
    print("outside", int(out.sum()), "stats", st, "oracle", rst.n_class1, rst.n_class2,
          rst.n_unclassified)
    assert np.array_equal(tiles["npoints"], rtiles["npoints"])
    assert np.array_equal(tiles["hs"], rtiles["hs"])
    assert int((cls != rcls).sum()) == 0
    assert rst.n_class2 > 0 and rst.n_class1 > 0


def test_points_outside_root_cube_pile_up(ctx):
    # a smaller box: more than 537 points outside one corner of the root
    # cube follow the same extreme path to level 21; neither tree exists
    cloud = header_box(synth.generate(seed=32, extent=40.0, density=20.0, embankments=1,
                                      outliers=0.01), 0.3, 0.65)
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    with pytest.raises(W.WbError, match="21 levels"):
        W.Octree.build(ctx, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                       cloud["offset"], center, side)
    x, y, z = decode(cloud)
    with pytest.raises(ValueError):
        O.Tree(x, y, z, center, side)


def test_context_close_frees_live_trees():
    # a tree still alive when its context closes is freed with it; freeing
    # it afterwards is a no-op, not a use of the closed context
    c = W.Context(0)
    cloud = synth.generate(seed=31, extent=10.0, density=25.0)
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    tree = W.Octree.build(c, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                          cloud["offset"], center, side)
    c.close()
    assert tree.h is None
    tree.free()


def normal_angle(ax, ay, bx, by):
    """Angle between the plane normals (-ax, -ay, 1) and (-bx, -by, 1)."""
    n1 = np.stack([-ax, -ay, np.ones_like(ax)], 1)
    n2 = np.stack([-bx, -by, np.ones_like(bx)], 1)
    cr = np.linalg.norm(np.cross(n1, n2), axis=1)
    return np.arctan2(cr, (n1 * n2).sum(1))


def test_scan_tiles_matches_classify_and_oracle(ctx):
    cloud = synth.generate(seed=41, extent=24.0, density=25.0, embankments=1)
    tree, _, _ = build_both(ctx, cloud)
    tiles, st = W.scan_tiles(tree, cloud["bbmin"], cloud["bbmax"])
    _, ctiles, _ = W.classify(tree, cloud["bbmin"], cloud["bbmax"])
    _, rtiles, _ = O.run(cloud)
    for f in ("npoints", "tree_flags", "density", "hs_scan", "hs", "height", "tilt_x", "tilt_y"):
        assert np.array_equal(tiles[f], ctiles[f]), f
        assert np.array_equal(tiles[f], rtiles[f]), f
    assert st["nonempty_tiles"] == int((tiles["npoints"] > 0).sum())


def line_cloud(seed):
    """Points only along parallel lines 3 m apart (x every 2 cm, y within
    2 mm of the line, z a gentle slope): every tile cylinder is near-collinear
    in xy, the case the FP32 fit hands to FP64."""
    rng = np.random.default_rng(seed)
    xs = np.arange(0, 30000, 20)
    ys = np.arange(0, 30000, 3000)
    x = np.repeat(xs[None, :], len(ys), 0).ravel()
    y = (ys[:, None] + rng.integers(0, 3, (len(ys), len(xs)))).ravel()
    z = (x // 10 + y // 20 + rng.integers(0, 5, x.size)).astype(np.int64)
    xi, yi, zi = x.astype(np.int32), y.astype(np.int32), z.astype(np.int32)
    scale = np.array([0.001, 0.001, 0.001])
    offset = np.array([500000.0, 4400000.0, 0.0])
    lo = offset + scale * np.array([xi.min(), yi.min(), zi.min()])
    hi = offset + scale * np.array([xi.max(), yi.max(), zi.max()])
    return dict(xi=xi, yi=yi, zi=zi, scale=scale, offset=offset, bbmin=lo, bbmax=hi)


# FP32 plane fits (an option for the precision study; FP64 is the reference).
# Tolerance: on tiles fitted in FP32 the plane normal is within 1e-3 rad of
# the FP64 normal with compensated sums; tiles that fall back are bit-equal.
@pytest.mark.parametrize("kind", ["terrain", "lines"])
def test_fp32_fits_against_fp64(ctx, kind):
    if kind == "lines":
        cloud = line_cloud(43)
    else:
        cloud = synth.generate(seed=42, extent=40.0, density=25.0, linear=0.05)
    tree, _, _ = build_both(ctx, cloud)
    ref, _ = W.scan_tiles(tree, cloud["bbmin"], cloud["bbmax"])
    comp, st = W.scan_tiles(tree, cloud["bbmin"], cloud["bbmax"],
                            W.default_params(fit_precision=W.WB_FIT_FP32_COMP))
    naive, _ = W.scan_tiles(tree, cloud["bbmin"], cloud["bbmax"],
                            W.default_params(fit_precision=W.WB_FIT_FP32))
    live = ref["npoints"] > 0
    assert np.array_equal(comp["npoints"], ref["npoints"])
    fb = comp["fit_fp64"] == 1
    assert st["fit_fallbacks"] == int(fb.sum())
    for f in ("tilt_x", "tilt_y", "density", "height"):
        assert np.array_equal(comp[f][fb], ref[f][fb]), f
    ok = live & ~fb
    ang = normal_angle(comp["tilt_x"][ok], comp["tilt_y"][ok], ref["tilt_x"][ok], ref["tilt_y"][ok])
    print("fallbacks", int(fb.sum()), "of", int(live.sum()), "max angle comp",
          ang.max() if ok.any() else None,
          "naive", normal_angle(naive["tilt_x"][live], naive["tilt_y"][live],
                                ref["tilt_x"][live], ref["tilt_y"][live]).max())
    if ok.any():
        assert ang.max() < 1e-3
    if kind == "lines":
        assert fb.sum() > 0.9 * live.sum()
