"""Pins the CPU oracle: against the reference's own sources built from
source (oracle/_ref, skipped where the reference tree is absent), against
the known answers of the reference's unit tests (wolkentest.cpp), and
against invariants and brute force."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from wolkenbase_amd import synth

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built here")
D = O._dp


# ---------------------------------------------------------------- atan2i --
@needs_ref
def test_tables_match_reference():
    R = O.ref()
    t, c, s = O.tables()
    for i in range(511):
        assert R.ref_tan_bin(i * 0x100000 - 0xff00000) == t[i]
    for i in range(512):
        assert R.ref_sin_bin(i * 0x100000 - 0xff80000) == s[i]
        assert R.ref_cos_bin(i * 0x100000 - 0xff80000) == c[i]


def atan_samples(rng, n):
    ys = np.concatenate([rng.normal(size=n), rng.integers(-3000, 3000, n) * 0.001,
                         np.zeros(50), rng.normal(size=50) * 1e-12, [0.0, -0.0, 1.0, -1.0]])
    xs = np.concatenate([rng.normal(size=n), rng.integers(-3000, 3000, n) * 0.001,
                         rng.normal(size=50), np.zeros(50), [-1.0, -1.0, 0.0, 0.0]])
    return ys, xs


@needs_ref
def test_atan2i_matches_reference():
    R, L = O.ref(), O.lib()
    ys, xs = atan_samples(np.random.default_rng(1), 100000)
    bad = [(y, x) for y, x in zip(ys, xs) if L.wbo_atan2i(y, x) != R.ref_atan2i(y, x)]
    assert not bad, bad[:5]


def foldangle(a):
    # angle.cpp:183-188
    a &= 0xffffffff
    if ((a >> 30) % 3) != 0:
        a ^= 0x80000000
    return a


def test_atan2i_round_trip():
    # wolkentest.cpp:122-132: foldangle(atan2i(cossin(i)))==foldangle(i)
    L = O.lib()
    rng = np.random.default_rng(2)
    for i in list(range(-2147400000, 2147400000, 9000017)) + list(rng.integers(-2**31, 2**31 - 1, 3000)):
        i = int(i)
        a = L.wbo_atan2i(L.wbo_sin_bin(i), L.wbo_cos_bin(i))
        assert foldangle(a) == foldangle(i), i


def test_atan2i_special_values():
    assert O.atan2i(0.0, 0.0) == 0
    assert O.atan2i(0.0, 1.0) == 0
    assert O.atan2i(1.0, 0.0) == 0x20000000
    assert O.atan2i(0.0, -1.0) == -0x40000000   # due west is exactly -180 deg
    assert O.atan2i(-1.0, 0.0) == -0x20000000


# ---------------------------------------------------------- pairwise sum --
def test_pairwisesum_odd_squares():
    # wolkentest.cpp:611-617
    odd = np.arange(32, dtype=np.float64) * 2 + 1
    for i in range(32):
        assert O.pairwisesum(odd[:i]) == i * i


def test_pairwisesum_tree_shape():
    # the reference tree: one balanced block per set bit of n, largest first
    a = np.array([1e16, 1.0, -1e16, 1.0, 3.0, 5.0, 7.0])
    blk4 = (a[0] + a[1]) + (a[2] + a[3])
    blk2 = a[4] + a[5]
    assert O.pairwisesum(a) == ((0 + a[6]) + blk2) + blk4


@needs_ref
def test_pairwisesum_matches_reference():
    R = O.ref()
    rng = np.random.default_rng(3)
    for n in list(range(0, 300)) + [1000, 1023, 1024, 1025, 4097, 12345]:
        a = np.ascontiguousarray(rng.normal(size=n) * 10.0 ** rng.integers(-5, 5, size=n))
        assert O.pairwisesum(a) == R.ref_pairwisesum(a.ctypes.data_as(D), n)


# ---------------------------------------------------------------- shapes --
def test_sphere_known_answers():
    # wolkentest.cpp:203-221: 89^2 = 15^2+36^2+80^2 = 39^2+48^2+64^2
    s = [100, 200, 300, 89]
    inside = [(100, 200, 300), (115, 164, 220), (36, 239, 252), (64, 120, 315), (164, 248, 339)]
    outside = [(115, 163, 220), (36, 240, 252), (63, 120, 315), (164, 248, 340)]
    assert all(O.shape_in(O.SH_SPH, s, q) for q in inside)
    assert not any(O.shape_in(O.SH_SPH, s, q) for q in outside)


def test_hyperboloid_known_answers():
    # wolkentest.cpp:223-263: 72^2 = 78^2-18^2-24^2 = 97^2-25^2-60^2
    h1 = [100, 200, 300, 72, 1]
    assert O.shape_in(O.SH_HYP, h1, (118, 176, 294))
    assert O.shape_in(O.SH_HYP, h1, (75, 260, 275))
    assert not O.shape_in(O.SH_HYP, h1, (118, 176, 295))
    assert not O.shape_in(O.SH_HYP, h1, (75, 260, 276))
    assert O.shape_in(O.SH_HYP, h1, (118, 176, 293))
    assert O.shape_in(O.SH_HYP, h1, (75, 260, 274))


def find_intersection(kind, p, a, b):
    # wolkentest.cpp:169-186: bisection on the segment a-b
    a, b = np.array(a, float), np.array(b, float)
    ain = O.shape_in(kind, p, a)
    while True:
        m = (a + b) / 2
        if np.array_equal(a, m) or np.array_equal(m, b):
            return m
        if O.shape_in(kind, p, m) == ain:
            a = m
        else:
            b = m


def test_hyperboloid_curvature_and_asymptote():
    # wolkentest.cpp:236-262
    topc, botc = (100.6, 200.8, 300), (100.6, 200.8, 228)
    for r, slope, ocz in ((72, 1, 228), (97, 2, 203)):
        h = [100, 200, 300, r, slope]
        s = [100, 200, ocz, r]
        sint = find_intersection(O.SH_SPH, s, topc, botc)[2]
        hint = find_intersection(O.SH_HYP, h, topc, botc)[2]
        assert abs(sint - hint) < 1 / r ** 2
        a1 = find_intersection(O.SH_HYP, h, (3000100, 4000200, 300), (3000100, 4000200, -4e9))[2]
        a2 = find_intersection(O.SH_HYP, h, (6000100, 8000200, 300), (6000100, 8000200, -4e9))[2]
        assert abs(slope - (a1 - a2) / 5e6) < 1e-6


def test_cylinder_known_answers():
    # wolkentest.cpp:265-279 (5-12-13)
    c = [0, 0, 13]
    assert O.shape_in(O.SH_CYL, c, (0, 0, 13))
    assert O.shape_in(O.SH_CYL, c, (5, 12, 0))
    assert not O.shape_in(O.SH_CYL, c, (5, 13, 4))
    assert O.shape_in(O.SH_CYL, c, (-9, -9, -3))
    assert not O.shape_in(O.SH_CYL, c, (-11, -7, 2))
    assert O.shape_intersect(O.SH_CYL, c, (0, 0, 13), 1)
    assert O.shape_intersect(O.SH_CYL, c, (5, 12, 0), 0.1)
    assert O.shape_intersect(O.SH_CYL, c, (-11, -7, 2), 1 / 16)
    assert not O.shape_intersect(O.SH_CYL, c, (-11, -7, 2), 1 / 32)


def test_paraboloid_known_answers():
    # wolkentest.cpp:188-201
    p = [0, 0, 13, 13]
    assert O.shape_in(O.SH_PAR, p, (0, 0, 13))
    assert O.shape_in(O.SH_PAR, p, (5, 0, 12))
    assert not O.shape_in(O.SH_PAR, p, (5, 2, 12))
    assert O.shape_in(O.SH_PAR, p, (13, 13, 0))
    assert not O.shape_in(O.SH_PAR, p, (-14, -12, 0))
    assert O.shape_in(O.SH_PAR, p, (9, 16, 0))
    assert O.shape_intersect(O.SH_PAR, p, (0, 0, 13), 1)
    assert not O.shape_intersect(O.SH_PAR, p, (5, 2, 12), 0.1)


def random_shape(rng, kind):
    c = rng.normal(size=3) * 3
    if kind == O.SH_CYL:
        return np.array([c[0], c[1], abs(rng.normal()) * 2])
    if kind == O.SH_SPH:
        return np.array([c[0], c[1], c[2], abs(rng.normal()) * 2])
    if kind == O.SH_HYP:
        return np.array([c[0], c[1], c[2], abs(rng.normal()), rng.choice([1.0, -1.0, 0.5])])
    return np.array([c[0], c[1], c[2], rng.normal() * 2])


@needs_ref
def test_shapes_match_reference():
    R = O.ref()
    rng = np.random.default_rng(4)
    for it in range(40000):
        kind = it % 4
        p = np.ascontiguousarray(random_shape(rng, kind))
        q = np.ascontiguousarray(rng.normal(size=3) * 3)
        cc = np.ascontiguousarray(np.round(rng.normal(size=3) * 4) / 4)
        side = 2.0 ** rng.integers(-3, 3)
        got = (O.shape_in(kind, p, q), O.shape_in_cube(kind, p, cc, side),
               O.shape_intersect(kind, p, cc, side))
        want = (R.ref_shape_in(kind, p.ctypes.data_as(D), q.ctypes.data_as(D)),
                R.ref_shape_in_cube(kind, p.ctypes.data_as(D), cc.ctypes.data_as(D), side),
                R.ref_shape_intersect(kind, p.ctypes.data_as(D), cc.ctypes.data_as(D), side))
        assert got == want, (kind, p, q, cc, side)


# --------------------------------------------------------- least squares --
def test_least_squares_known_answer():
    # wolkentest.cpp:744-759
    x = O.lls(np.array([[1, 3], [2, 4], [1, 6]], float), np.array([4, 1, 3], float))
    assert math.hypot(x[0] + 29 / 77, x[1] - 51 / 77) < 1e-9


def test_plane_fit_exact_plane():
    # noise-free points of a plane give back its coefficients (to rounding:
    # the elimination divides by pivots)
    rng = np.random.default_rng(5)
    for _ in range(20):
        a0, b0, c0 = rng.normal(size=3)
        x = rng.integers(-8, 8, 40).astype(float)
        y = rng.integers(-8, 8, 40).astype(float)
        z = a0 * x + b0 * y + c0
        a, b, c = O.lls3(x, y, z)
        assert abs(a - a0) < 1e-12 and abs(b - b0) < 1e-12 and abs(c - c0) < 1e-12


def test_plane_fit_collinear_is_nan():
    # all points on a line: a zero pivot gives NaN (leastsquares.cpp:39-41)
    x = np.arange(10, dtype=float)
    out = O.lls3(x, x, x)
    assert np.isnan(out).any()


# ------------------------------------------------------------- flowsnake --
def test_flowsnake_base_seven_inverts_base_flow():
    # wolkentest.cpp:792-796
    rng = np.random.default_rng(6)
    for m in list(range(-3000, 3000)) + list(rng.integers(-10**9, 10**9, 2000)):
        ex, ey = O.base_flow(int(m))
        assert O.base_seven(ex, ey) == m


def test_flowsnake_tile_counts_and_bijection():
    lo = [0, -4, -18, -214, -900, -10504, -44118]
    hi = [0, 2, 30, 128, 1500, 6302, 73530]
    for k in range(7):
        assert hi[k] - lo[k] + 1 == 7 ** k
        addrs = {O.to_flowsnake(n) for n in range(lo[k], hi[k] + 1)}
        assert len(addrs) == 7 ** k


def test_snake_setup_spacing():
    # a 200 m square tile picks 7^6 tiles of spacing 200/243.03
    sn = O.snake_setup([0, 0, 0], [200, 200, 30], 1.0)
    assert sn.index == 6 and (sn.lo, sn.hi) == (-44118, 73530)
    assert sn.spacing == 200 / 243.0343734125204


# ---------------------------------------------------------------- octree --
def morton_prefixes(x, y, z, center, side, levels):
    """Independent restatement of the compare chain: per level the child
    index z*4+y*2+x of the node containing each point."""
    cx, cy, cz = (np.full(len(x), center[i], dtype=np.float64) for i in range(3))
    s = side
    digits = []
    for _ in range(levels):
        xb = (x >= cx).astype(np.int64)
        yb = (y >= cy).astype(np.int64)
        zb = (z >= cz).astype(np.int64)
        digits.append(zb * 4 + yb * 2 + xb)
        cx = cx + (2 * xb - 1) * s / 4
        cy = cy + (2 * yb - 1) * s / 4
        cz = cz + (2 * zb - 1) * s / 4
        s = s / 2
    return np.stack(digits, 1)


@pytest.mark.parametrize("seed,extent,density", [(31, 20.0, 25.0), (32, 12.0, 80.0)])
def test_octree_invariants(seed, extent, density):
    cloud = synth.generate(seed=seed, extent=extent, density=density)
    x = cloud["offset"][0] + cloud["scale"][0] * cloud["xi"].astype(np.float64)
    y = cloud["offset"][1] + cloud["scale"][1] * cloud["yi"].astype(np.float64)
    z = cloud["offset"][2] + cloud["scale"][2] * cloud["zi"].astype(np.float64)
    center, side = O.sizefit(cloud["bbmin"], cloud["bbmax"])
    t = O.Tree(x, y, z, center, side)
    lv, order = t.leaves(), t.order()
    assert np.array_equal(np.sort(order), np.arange(len(x)))           # a partition
    dig = morton_prefixes(x, y, z, center, side, t.maxdepth)
    prefixes = [tuple(dig[order[f], :d]) for f, d in zip(lv["first"], lv["depth"])]
    assert all(a < b for a, b in zip(prefixes, prefixes[1:]))            # DFS = Morton order
    for k in range(len(lv["count"])):
        d, f, c = lv["depth"][k], lv["first"][k], lv["count"][k]
        pts = order[f:f + c]
        assert 1 <= c <= 537
        pref = dig[pts[0], :d]
        assert (dig[pts, :d] == pref).all()                               # same cell
        assert (order[f:f + c] == np.sort(pts)).all()                     # input order in a leaf
        same_leaf = (dig[:, :d] == pref).all(1).sum()
        assert same_leaf == c                                             # whole cell
        if d > 1:
            parent = (dig[:, :d - 1] == pref[:d - 1]).all(1).sum()
            assert parent > 537                                           # parent was split


def test_octree_split_boundaries():
    # 537 points stay one leaf under the root child; 538 split it
    # (all points in the root's +x+y+z child, octree.cpp:620-662, 849-876)
    rng = np.random.default_rng(7)
    for n in (537, 538, 4296, 4297):
        x = 10 + rng.random(n)
        y = 10 + rng.random(n)
        z = 10 + rng.random(n)
        t = O.Tree(x, y, z, [0.0, 0.0, 0.0], 64.0)
        lv = t.leaves()
        if n == 537:
            assert t.nleaves == 1 and lv["depth"][0] == 1 and lv["count"][0] == 537
        else:
            assert lv["depth"].min() > 1 and lv["count"].max() <= 537
        assert lv["count"].sum() == n


def test_octree_points_outside_root_cube():
    # points outside the root cube follow the compares to its extreme cells:
    # 537 of them beyond one corner are one leaf; 538 would need a level-21
    # cell of more than 537 points, which the canonical tree refuses; points
    # beyond one face but spread along it split as usual
    rng = np.random.default_rng(8)
    for n in (537, 538):
        x, y, z = -50 - rng.random(n), -50 - rng.random(n), -50 - rng.random(n)
        if n == 537:
            t = O.Tree(x, y, z, [0.0, 0.0, 0.0], 64.0)
            lv = t.leaves()
            assert t.nleaves == 1 and lv["depth"][0] == 1 and lv["count"][0] == 537
        else:
            with pytest.raises(ValueError):
                O.Tree(x, y, z, [0.0, 0.0, 0.0], 64.0)
    n = 5000
    x = -50 - rng.random(n)
    y, z = rng.uniform(-32, 32, n), rng.uniform(-32, 32, n)
    t = O.Tree(x, y, z, [0.0, 0.0, 0.0], 64.0)
    lv = t.leaves()
    assert lv["count"].sum() == n and lv["count"].max() <= 537
    # every leaf lies on the root's -x face
    assert np.all(lv["center"][:, 0] - lv["side"] / 2 == -32.0)


# ------------------------------------------------------- ground test ----
def dir_deg(deg):
    return int(round(deg / 360 * 2 ** 31)) - (2 ** 31 if deg >= 180 else 0)


def test_surround_known_sets():
    assert not O.surround([])
    assert not O.surround([5])
    assert not O.surround([0, 2 ** 30 - 1])
    assert O.surround([dir_deg(0), dir_deg(120), dir_deg(240)])
    # gaps 143.9, 72.1, 144.0 -> a 144 deg gap fails
    a, b = 0, 0x33333333 - 100
    c = b + (0x80000000 - 0x33333333 - (0x33333333 - 100))
    assert not O.surround([a, b, (c - 2 ** 31)])
    assert O.surround([0, 0x33333333 - 1, 0x66666666 - 2])
    # the degenerate pair at +-180 deg: the sorted-int test sees two equal gaps of 0
    assert O.surround([-2 ** 30, 2 ** 30])


def flat_cloud(n_side=40, spacing=0.5, z0=100.0):
    g = np.arange(n_side) * spacing
    X, Y = np.meshgrid(g, g)
    xi = np.round(X.ravel() / 0.001).astype(np.int32)
    yi = np.round(Y.ravel() / 0.001).astype(np.int32)
    zi = np.full(xi.shape, int(z0 / 0.001), np.int32)
    return xi, yi, zi


def make_cloud(xi, yi, zi):
    c = dict(xi=xi, yi=yi, zi=zi, scale=np.array([0.001] * 3), offset=np.array([1000.0, 2000.0, 0.0]))
    x = c["offset"][0] + 0.001 * xi.astype(float)
    y = c["offset"][1] + 0.001 * yi.astype(float)
    z = c["offset"][2] + 0.001 * zi.astype(float)
    c["bbmin"] = np.array([x.min(), y.min(), z.min()])
    c["bbmax"] = np.array([x.max(), y.max(), z.max()])
    return c


def test_flat_ground_is_ground():
    cls, tiles, st = O.run(make_cloud(*flat_cloud()))
    assert st.n_unclassified == 0
    assert (cls == 2).all()


def test_isolated_point_above_ground_is_not_ground():
    xi, yi, zi = flat_cloud()
    k = len(xi) // 2 + 20
    zi2 = zi.copy()
    zi2[k] += 5000                       # 5 m above the plane, in the middle
    cls, _, _ = O.run(make_cloud(xi, yi, zi2))
    assert cls[k] == 1
    assert (np.delete(cls, k) == 2).all()


def test_classification_order_independent():
    cloud = synth.generate(seed=41, extent=14.0, density=25.0)
    cls, _, _ = O.run(cloud)
    perm = np.random.default_rng(8).permutation(len(cloud["xi"]))
    c2 = dict(cloud)
    for k in ("xi", "yi", "zi"):
        c2[k] = cloud[k][perm]
    cls2, _, _ = O.run(c2)
    assert np.array_equal(cls2, cls[perm])


def test_early_exit_equals_exhaustive():
    cloud = synth.generate(seed=42, extent=14.0, density=25.0, outliers=0.02)
    fast, tf, _ = O.run(cloud)
    O.set_exhaustive(True)
    try:
        full, tx, _ = O.run(cloud)
    finally:
        O.set_exhaustive(False)
    assert np.array_equal(fast, full)
