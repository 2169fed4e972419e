"""The oracle's SCAN, POSTSCAN and CLASSIFY pinned against what the
mathematics fixes, on constructed clouds (CPU only).

The reference's scan.cpp, classify.cpp and matrix.cpp cannot be compiled here
(they need the cmake-generated config.h), so these restate no formula of the
oracle: expected values come from the construction itself (points placed in
known parts of a tile, exact planes), from the documented intent (the
density is the RMS of the densities of the tile's seven equal-area parts,
scan.cpp:38-52), from an independent walk of the six rays (the Eisenstein
units derived from exp(i pi k/3)), or from brute force over every pair of
points (classify.cpp:96-173 without the octree, the early exit or the
surround() bookkeeping).
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from wolkenbase_amd import synth

SQRT7 = math.sqrt(7.0)
DEG144 = 0x33333333  # angle.h:115, 2^31 units per turn
INT_MAX = 0x7FFFFFFF


def tiles_of(bbmin, bbmax, tile_size):
    """(counter-order list of (ex, ey, cx, cy, r), snake) of the tiling."""
    sn = O.snake_setup(bbmin, bbmax, tile_size)
    out = []
    for k in range(sn.hi - sn.lo + 1):
        ex, ey = O.to_flowsnake(sn.lo + k)
        out.append((ex, ey) + O.snake_cyl(sn, ex, ey))
    return out, sn


def cloud_mm(x, y, z):
    """A cloud of metre coordinates quantised to 1 mm (scale 0.001, offset 0)."""
    xi = np.round(np.asarray(x) * 1000).astype(np.int32)
    yi = np.round(np.asarray(y) * 1000).astype(np.int32)
    zi = np.round(np.asarray(z) * 1000).astype(np.int32)
    X, Y, Z = xi * 0.001, yi * 0.001, zi * 0.001
    return dict(xi=xi, yi=yi, zi=zi, scale=np.array([0.001] * 3), offset=np.zeros(3),
                bbmin=np.array([X.min(), Y.min(), Z.min()]),
                bbmax=np.array([X.max(), Y.max(), Z.max()]))


def place_parts(rng, cx, cy, r, counts):
    """counts[0] points in the inner disk (radius < r/sqrt7), counts[j] in
    the sector centred on (j-1)*60 deg (j = 1..6), well inside each part."""
    xs, ys = [], []
    for j, m in enumerate(counts):
        if j == 0:
            rho = rng.uniform(0.15, 0.8, m) * r / SQRT7
            th = rng.uniform(-np.pi, np.pi, m)
        else:
            rho = rng.uniform(1.15 / SQRT7, 0.92, m) * r
            th = np.radians((j - 1) * 60 + rng.uniform(-24, 24, m))
        xs.append(cx + rho * np.cos(th))
        ys.append(cy + rho * np.sin(th))
    return np.concatenate(xs), np.concatenate(ys)


# the corners of every constructed scan case: they fix the root cube
# (sizeFit of the box, z included) and so the tiling
ZLO, ZHI = -10.0, 60.0


def centre_tile(tile_size=1.0, E=20.0):
    tl, sn = tiles_of([0, 0, ZLO], [E, E, ZHI], tile_size)
    k = min(range(len(tl)), key=lambda i: math.hypot(tl[i][2] - E / 2, tl[i][3] - E / 2))
    return k, tl[k], sn


def scan_case(rng, counts, plane=(0.3, -0.2, 5.0), extra=(), tile_size=1.0, E=20.0):
    """A cloud with points only in one interior tile (placed by `counts`) and
    two far corners that fix the tiling; returns (tile record, r, spacing)."""
    k, _, sn = centre_tile(tile_size, E)
    tl, _ = tiles_of([0, 0, ZLO], [E, E, ZHI], tile_size)
    _, _, cx, cy, r = tl[k]
    x, y = place_parts(rng, cx, cy, r, counts)
    a, b, c = plane
    z = a * x + b * y + c
    ex_pts = [(x0, y0, a * x0 + b * y0 + c + dz) for (x0, y0, dz) in extra]
    xs = np.concatenate([x, [p[0] for p in ex_pts], [0.0, E]])
    ys = np.concatenate([y, [p[1] for p in ex_pts], [0.0, E]])
    zs = np.concatenate([z, [p[2] for p in ex_pts], [ZLO, ZHI]])
    cloud = cloud_mm(xs, ys, zs)
    assert ZLO < zs[:-2].min() and zs[:-2].max() < ZHI
    p = O.default_params()
    p.tile_size = tile_size
    _, tiles, _ = O.run(cloud, p)
    return tiles[k], r, sn.spacing


@pytest.mark.parametrize("counts", [[5] * 7, [3, 5, 7, 2, 4, 6, 1], [9, 0, 4, 0, 4, 0, 4]])
def test_scan_density_is_rms_of_seven_part_densities(counts):
    rng = np.random.default_rng(sum(counts))
    t, r, _ = scan_case(rng, counts)
    area = math.pi * r * r / 7  # each part: the inner disk or one sixth of the ring
    rms = math.sqrt(sum((h / area) ** 2 for h in counts) / 7)
    assert t["npoints"] == sum(counts)
    assert t["density"] == pytest.approx(rms, rel=1e-13)
    assert t["hs_scan"] == pytest.approx(math.sqrt(1 / rms + 0.1 ** 2), rel=1e-13)
    assert t["hs"] == t["hs_scan"]  # not a tree tile: postscan leaves it
    assert t["tree_flags"] == 0
    # the fitted plane is the one the points were placed on (to 1 mm rounding)
    assert abs(t["tilt_x"] - 0.3) < 2e-3 and abs(t["tilt_y"] + 0.2) < 2e-3
    assert t["height"] < 3e-3


def test_scan_uniform_tile_density_is_point_density():
    # equal counts in the seven equal-area parts: the RMS equals the mean,
    # N / (pi r^2)
    rng = np.random.default_rng(1)
    t, r, _ = scan_case(rng, [6] * 7)
    assert t["density"] == pytest.approx(42 / (math.pi * r * r), rel=1e-13)


def test_scan_slope_clamped_to_unit_length():
    rng = np.random.default_rng(2)
    t, _, _ = scan_case(rng, [6] * 7, plane=(1.5, 0.8, 1.0))
    g = math.hypot(1.5, 0.8)
    assert math.hypot(t["tilt_x"], t["tilt_y"]) == pytest.approx(1.0, abs=1e-12)
    assert t["tilt_x"] == pytest.approx(1.5 / g, abs=2e-3)
    assert t["tilt_y"] == pytest.approx(0.8 / g, abs=2e-3)


def test_scan_high_points_leave_the_bottom_layer():
    # a point 3 m above the plane is no part of the bottom density and makes
    # the tile a tree tile (height above 1.5 m)
    rng = np.random.default_rng(3)
    base, r, _ = scan_case(rng, [5] * 7)
    rng = np.random.default_rng(3)
    _, (_, _, cx, cy, _), _ = centre_tile()
    t, _, _ = scan_case(rng, [5] * 7, extra=[(cx + 0.01, cy + 0.02, 3.0)])
    assert t["npoints"] == 36
    assert t["density"] == base["density"]
    assert t["tree_flags"] == 1
    # (the plane is fitted to every point, the high one too, so the height
    # above it is 3 m only to within the tilt the point adds)
    assert t["height"] == pytest.approx(3.0, abs=0.1)


def test_scan_sparse_bottom_under_canopy_is_tree():
    # two bottom points (sum of squared part counts 2 < 7) and one point just
    # above the bottom layer (within 1.5 m): a tree tile by the count rule alone
    rng = np.random.default_rng(4)
    _, (_, _, cx, cy, r), _ = centre_tile(0.5)
    assert 2 * r + 0.1 < 1.5
    t, _, _ = scan_case(rng, [1, 1, 0, 0, 0, 0, 0], tile_size=0.5,
                        extra=[(cx + 0.01, cy - 0.01, 2 * r + 0.1)])
    assert t["npoints"] == 3 and t["height"] < 1.5
    assert t["tree_flags"] == 1
    # the same without the high point is not a tree tile
    rng = np.random.default_rng(4)
    t, _, _ = scan_case(rng, [1, 1, 0, 0, 0, 0, 0], tile_size=0.5)
    assert t["tree_flags"] == 0


def test_scan_second_lowest_point_sets_the_bottom_layer():
    # one stray point 5 m below: the layer is measured from the second
    # lowest point, so the plane points all stay in it
    rng = np.random.default_rng(5)
    base, _, _ = scan_case(rng, [4] * 7)
    rng = np.random.default_rng(5)
    _, (_, _, cx, cy, _), _ = centre_tile()
    t, r, _ = scan_case(rng, [4] * 7, extra=[(cx + 0.01, cy, -5.0)])
    # the stray point sits in the inner disk: that part holds 5, the others 4
    area = math.pi * r * r / 7
    rms = math.sqrt((5 ** 2 + 6 * 4 ** 2) / 7) / area
    assert t["density"] == pytest.approx(rms, rel=1e-13)
    assert t["tree_flags"] == 1 and t["height"] == pytest.approx(5.0, abs=0.2)


def eisenstein_units():
    """The six unit steps (a, b), a + b*omega = exp(i pi k / 3), omega =
    exp(2 pi i / 3), k = 0..5."""
    w = complex(math.cos(2 * math.pi / 3), math.sin(2 * math.pi / 3))
    out = []
    for k in range(6):
        u = complex(math.cos(math.pi * k / 3), math.sin(math.pi * k / 3))
        b = u.imag / w.imag
        a = u.real - b * w.real
        out.append((round(a), round(b)))
    return out


def test_postscan_rays_independent_walk():
    # flat ground with a canopy disk 5 m up: tiles under the canopy are tree
    # tiles; each widens its hyperboloid by the tree tiles met along six rays
    # before a ring holds a non-tree tile or nothing
    rng = np.random.default_rng(7)
    E = 24.0
    n = int(E * E * 25)
    x, y = rng.uniform(0, E, n), rng.uniform(0, E, n)
    z = 0.02 * rng.standard_normal(n)
    m = int(math.pi * 6.0 ** 2 * 12)
    rho, th = 6.0 * np.sqrt(rng.uniform(0, 1, m)), rng.uniform(-np.pi, np.pi, m)
    x = np.concatenate([x, E / 2 + rho * np.cos(th)])
    y = np.concatenate([y, E / 2 + rho * np.sin(th)])
    z = np.concatenate([z, 5.0 + 0.3 * rng.standard_normal(m)])
    cloud = cloud_mm(x, y, z)
    _, tiles, _ = O.run(cloud)
    tl, sn = tiles_of(cloud["bbmin"], cloud["bbmax"], 1.0)
    at = {(ex, ey): k for k, (ex, ey, _, _, _) in enumerate(tl)}
    units = eisenstein_units()
    counts = []
    for k, (ex, ey, _, _, _) in enumerate(tl):
        t = tiles[k]
        if t["npoints"] == 0:
            continue
        count = 0
        if t["tree_flags"] & 1:
            i = 1
            while True:
                ring = nontree = 0
                for (a, b) in units:
                    o = at.get((ex + a * i, ey + b * i))
                    if o is not None and tiles[o]["npoints"]:
                        ring += 1
                        if tiles[o]["tree_flags"] & 1:
                            count += 1
                        else:
                            nontree += 1
                i += 1
                if not ring or nontree:
                    break
        counts.append(count)
        want = math.sqrt(t["hs_scan"] ** 2 + (count * sn.spacing / 6) ** 2)
        assert t["hs"] == pytest.approx(want, rel=1e-15), (k, count)
    assert max(counts) >= 12  # deep inside the canopy the rays run several rings


def test_classification_brute_force():
    # pointsIn(cylinder) and pointsIn(hyperboloid) by testing every point,
    # the deciding tile as the last one in counter order, and the 144 deg
    # rule on the sorted directions, against the oracle's octree walk with
    # its early exit
    cloud = synth.generate(seed=88, extent=11.0, density=22.0, buildings=40, trees=300,
                           embankments=1, outliers=0.02)
    cls, tiles, _ = O.run(cloud)
    sc, of = cloud["scale"], cloud["offset"]
    x = of[0] + sc[0] * cloud["xi"].astype(np.float64)
    y = of[1] + sc[1] * cloud["yi"].astype(np.float64)
    z = of[2] + sc[2] * cloud["zi"].astype(np.float64)
    n = len(x)
    tl, _ = tiles_of(cloud["bbmin"], cloud["bbmax"], 1.0)
    owner = np.full(n, -1)
    for k, (_, _, cx, cy, r) in enumerate(tl):  # Cylinder::in, shape.cpp:214-218
        owner[np.hypot(cx - x, cy - y) <= r] = k
    assert (cls[owner < 0] == 0).all()
    pts = np.flatnonzero(owner >= 0)
    rng = np.random.default_rng(3)
    pts = pts if len(pts) <= 1500 else rng.choice(pts, 1500, replace=False)
    kinds = {1: 0, 2: 0}
    for p in pts:
        hs, s, th = tiles[owner[p]]["hs"], 1.0, 0.0
        por = hs * s * s  # Hyperboloid(v, r, s), shape.cpp:119-125
        czz = (z[p] - th) + por
        zd = czz - z
        xy = np.hypot(x[p] - x, y[p] - y)
        inside = (zd > 0) & (zd * zd - (xy * s) * (xy * s) >= por * por)
        inside &= np.hypot(x[p] - x, y[p] - y) != 0
        q = np.flatnonzero(inside)
        d = np.unique(O.atan2i_batch(y[q] - y[p], x[q] - x[p]))  # std::set<int>
        ok = len(d) > 1
        if ok:
            gaps = np.concatenate([(d[1:] - d[:-1]) & INT_MAX, [(int(d[0]) - int(d[-1])) & INT_MAX]])
            ok = bool((gaps < DEG144).all())
        want = 1 if ok else 2
        kinds[want] += 1
        assert cls[p] == want, (p, len(q))
    assert kinds[1] > 100 and kinds[2] > 100
