"""The device math (wb_math.h, wb_flowsnake.h, wb_host.cpp) built for the
host, pinned against glibc and the oracle without a GPU."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle_lib as O
import wolkenbase_amd

HM = os.path.join(os.path.dirname(wolkenbase_amd.__file__), "libwb_hostmath.so")
D = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def hm():
    L = C.CDLL(HM)
    L.wbh_hypot.restype = C.c_double
    L.wbh_hypot.argtypes = [C.c_double, C.c_double]
    L.wbh_atan2i.restype = C.c_int
    L.wbh_atan2i.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_int)]
    L.wbh_atan2i_forced.restype = C.c_int
    L.wbh_atan2i_forced.argtypes = [C.c_double, C.c_double]
    L.wbh_last_exact.restype = C.c_long
    L.wbh_last_exact.argtypes = [C.c_double, C.c_double]
    L.wbh_last_native.restype = C.c_long
    L.wbh_last_native.argtypes = [C.c_double, C.c_double]
    L.wbh_x80_op_check.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    L.wbh_x80_lrint_check.argtypes = [C.c_double, C.c_double]
    for f in ("wbh_shape_in",):
        getattr(L, f).argtypes = [C.c_int, D, D]
    L.wbh_shape_in_cube.argtypes = [C.c_int, D, D, C.c_double]
    L.wbh_shape_intersect.argtypes = [C.c_int, D, D, C.c_double]
    L.wbh_to_flowsnake.argtypes = [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.wbh_from_flowsnake.restype = C.c_int
    L.wbh_from_flowsnake.argtypes = [C.c_int] * 4
    L.wbh_snake_setup.argtypes = [D, D, C.c_double, D, C.POINTER(C.c_int)]
    L.wbh_size_fit.argtypes = [D, C.c_int, D, D]
    return L


def glibc_hypot():
    m = C.CDLL("libm.so.6")
    m.hypot.restype = C.c_double
    m.hypot.argtypes = [C.c_double, C.c_double]
    return m.hypot


def test_hypot_matches_glibc(hm):
    gh = glibc_hypot()
    rng = np.random.default_rng(11)
    n = 200000
    a = np.concatenate([
        (rng.integers(-10**6, 10**6, n) * 0.001 + 123456.0) - (rng.integers(-10**6, 10**6, n) * 0.001 + 123456.0),
        rng.random(n) * 100,
        np.ldexp(rng.random(n), rng.integers(-1060, 1060, n)),
        rng.integers(0, 2**62, n).astype(np.uint64).view(np.float64)])
    b = np.concatenate([
        rng.integers(-3000, 3000, n) * 0.001,
        rng.random(n) * 100,
        np.ldexp(rng.random(n), rng.integers(-1060, 1060, n)),
        rng.integers(0, 2**62, n).astype(np.uint64).view(np.float64)])
    bad = 0
    for x, y in zip(a, b):
        if not (np.isfinite(x) and np.isfinite(y)):
            continue
        u, v = hm.wbh_hypot(x, y), gh(x, y)
        if np.float64(u).view(np.uint64) != np.float64(v).view(np.uint64):
            bad += 1
    assert bad == 0


def test_hypot_special(hm):
    inf, nan = float("inf"), float("nan")
    assert hm.wbh_hypot(inf, nan) == inf
    assert np.isnan(hm.wbh_hypot(nan, 1.0))
    assert hm.wbh_hypot(0.0, -0.0) == 0.0
    assert hm.wbh_hypot(3.0, 4.0) == 5.0


def test_atan2i_matches_oracle(hm):
    L = O.lib()
    rng = np.random.default_rng(12)
    n = 100000
    ys = np.concatenate([rng.normal(size=n), rng.integers(-4000, 4000, n) * 0.001, np.zeros(20)])
    xs = np.concatenate([rng.normal(size=n), rng.integers(-4000, 4000, n) * 0.001,
                         rng.normal(size=20)])
    used = C.c_int(0)
    bad = [(y, x) for y, x in zip(ys, xs) if hm.wbh_atan2i(y, x, C.byref(used)) != L.wbo_atan2i(y, x)]
    assert not bad, bad[:5]


def test_atan2i_exact_path_matches_oracle(hm):
    # every last step through the 64-bit-significand emulation
    L = O.lib()
    rng = np.random.default_rng(13)
    for y, x in zip(rng.normal(size=30000), rng.normal(size=30000)):
        assert hm.wbh_atan2i_forced(y, x) == L.wbo_atan2i(y, x)


def exact_bin(a):
    return (((a + 0x40000000) & 0xffffffff) & 0x7fffffff) >> 25


def test_dir_bin_matches_exact_bin(hm):
    hm.wbh_dir_bin.restype = C.c_int
    hm.wbh_dir_bin.argtypes = [C.c_double, C.c_double]
    L = O.lib()
    rng = np.random.default_rng(19)
    n = 100000
    # random, LAS-like, axis-aligned, diagonal, and directions placed near bin
    # edges: within 70000 units, and within 4096 units (the float path's
    # margin is 2048 units, its error a few hundred)
    base = (np.arange(64) * 2 ** 25 - 2 ** 30)
    edges = np.concatenate([base[rng.integers(0, 64, n)] + rng.integers(-70000, 70000, n),
                            base[rng.integers(0, 64, n)] + rng.integers(-4096, 4097, n)])
    ang = edges / 2 ** 31 * 2 * np.pi
    r = rng.uniform(0.001, 500, 2 * n)
    d = rng.integers(1, 4000, 20) * 0.001
    sx = rng.choice([-1.0, 1.0], 20)
    sy = rng.choice([-1.0, 1.0], 20)
    ys = np.concatenate([rng.normal(size=n), rng.integers(-4000, 4000, n) * 0.001, r * np.sin(ang),
                         np.zeros(10), rng.normal(size=10), sy * d])
    xs = np.concatenate([rng.normal(size=n), rng.integers(-4000, 4000, n) * 0.001, r * np.cos(ang),
                         rng.normal(size=10), np.zeros(10), sx * d])
    # and magnitudes a float cannot carry, or only as a subnormal
    k = rng.integers(0, 2000, 400)
    sc = np.array([1e-45, 1e-38, 1e-31, 1e31, 1e38, 1e45])[rng.integers(0, 6, 400)]
    ys = np.concatenate([ys, ys[k] * sc, ys[k] * sc[::-1]])
    xs = np.concatenate([xs, xs[k] * sc, xs[k]])
    bad = 0
    for y, x in zip(ys, xs):
        if x == 0 and y == 0:
            continue
        if hm.wbh_dir_bin(y, x) != exact_bin(L.wbo_atan2i(y, x)):
            bad += 1
    assert bad == 0


def care_bins(occ):
    """Bins of the empty circular runs of at least 24 bins."""
    empty = [not (occ >> b) & 1 for b in range(64)]
    care = set()
    if not any(empty):
        return care
    if all(empty):
        return set(range(64))
    for st in range(64):
        if empty[st] and not empty[(st - 1) % 64]:
            run = []
            while empty[(st + len(run)) % 64]:
                run.append((st + len(run)) % 64)
            if len(run) >= 24:
                care.update(run)
    return care


def test_care_runs_match_brute_force(hm):
    hm.wbh_care_mask.restype = C.c_uint64
    hm.wbh_care_mask.argtypes = [C.c_uint64]
    rng = np.random.default_rng(22)
    masks = [0, 1, 1 << 63, (1 << 64) - 1, 1 | (1 << 32), 1 | (1 << 25), 1 | (1 << 24)]
    for _ in range(20000):
        nb = int(rng.integers(1, 10))
        occ = 0
        for b in rng.integers(0, 64, nb):
            occ |= 1 << int(b)
        masks.append(occ)
    for occ in masks:
        want = sum(1 << b for b in care_bins(occ)) if occ else (1 << 64) - 1
        assert hm.wbh_care_mask(occ) == want, hex(occ)


def test_wedge_pruning_is_conservative(hm):
    # a box the wedge test rejects holds no direction in a run of >= 24 empty
    # bins: brute force over points inside rejected boxes (corners, edges and
    # random interior points), with bins from the exact direction binning
    hm.wbh_wedge_hit.restype = C.c_int
    hm.wbh_wedge_hit.argtypes = [C.c_uint64] + [C.c_double] * 6
    hm.wbh_dir_bin.restype = C.c_int
    hm.wbh_dir_bin.argtypes = [C.c_double, C.c_double]
    rng = np.random.default_rng(21)
    rejected = 0
    for it in range(3000):
        nb = int(rng.integers(1, 12))
        occ = 0
        start = int(rng.integers(0, 64))
        for b in rng.integers(0, 40, nb):       # bins clustered on one side
            occ |= 1 << int((start + b) % 64)
        care = care_bins(occ)
        px, py = rng.normal(size=2) * 50 + np.array([500000.0, 4400000.0])
        w, h = rng.uniform(0.01, 40, 2)
        cx, cy = px + rng.normal() * 60, py + rng.normal() * 60
        x0, x1, y0, y1 = cx - w, cx + w, cy - h, cy + h
        if hm.wbh_wedge_hit(occ, px, py, x0, x1, y0, y1):
            continue
        rejected += 1
        us = np.concatenate([[0, 1, 0, 1, 0.5, 0.5, 0, 1], rng.random(60)])
        vs = np.concatenate([[0, 0, 1, 1, 0, 1, 0.5, 0.5], rng.random(60)])
        for u, v in zip(us, vs):
            qx, qy = x0 + u * (x1 - x0), y0 + v * (y1 - y0)
            if qx == px and qy == py:
                continue
            assert hm.wbh_dir_bin(qy - py, qx - px) not in care, (it, occ, u, v)
    assert rejected > 300
    # no directions yet: nothing may be skipped
    assert hm.wbh_wedge_hit(0, 0.0, 0.0, 5.0, 6.0, 5.0, 6.0) == 1
    # runs of exactly 24, 31, 32 (more than pi with the margins), 33, 40 and
    # two runs, with small boxes all round p at every angle
    specials = []
    for run in (24, 31, 32, 33, 40, 62, 63):
        s = int(rng.integers(0, 64))
        specials.append(sum(1 << ((s + run + k) % 64) for k in range(64 - run)))
    specials.append(((1 << 64) - 1) & ~(((1 << 24) - 1) | (((1 << 25) - 1) << 30)))
    for occ in specials:
        care = care_bins(occ)
        # run edges, the edges half a turn away, and sub-margin offsets
        starts = [b for b in care if (b - 1) % 64 not in care]
        ends = [(b + 1) % 64 for b in care if (b + 1) % 64 not in care]
        # and the middle edge of every run of 32 bins or more, where it is split
        mids = []
        for b in starts:
            n = 0
            while (b + n) % 64 in care and n < 64:
                n += 1
            if n >= 32:
                mids.append((b + n // 2) % 64)
        edges = [-np.pi + k * np.pi / 32 for k in starts + ends + mids]
        edges += [e + np.pi for e in edges]
        for e in edges:
            for off in (-3e-3, -1e-3, -6e-4, -2e-4, -1e-5, 0.0, 1e-5, 2e-4, 6e-4, 1e-3, 3e-3):
                r = rng.uniform(5, 80)
                cx, cy = r * np.cos(e + off), r * np.sin(e + off)
                w = 1e-5 * r
                x0, x1, y0, y1 = cx - w, cx + w, cy - w, cy + w
                if hm.wbh_wedge_hit(occ, 0.0, 0.0, x0, x1, y0, y1):
                    continue
                for qx, qy in ((x0, y0), (x0, y1), (x1, y0), (x1, y1), (cx, cy)):
                    assert hm.wbh_dir_bin(qy, qx) not in care, (hex(occ), e, off)
        for ang in np.linspace(-np.pi, np.pi, 721):
            r = rng.uniform(0.5, 80)
            cx, cy = r * np.cos(ang), r * np.sin(ang)
            w = rng.uniform(0.001, 0.3 * r)
            x0, x1, y0, y1 = cx - w, cx + w, cy - w, cy + w
            if hm.wbh_wedge_hit(occ, 0.0, 0.0, x0, x1, y0, y1):
                continue
            for u, v in zip(rng.random(40), rng.random(40)):
                qx, qy = x0 + u * (x1 - x0), y0 + v * (y1 - y0)
                assert hm.wbh_dir_bin(qy, qx) not in care, (hex(occ), ang, u, v)


def test_scan_sector_fast_form(hm):
    # the comparison form of the scan's sector (sqrt(3)|vy| against |vx|, the
    # signs, and the squared radius against (r/sqrt 7)^2) equals the atan2 /
    # lrint / hypot form on random points and on points placed just either
    # side of every sector boundary, of the axis lines and of the inner radius
    hm.wbh_scan_sector.restype = C.c_int
    hm.wbh_scan_sector.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
    rng = np.random.default_rng(29)
    pts = []
    for _ in range(20000):
        r = float(rng.uniform(0.3, 20))
        q = rng.normal(size=2) * r
        pts.append((float(q[0]), float(q[1]), r))
    for k in range(12):  # boundaries every 30 degrees (half of them sector edges)
        for off in (0.0, 1e-15, -1e-15, 1e-12, -1e-12, 1e-10, -1e-10, 1e-8, -1e-8, 1e-4):
            for rad in (0.05, 1.0, 7.0):
                t = k * np.pi / 6 + off
                pts.append((rad * np.cos(t), rad * np.sin(t), 2.0))
    for vy in (0.0, -0.0, 1e-300, -1e-300):
        for vx in (1.0, -1.0, 0.0, -0.0):
            pts.append((vx, vy, 1.0))
    for off in (0.0, 1e-16, -1e-16, 1e-13, -1e-13, 1e-11, -1e-11):  # the inner radius
        r = 3.0
        rr = r / 2.6457513110645905905016 * (1 + off)
        for t in rng.uniform(-np.pi, np.pi, 20):
            pts.append((rr * np.cos(t), rr * np.sin(t), r))
    fast = exact = 0
    for vx, vy, r in pts:
        a = hm.wbh_scan_sector(0, vx, vy, r)
        b = hm.wbh_scan_sector(1, vx, vy, r)
        assert a == b, (vx, vy, r, a, b)
        fast += 1
    assert fast == len(pts)


def test_x87_emulation_ops(hm):
    rng = np.random.default_rng(14)
    for _ in range(40000):
        a = rng.normal() * 2.0 ** rng.integers(-30, 30)
        b = rng.normal() * 2.0 ** rng.integers(-30, 30)
        c = rng.normal() * 2.0 ** rng.integers(-5, 5)
        for op in range(3):
            assert hm.wbh_x80_op_check(op, a, b, c), (op, a, b, c)
        assert hm.wbh_x80_lrint_check(a if abs(a) < 1e15 else 1.0, c)
    for k in range(-3000, 3000):               # exact ties round to even
        assert hm.wbh_x80_lrint_check(k + 0.5, 1.0)


def test_last_step_exact_vs_native(hm):
    rng = np.random.default_rng(15)
    for _ in range(50000):
        x = abs(rng.normal()) + 1e-3
        y = x * rng.uniform(-0.004, 0.004)     # the residual range after the table search
        assert hm.wbh_last_exact(y, x) == hm.wbh_last_native(y, x)


def test_shapes_match_oracle(hm):
    rng = np.random.default_rng(16)
    for it in range(30000):
        kind = it % 3
        c = rng.normal(size=3) * 3
        if kind == 0:
            p = np.array([c[0], c[1], abs(rng.normal()) * 2, 0, 0])
        elif kind == 1:
            p = np.array([c[0], c[1], c[2], abs(rng.normal()) * 2, 0])
        else:
            p = np.array([c[0], c[1], c[2], abs(rng.normal()), rng.choice([1.0, -1.0, 0.5])])
        q = rng.normal(size=3) * 3
        cc = np.round(rng.normal(size=3) * 4) / 4
        side = 2.0 ** rng.integers(-3, 3)
        pp, qq, ccc = (np.ascontiguousarray(v) for v in (p, q, cc))
        got = (hm.wbh_shape_in(kind, pp.ctypes.data_as(D), qq.ctypes.data_as(D)),
               hm.wbh_shape_in_cube(kind, pp.ctypes.data_as(D), ccc.ctypes.data_as(D), side),
               hm.wbh_shape_intersect(kind, pp.ctypes.data_as(D), ccc.ctypes.data_as(D), side))
        want = (O.shape_in(kind, p[:4 if kind else 3], q), O.shape_in_cube(kind, p[:4 if kind else 3], cc, side),
                O.shape_intersect(kind, p[:4 if kind else 3], cc, side)) if kind < 2 else \
               (O.shape_in(kind, p, q), O.shape_in_cube(kind, p, cc, side), O.shape_intersect(kind, p, cc, side))
        assert got == want, (kind, p, q, cc, side)


def test_shape_boundary_cases_take_exact_path(hm):
    # points placed on the cylinder boundary to within an ulp exercise the
    # hypot fallback of the fast filter
    rng = np.random.default_rng(17)
    for _ in range(20000):
        r = rng.uniform(0.1, 5)
        th = rng.uniform(0, 2 * np.pi)
        q = np.array([r * np.cos(th), r * np.sin(th), 0.0])
        p = np.array([0.0, 0.0, r, 0, 0])
        got = hm.wbh_shape_in(0, p.ctypes.data_as(D), q.ctypes.data_as(D))
        assert got == O.shape_in(0, p[:3], q)


def test_flowsnake_round_trip(hm):
    lo = [0, -4, -18, -214, -900, -10504, -44118]
    hi = [0, 2, 30, 128, 1500, 6302, 73530]
    for k in range(7):
        for n in range(lo[k], hi[k] + 1, max(1, (hi[k] - lo[k]) // 3000)):
            ex, ey = C.c_int(), C.c_int()
            hm.wbh_to_flowsnake(n, C.byref(ex), C.byref(ey))
            assert (ex.value, ey.value) == O.to_flowsnake(n)
            assert hm.wbh_from_flowsnake(ex.value, ey.value, lo[k], hi[k]) == n
    # a lattice point outside the island has no counter
    assert hm.wbh_from_flowsnake(10**6, 0, lo[6], hi[6]) == -2**31


def test_snake_setup_and_size_fit_match_oracle(hm):
    rng = np.random.default_rng(18)
    for _ in range(200):
        lo = rng.normal(size=3) * 1e5 + np.array([5e5, 4.4e6, 0])
        hi = lo + rng.uniform(1, 3000, 3)
        out4 = np.zeros(4)
        out3 = (C.c_int * 3)()
        hm.wbh_snake_setup(lo.ctypes.data_as(D), hi.ctypes.data_as(D), 1.0,
                           out4.ctypes.data_as(D), out3)
        sn = O.snake_setup(lo, hi, 1.0)
        assert (out4[0], out4[1], out4[2]) == (sn.cx, sn.cy, sn.spacing)
        assert tuple(out3) == (sn.lo, sn.hi, sn.index)
        pts = np.ascontiguousarray(np.concatenate([lo, hi]))
        c, s = np.zeros(3), np.zeros(1)
        hm.wbh_size_fit(pts.ctypes.data_as(D), 2, c.ctypes.data_as(D), s.ctypes.data_as(D))
        oc, os_ = O.sizefit(lo, hi)
        assert np.array_equal(c, oc) and s[0] == os_


def border_care_bins(occ):
    """Runs of 24 or more empty bins with the occupied bin on each side."""
    care = set()
    for st in range(64):
        if not (occ >> st) & 1 and (occ >> ((st - 1) % 64)) & 1:
            n = 0
            while not (occ >> ((st + n) % 64)) & 1:
                n += 1
            if n >= 24:
                care |= {(st + k) % 64 for k in range(-1, n + 1)}
    return care


def test_border_wedge_pruning_is_conservative(hm):
    # the exact border pass (a longest empty run of 24 or 25 bins) skips a box
    # only if it holds no direction in the run or its two border bins
    hm.wbh_border_wedge_hit.restype = C.c_int
    hm.wbh_border_wedge_hit.argtypes = [C.c_uint64] + [C.c_double] * 6
    hm.wbh_dir_bin.restype = C.c_int
    hm.wbh_dir_bin.argtypes = [C.c_double, C.c_double]
    rng = np.random.default_rng(23)
    masks = []
    for it in range(400):
        runs = [int(rng.choice([24, 25]))] + ([int(rng.choice([24, 25]))] if it % 3 == 0 else [])
        occ = (1 << 64) - 1
        s = int(rng.integers(0, 64))
        for r in runs:
            occ &= ~sum(1 << ((s + k) % 64) for k in range(r))
            s = (s + r + 1 + int(rng.integers(0, 6))) % 64
        # thin out the occupied bins away from the runs' borders
        care = border_care_bins(occ)
        for b in range(64):
            if b not in care and rng.random() < 0.3:
                occ &= ~(1 << b)
        # the exact pass sees a longest run of 24 or 25 bins, at most two of them
        lens = []
        for st in range(64):
            if not (occ >> st) & 1 and (occ >> ((st - 1) % 64)) & 1:
                n = 0
                while not (occ >> ((st + n) % 64)) & 1:
                    n += 1
                lens.append(n)
        big = [n for n in lens if n >= 24]
        if not big or max(big) > 25 or len(big) > 2:
            continue
        masks.append(occ)
    rejected = 0
    for it, occ in enumerate(masks):
        care = border_care_bins(occ)
        for rep in range(12):
            px, py = rng.normal(size=2) * 50 + np.array([500000.0, 4400000.0])
            w, h = rng.uniform(0.01, 40, 2)
            cx, cy = px + rng.normal() * 60, py + rng.normal() * 60
            x0, x1, y0, y1 = cx - w, cx + w, cy - h, cy + h
            if hm.wbh_border_wedge_hit(occ, px, py, x0, x1, y0, y1):
                continue
            rejected += 1
            us = np.concatenate([[0, 1, 0, 1, 0.5, 0.5, 0, 1], rng.random(40)])
            vs = np.concatenate([[0, 0, 1, 1, 0, 1, 0.5, 0.5], rng.random(40)])
            for u, v in zip(us, vs):
                qx, qy = x0 + u * (x1 - x0), y0 + v * (y1 - y0)
                if qx == px and qy == py:
                    continue
                assert hm.wbh_dir_bin(qy - py, qx - px) not in care, (it, hex(occ), u, v)
        # tiny boxes at every care edge and its opposite, sub-margin offsets
        starts = [b for b in care if (b - 1) % 64 not in care]
        ends = [(b + 1) % 64 for b in care if (b + 1) % 64 not in care]
        # and the middle edge of every run of 32 bins or more, where it is split
        mids = []
        for b in starts:
            n = 0
            while (b + n) % 64 in care and n < 64:
                n += 1
            if n >= 32:
                mids.append((b + n // 2) % 64)
        edges = [-np.pi + k * np.pi / 32 for k in starts + ends + mids]
        edges += [e + np.pi for e in edges]
        for e in edges:
            for off in (-2e-3, -1e-3, -5e-4, -1e-5, 0.0, 1e-5, 5e-4, 1e-3, 2e-3):
                r = rng.uniform(5, 80)
                cx, cy = r * np.cos(e + off), r * np.sin(e + off)
                d = 1e-5 * r
                x0, x1, y0, y1 = cx - d, cx + d, cy - d, cy + d
                if hm.wbh_border_wedge_hit(occ, 0.0, 0.0, x0, x1, y0, y1):
                    continue
                for qx, qy in ((x0, y0), (x0, y1), (x1, y0), (x1, y1), (cx, cy)):
                    assert hm.wbh_dir_bin(qy, qx) not in care, (hex(occ), e, off)
    assert len(masks) > 200 and rejected > 500


def test_membership_candidates_cover_every_tile_within_reach():
    """k_membership takes as candidates the lattice point nearest to a point
    (cube rounding of its Eisenstein coordinates) and that point's six
    neighbours.  Every tile centre within the candidate bound (spacing*41/71
    with its margin) must be among them: checked against all lattice points
    within two steps, on random points and on points placed at hexagon
    vertices (three-way ties of the rounding) within 1e-12 relative."""
    rng = np.random.default_rng(71)
    s = 1.518
    lim = s * 41 / 71 * (1 + 1e-9) + 1e-9
    k34 = math.sqrt(3) / 2

    def centre(a, b):
        return s * (a - b / 2), s * (b * k34)

    pts = [tuple(p) for p in rng.uniform(-60, 60, (20000, 2))]
    for _ in range(4000):
        a, b = (int(v) for v in rng.integers(-30, 30, 2))
        cx, cy = centre(a, b)
        ang = int(rng.integers(0, 12)) * math.pi / 6
        rr = s / math.sqrt(3) * (1 + rng.normal() * 1e-12)
        pts.append((cx + rr * math.cos(ang), cy + rr * math.sin(ang)))
    for x, y in pts:
        fb = (y / s) / k34
        fa = x / s + fb / 2
        rx, ry, rz = np.rint(fa - fb), np.rint(fb), np.rint(-fa)
        ddx, ddy, ddz = abs(rx - (fa - fb)), abs(ry - fb), abs(rz + fa)
        if ddy > ddx and ddy > ddz:
            ry = -rx - rz
        elif ddz >= ddx and ddz >= ddy:
            rz = -rx - ry
        X, Y = int(-rz), int(ry)
        seven = {(X + da, Y + db) for da, db in
                 [(0, 0), (1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, -1)]}
        b0, a0 = round(fb), round(fa)
        for b in range(b0 - 3, b0 + 4):
            for a in range(a0 - 3, a0 + 4):
                cx, cy = centre(a, b)
                if (cx - x) ** 2 + (cy - y) ** 2 <= lim * lim:
                    assert (a, b) in seven, (x, y, a, b)
