"""Writes tests/golden/ref_vectors.npz from the reference's own sources
(angle.cpp, point.cpp, shape.cpp, manysum.cpp built by `make -C oracle ref`
into oracle/_ref/libwbref.so).  Run in the container that has the reference
tree; the fixture is data (inputs and the reference's outputs).

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402


def main():
    R = O.ref()
    D = O._dp
    rng = np.random.default_rng(20240)
    # atan2i: random, LAS-like differences, axis and tiny cases
    n = 4000
    ys = np.concatenate([rng.normal(size=n), rng.integers(-3000, 3000, n) * 0.001,
                         [0.0, 0.0, 1.0, -1.0, 1e-300, -1e-300, 0.0], rng.normal(size=20) * 1e-12])
    xs = np.concatenate([rng.normal(size=n), rng.integers(-3000, 3000, n) * 0.001,
                         [1.0, -1.0, 0.0, 0.0, -1.0, -1.0, 0.0], rng.normal(size=20)])
    at = np.array([R.ref_atan2i(y, x) for y, x in zip(ys, xs)], dtype=np.int32)
    # dist (glibc hypot through xy::dist)
    da = rng.integers(-10**6, 10**6, (n, 4)) * 0.001 + np.array([5e5, 4.4e6, 5e5, 4.4e6])
    dist = np.array([R.ref_dist_xy(*r) for r in da])
    # shapes
    kinds, params, pts, cubes, sides, res = [], [], [], [], [], []
    for it in range(3000):
        kind = it % 4
        c = rng.normal(size=3) * 3
        if kind == 0:
            p = [c[0], c[1], abs(rng.normal()) * 2, 0, 0]
        elif kind == 1:
            p = [c[0], c[1], c[2], abs(rng.normal()) * 2, 0]
        elif kind == 2:
            p = [c[0], c[1], c[2], abs(rng.normal()), float(rng.choice([1.0, -1.0, 0.5]))]
        else:
            p = [c[0], c[1], c[2], rng.normal() * 2, 0]
        p = np.ascontiguousarray(p, dtype=np.float64)
        q = np.ascontiguousarray(rng.normal(size=3) * 3)
        cc = np.ascontiguousarray(np.round(rng.normal(size=3) * 4) / 4)
        side = float(2.0 ** rng.integers(-3, 3))
        res.append((R.ref_shape_in(kind, p.ctypes.data_as(D), q.ctypes.data_as(D)),
                    R.ref_shape_in_cube(kind, p.ctypes.data_as(D), cc.ctypes.data_as(D), side),
                    R.ref_shape_intersect(kind, p.ctypes.data_as(D), cc.ctypes.data_as(D), side)))
        kinds.append(kind); params.append(p); pts.append(q); cubes.append(cc); sides.append(side)
    # pairwise sums
    lens = np.array(list(range(0, 70)) + [127, 128, 129, 1000, 1025], dtype=np.int64)
    vals = rng.normal(size=int(lens.sum())) * 10.0 ** rng.integers(-6, 6, int(lens.sum()))
    offs = np.concatenate([[0], np.cumsum(lens)])
    sums = []
    for k in range(len(lens)):
        a = np.ascontiguousarray(vals[offs[k]:offs[k + 1]])
        sums.append(R.ref_pairwisesum(a.ctypes.data_as(D), len(a)))
    out = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(out, atan_y=ys, atan_x=xs, atan=at, dist_in=da, dist=dist,
                        shape_kind=np.array(kinds, np.int32), shape_params=np.array(params),
                        shape_point=np.array(pts), shape_cube=np.array(cubes),
                        shape_side=np.array(sides), shape_res=np.array(res, np.int8),
                        pw_len=lens, pw_vals=vals, pw_sum=np.array(sums))
    print(out, os.path.getsize(out), "bytes")


if __name__ == "__main__":
    main()
