"""The oracle and the host build of the device math against vectors the
reference's own sources produced (tests/golden/ref_vectors.npz, written by
tests/golden/make_ref_vectors.py).  Runs without the reference tree."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as O
import wolkenbase_amd

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_vectors.npz")
HM = os.path.join(os.path.dirname(wolkenbase_amd.__file__), "libwb_hostmath.so")
D = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def g():
    return np.load(GOLD)


def test_oracle_atan2i_golden(g):
    got = [O.atan2i(y, x) for y, x in zip(g["atan_y"], g["atan_x"])]
    assert np.array_equal(np.array(got, np.int32), g["atan"])


def test_device_atan2i_golden(g):
    L = C.CDLL(HM)
    L.wbh_atan2i.restype = C.c_int
    L.wbh_atan2i.argtypes = [C.c_double, C.c_double, C.POINTER(C.c_int)]
    L.wbh_atan2i_forced.restype = C.c_int
    L.wbh_atan2i_forced.argtypes = [C.c_double, C.c_double]
    u = C.c_int()
    got = [L.wbh_atan2i(y, x, C.byref(u)) for y, x in zip(g["atan_y"], g["atan_x"])]
    assert np.array_equal(np.array(got, np.int32), g["atan"])
    got = [L.wbh_atan2i_forced(y, x) for y, x in zip(g["atan_y"], g["atan_x"])]
    assert np.array_equal(np.array(got, np.int32), g["atan"])


def test_device_hypot_golden(g):
    L = C.CDLL(HM)
    L.wbh_hypot.restype = C.c_double
    L.wbh_hypot.argtypes = [C.c_double, C.c_double]
    a = g["dist_in"]
    got = np.array([L.wbh_hypot(r[0] - r[2], r[1] - r[3]) for r in a])
    assert np.array_equal(got.view(np.uint64), g["dist"].view(np.uint64))


def test_oracle_shapes_golden(g):
    for k, p, q, c, s, r in zip(g["shape_kind"], g["shape_params"], g["shape_point"],
                                g["shape_cube"], g["shape_side"], g["shape_res"]):
        got = (O.shape_in(int(k), p, q), O.shape_in_cube(int(k), p, c, s),
               O.shape_intersect(int(k), p, c, s))
        assert got == tuple(int(v) for v in r)


def test_device_shapes_golden(g):
    L = C.CDLL(HM)
    L.wbh_shape_in.argtypes = [C.c_int, D, D]
    L.wbh_shape_in_cube.argtypes = [C.c_int, D, D, C.c_double]
    L.wbh_shape_intersect.argtypes = [C.c_int, D, D, C.c_double]
    for k, p, q, c, s, r in zip(g["shape_kind"], g["shape_params"], g["shape_point"],
                                g["shape_cube"], g["shape_side"], g["shape_res"]):
        if k == 3:      # the paraboloid is not a device shape
            continue
        p, q, c = (np.ascontiguousarray(v) for v in (p, q, c))
        got = (L.wbh_shape_in(int(k), p.ctypes.data_as(D), q.ctypes.data_as(D)),
               L.wbh_shape_in_cube(int(k), p.ctypes.data_as(D), c.ctypes.data_as(D), s),
               L.wbh_shape_intersect(int(k), p.ctypes.data_as(D), c.ctypes.data_as(D), s))
        assert got == tuple(int(v) for v in r)


def test_oracle_pairwisesum_golden(g):
    offs = np.concatenate([[0], np.cumsum(g["pw_len"])])
    for k in range(len(g["pw_len"])):
        assert O.pairwisesum(g["pw_vals"][offs[k]:offs[k + 1]]) == g["pw_sum"][k]
