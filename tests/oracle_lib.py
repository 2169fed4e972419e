"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and, where it was
built, to oracle/_ref/libwbref.so (the reference's own geometry sources).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  Never imported by the wolkenbase_amd package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libwbref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_lp = C.POINTER(C.c_long)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)

SH_CYL, SH_SPH, SH_HYP, SH_PAR = 0, 1, 2, 3


class Params(C.Structure):
    _fields_ = [("tile_size", C.c_double), ("max_slope", C.c_double),
                ("thickness", C.c_double), ("min_hs", C.c_double)]


class Tile(C.Structure):
    _fields_ = [("npoints", C.c_int), ("nground", C.c_int), ("tree_flags", C.c_int),
                ("density", C.c_double), ("hs_scan", C.c_double), ("hs", C.c_double),
                ("height", C.c_double), ("tilt_x", C.c_double), ("tilt_y", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("n_class1", C.c_long), ("n_class2", C.c_long), ("n_unclassified", C.c_long),
                ("ntiles", C.c_long), ("nonempty_tiles", C.c_long),
                ("sector_borderline", C.c_long)]


class Snake(C.Structure):
    _fields_ = [("cx", C.c_double), ("cy", C.c_double), ("spacing", C.c_double),
                ("lo", C.c_int), ("hi", C.c_int), ("index", C.c_int)]


def default_params():
    """Reference defaults, mainwindow.cpp:403-410."""
    return Params(1.0, 1.0, 0.0, 0.1)


def _ptr(a, t):
    return a.ctypes.data_as(t)


def ensure_built():
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "all"])


_lib = None


def lib():
    global _lib
    if _lib is None:
        ensure_built()
        L = C.CDLL(ORACLE_SO)
        L.wbo_atan2i.restype = C.c_int
        L.wbo_atan2i.argtypes = [C.c_double, C.c_double]
        L.wbo_pairwisesum.restype = C.c_double
        L.wbo_pairwisesum.argtypes = [_dp, C.c_uint]
        L.wbo_lls3.argtypes = [_dp, _dp, _dp, C.c_uint, _dp]
        L.wbo_lls.argtypes = [_dp, C.c_int, C.c_int, _dp, _dp]
        for f in ("wbo_in_flat",):
            getattr(L, f).argtypes = [C.c_int, _dp, _dp]
        L.wbo_in_cube_flat.argtypes = [C.c_int, _dp, _dp, C.c_double]
        L.wbo_intersect_flat.argtypes = [C.c_int, _dp, _dp, C.c_double]
        L.wbo_sizefit.argtypes = [_dp, C.c_int, _dp, _dp]
        L.wbo_build.restype = C.c_void_p
        L.wbo_failed.argtypes = [C.c_void_p]
        L.wbo_failed.restype = C.c_int
        L.wbo_build.argtypes = [_dp, _dp, _dp, C.c_long, _dp, C.c_double]
        L.wbo_free.argtypes = [C.c_void_p]
        L.wbo_nleaves.argtypes = [C.c_void_p]
        L.wbo_nnodes.argtypes = [C.c_void_p]
        L.wbo_maxdepth.argtypes = [C.c_void_p]
        L.wbo_leaves.argtypes = [C.c_void_p, _dp, _dp, _ip, _lp, _lp]
        L.wbo_order.argtypes = [C.c_void_p, _lp]
        L.wbo_points_in_flat.restype = C.c_long
        L.wbo_points_in_flat.argtypes = [C.c_void_p, C.c_int, _dp, _lp, C.c_long]
        L.wbo_to_flowsnake.argtypes = [C.c_int, _ip, _ip]
        L.wbo_base_seven.restype = C.c_int
        L.wbo_base_seven.argtypes = [C.c_int, C.c_int]
        L.wbo_snake_setup.argtypes = [_dp, _dp, C.c_double, C.POINTER(Snake)]
        L.wbo_snake_cyl.argtypes = [C.POINTER(Snake), C.c_int, C.c_int, _dp, _dp, _dp]
        L.wbo_surround.restype = C.c_int
        L.wbo_surround.argtypes = [_ip, C.c_int]
        L.wbo_run.restype = C.c_long
        L.wbo_run.argtypes = [_i32p, _i32p, _i32p, C.c_long, _dp, _dp, _dp, _dp,
                              C.POINTER(Params), _u8p, C.POINTER(Tile), C.POINTER(Stats),
                              _dp, _dp, _dp]
        L.wbo_classify_points.argtypes = [C.c_void_p, _lp, C.c_long, _dp,
                                          C.POINTER(Params), _u8p]
        L.wbo_get_tables.argtypes = [_dp, _dp, _dp]
        L.wbo_sin_bin.restype = C.c_double
        L.wbo_sin_bin.argtypes = [C.c_int]
        L.wbo_cos_bin.restype = C.c_double
        L.wbo_cos_bin.argtypes = [C.c_int]
        L.wbo_base_flow.argtypes = [C.c_int, _ip, _ip]
        L.wbo_i_to_flowsnake.restype = C.c_int
        L.wbo_i_to_flowsnake.argtypes = [C.c_int]
        L.wbo_set_exhaustive.argtypes = [C.c_int]
        L.wbo_atan2i_batch.argtypes = [_dp, _dp, C.c_long, _ip]
        L.wbo_fill_tables()
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


_ref = None


def ref():
    """The reference's own sources built from source (oracle/_ref)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_SO)
        R.ref_atan2i.restype = C.c_int
        R.ref_atan2i.argtypes = [C.c_double, C.c_double]
        for f in ("ref_sin_bin", "ref_cos_bin", "ref_tan_bin"):
            getattr(R, f).restype = C.c_double
            getattr(R, f).argtypes = [C.c_int]
        R.ref_pairwisesum.restype = C.c_double
        R.ref_pairwisesum.argtypes = [_dp, C.c_uint]
        R.ref_dist_xy.restype = C.c_double
        R.ref_dist_xy.argtypes = [C.c_double] * 4
        R.ref_dir.restype = C.c_int
        R.ref_dir.argtypes = [C.c_double] * 4
        R.ref_shape_in.argtypes = [C.c_int, _dp, _dp]
        R.ref_shape_in_cube.argtypes = [C.c_int, _dp, _dp, C.c_double]
        R.ref_shape_intersect.argtypes = [C.c_int, _dp, _dp, C.c_double]
        R.ref_fill_tables()
        _ref = R
    return _ref


def tables():
    L = lib()
    t = np.zeros(511)
    c = np.zeros(512)
    s = np.zeros(512)
    L.wbo_get_tables(_ptr(t, _dp), _ptr(c, _dp), _ptr(s, _dp))
    return t, c, s


def atan2i(y, x):
    return lib().wbo_atan2i(float(y), float(x))


def pairwisesum(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().wbo_pairwisesum(_ptr(a, _dp), len(a))


def lls3(x, y, z):
    x, y, z = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, z))
    out = np.zeros(3)
    lib().wbo_lls3(_ptr(x, _dp), _ptr(y, _dp), _ptr(z, _dp), len(x), _ptr(out, _dp))
    return out


def lls(m, v):
    m = np.ascontiguousarray(m, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(m.shape[1])
    lib().wbo_lls(_ptr(m, _dp), m.shape[0], m.shape[1], _ptr(v, _dp), _ptr(out, _dp))
    return out


def shape_in(kind, params, q):
    p = np.ascontiguousarray(params, dtype=np.float64)
    qq = np.ascontiguousarray(q, dtype=np.float64)
    return lib().wbo_in_flat(kind, _ptr(p, _dp), _ptr(qq, _dp))


def shape_in_cube(kind, params, cc, side):
    p = np.ascontiguousarray(params, dtype=np.float64)
    c = np.ascontiguousarray(cc, dtype=np.float64)
    return lib().wbo_in_cube_flat(kind, _ptr(p, _dp), _ptr(c, _dp), float(side))


def shape_intersect(kind, params, cc, side):
    p = np.ascontiguousarray(params, dtype=np.float64)
    c = np.ascontiguousarray(cc, dtype=np.float64)
    return lib().wbo_intersect_flat(kind, _ptr(p, _dp), _ptr(c, _dp), float(side))


def sizefit(bbmin, bbmax):
    pts = np.ascontiguousarray(np.concatenate([bbmin, bbmax]), dtype=np.float64)
    c = np.zeros(3)
    s = np.zeros(1)
    lib().wbo_sizefit(_ptr(pts, _dp), 2, _ptr(c, _dp), _ptr(s, _dp))
    return c, float(s[0])


def to_flowsnake(n):
    ex, ey = C.c_int(), C.c_int()
    lib().wbo_to_flowsnake(int(n), C.byref(ex), C.byref(ey))
    return ex.value, ey.value


def base_flow(m):
    ex, ey = C.c_int(), C.c_int()
    lib().wbo_base_flow(int(m), C.byref(ex), C.byref(ey))
    return ex.value, ey.value


def set_exhaustive(on):
    lib().wbo_set_exhaustive(1 if on else 0)


def base_seven(ex, ey):
    return lib().wbo_base_seven(int(ex), int(ey))


def snake_setup(bbmin, bbmax, tile_size=1.0):
    a = np.ascontiguousarray(bbmin, dtype=np.float64)
    b = np.ascontiguousarray(bbmax, dtype=np.float64)
    sn = Snake()
    lib().wbo_snake_setup(_ptr(a, _dp), _ptr(b, _dp), float(tile_size), C.byref(sn))
    return sn


def atan2i_batch(y, x):
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(y), dtype=np.int32)
    lib().wbo_atan2i_batch(_ptr(y, _dp), _ptr(x, _dp), len(y), out.ctypes.data_as(_ip))
    return out


def snake_cyl(sn, ex, ey):
    """Flowsnake::cyl (flowsnake.cpp:263-271): centre x, y and radius of the
    tile at Eisenstein address (ex, ey)."""
    cx, cy, r = C.c_double(), C.c_double(), C.c_double()
    lib().wbo_snake_cyl(C.byref(sn), int(ex), int(ey), C.byref(cx), C.byref(cy), C.byref(r))
    return cx.value, cy.value, r.value


def surround(dirs):
    d = np.ascontiguousarray(dirs, dtype=np.int32).copy()
    return bool(lib().wbo_surround(d.ctypes.data_as(_ip), len(d)))


class Tree:
    """Canonical octree of decoded double coordinates (input numbering)."""

    def __init__(self, x, y, z, center, side):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.y = np.ascontiguousarray(y, dtype=np.float64)
        self.z = np.ascontiguousarray(z, dtype=np.float64)
        c = np.ascontiguousarray(center, dtype=np.float64)
        self.h = lib().wbo_build(_ptr(self.x, _dp), _ptr(self.y, _dp), _ptr(self.z, _dp),
                                 len(self.x), _ptr(c, _dp), float(side))
        self.n = len(self.x)
        if lib().wbo_failed(self.h):
            raise ValueError("more than 537 points in a level-21 cell")

    def __del__(self):
        if getattr(self, "h", None):
            lib().wbo_free(self.h)
            self.h = None

    @property
    def nleaves(self):
        return lib().wbo_nleaves(self.h)

    @property
    def maxdepth(self):
        return lib().wbo_maxdepth(self.h)

    def leaves(self):
        n = self.nleaves
        cen = np.zeros((n, 3))
        side = np.zeros(n)
        depth = np.zeros(n, dtype=np.int32)
        first = np.zeros(n, dtype=np.int64)
        count = np.zeros(n, dtype=np.int64)
        lib().wbo_leaves(self.h, _ptr(cen, _dp), _ptr(side, _dp), depth.ctypes.data_as(_ip),
                         first.ctypes.data_as(_lp), count.ctypes.data_as(_lp))
        return dict(center=cen, side=side, depth=depth, first=first, count=count)

    def order(self):
        out = np.zeros(self.n, dtype=np.int64)
        lib().wbo_order(self.h, out.ctypes.data_as(_lp))
        return out

    def points_in(self, kind, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        cnt = lib().wbo_points_in_flat(self.h, kind, _ptr(p, _dp), None, 0)
        out = np.zeros(max(cnt, 1), dtype=np.int64)
        lib().wbo_points_in_flat(self.h, kind, _ptr(p, _dp), out.ctypes.data_as(_lp), cnt)
        return out[:cnt]

    def count_points_in(self, kind, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        return lib().wbo_points_in_flat(self.h, kind, _ptr(p, _dp), None, 0)

    def classify_points(self, idx, hs, params=None):
        params = params or default_params()
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        hs = np.ascontiguousarray(hs, dtype=np.float64)
        out = np.zeros(len(idx), dtype=np.uint8)
        lib().wbo_classify_points(self.h, idx.ctypes.data_as(_lp), len(idx), _ptr(hs, _dp),
                                  C.byref(params), out.ctypes.data_as(_u8p))
        return out


def run(cloud, params=None, want_xyz=False):
    """Full single-threaded pass (build, scan, postscan, classify).

    cloud: dict with int32 arrays xi, yi, zi and scale, offset, bbmin, bbmax.
    Returns (class bytes, tiles structured array, stats, [x, y, z])."""
    params = params or default_params()
    xi = np.ascontiguousarray(cloud["xi"], dtype=np.int32)
    yi = np.ascontiguousarray(cloud["yi"], dtype=np.int32)
    zi = np.ascontiguousarray(cloud["zi"], dtype=np.int32)
    n = len(xi)
    sc = np.ascontiguousarray(cloud["scale"], dtype=np.float64)
    of = np.ascontiguousarray(cloud["offset"], dtype=np.float64)
    bmin = np.ascontiguousarray(cloud["bbmin"], dtype=np.float64)
    bmax = np.ascontiguousarray(cloud["bbmax"], dtype=np.float64)
    sn = snake_setup(bmin, bmax, params.tile_size)
    ntiles = sn.hi - sn.lo + 1
    tiles = (Tile * ntiles)()
    st = Stats()
    cls = np.zeros(n, dtype=np.uint8)
    x = np.zeros(n) if want_xyz else None
    y = np.zeros(n) if want_xyz else None
    z = np.zeros(n) if want_xyz else None
    rc = lib().wbo_run(xi.ctypes.data_as(_i32p), yi.ctypes.data_as(_i32p), zi.ctypes.data_as(_i32p),
                  n, _ptr(sc, _dp), _ptr(of, _dp), _ptr(bmin, _dp), _ptr(bmax, _dp),
                  C.byref(params), cls.ctypes.data_as(_u8p), tiles, C.byref(st),
                  _ptr(x, _dp) if want_xyz else None, _ptr(y, _dp) if want_xyz else None,
                  _ptr(z, _dp) if want_xyz else None)
    if rc < 0:
        raise ValueError("more than 537 points in a level-21 cell")
    tarr = np.ctypeslib.as_array(tiles)
    out = (cls, tarr.copy(), st)
    if want_xyz:
        out = out + (x, y, z)
    return out
