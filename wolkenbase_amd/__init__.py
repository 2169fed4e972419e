"""MI355X octree build and ground classification (wolkenbase's hot path).

Python mirror of the reference's Octree / Shape / classify interface over
the C ABI of include/wb_gpu.h (libwb_gpu.so, built in-tree for gfx950).
There is no CPU fallback: every call runs the HIP kernels, and importing a
GPU entry point without the built library raises.

    tree = Octree.build(ctx, xi, yi, zi, scale, offset, *Octree.size_fit(bbmin, bbmax))
    counts = tree.count_points_in(Cylinder(cx, cy, r), ...)
    cls, tiles, stats = classify(tree, bbmin, bbmax)
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GPU_SO = os.path.join(HERE, "libwb_gpu.so")
# another build of the library (e.g. one compiled with different settings,
# to compare two builds on one device); the in-tree build by default
GPU_SO = os.environ.get("WB_GPU_LIB", GPU_SO)

WB_SHAPE_CYLINDER, WB_SHAPE_SPHERE, WB_SHAPE_HYPERBOLOID = 0, 1, 2
WB_SHAPE_NPARAMS = 5
WB_NOALLOC_MAX_SHAPES = 65536
WB_DEVICE_INPUT = 1
WB_BUILD_DEDUPE = 2
WB_WRITE_DRY = 4
WB_ZEROS_READ, WB_ZEROS_KEEP = 0, 1
WB_FIT_FP64, WB_FIT_FP32, WB_FIT_FP32_COMP = 0, 1, 2
STORE_RECORD, STORE_RECORDS, STORE_BLOCK = 61, 537, 32768  # octree.h:35-43


def _class_arg(cls, n):
    """A host class-byte array of n entries as a ctypes pointer (None
    passes through); the array must outlive the call."""
    if cls is None:
        return None
    a = np.ascontiguousarray(cls, dtype=np.uint8)
    if a.size != n:
        raise ValueError(f"class bytes: {a.size} for {n} input points")
    _class_arg.keep = a
    return a.ctypes.data_as(_vp)


class WbError(RuntimeError):
    pass


class LasXform(C.Structure):
    _fields_ = [("scale", C.c_double * 3), ("offset", C.c_double * 3), ("unit", C.c_double)]


class Leaf(C.Structure):
    _fields_ = [("center", C.c_double * 3), ("side", C.c_double), ("first", C.c_uint64),
                ("count", C.c_uint32), ("depth", C.c_int32)]


class ClassifyParams(C.Structure):
    _fields_ = [("tile_size", C.c_double), ("max_slope", C.c_double),
                ("thickness", C.c_double), ("min_hs", C.c_double),
                ("fit_precision", C.c_int32)]


class TileRecord(C.Structure):
    _fields_ = [("npoints", C.c_int32), ("nground", C.c_int32), ("tree_flags", C.c_int32),
                ("flowsnake_x", C.c_int32), ("flowsnake_y", C.c_int32), ("fit_fp64", C.c_int32),
                ("density", C.c_double), ("hs_scan", C.c_double), ("hs", C.c_double),
                ("height", C.c_double), ("tilt_x", C.c_double), ("tilt_y", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("n_class1", C.c_uint64), ("n_class2", C.c_uint64),
                ("n_unclassified", C.c_uint64), ("ntiles", C.c_uint64),
                ("nonempty_tiles", C.c_uint64), ("sector_borderline", C.c_uint64),
                ("exact_gap_points", C.c_uint64), ("atan_exact_calls", C.c_uint64),
                ("heavy_points", C.c_uint64),
                ("sort_passes", C.c_uint32), ("maxdepth", C.c_uint32),
                ("nleaves", C.c_uint64), ("nnodes", C.c_uint64),
                ("ms_build", C.c_double), ("ms_sort", C.c_double), ("ms_tiles", C.c_double),
                ("ms_scan", C.c_double), ("ms_classify", C.c_double),
                ("ms_classify_kernel", C.c_double), ("ms_heavy_kernel", C.c_double),
                ("ms_total", C.c_double), ("fit_fallbacks", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LasHeader(C.Structure):
    """wb_las_header: a LAS header after LasHeader::openRead (las.cpp:272-428)."""
    _fields_ = [("source_id", C.c_uint16), ("global_encoding", C.c_uint16),
                ("guid", C.c_uint8 * 16), ("version_major", C.c_uint8),
                ("version_minor", C.c_uint8), ("system_id", C.c_char * 33),
                ("software", C.c_char * 33), ("creation_day", C.c_uint16),
                ("creation_year", C.c_uint16), ("header_size", C.c_uint32),
                ("point_offset", C.c_uint32), ("n_vlr", C.c_uint32),
                ("point_format", C.c_uint16), ("point_length", C.c_uint16),
                ("scale", C.c_double * 3), ("offset", C.c_double * 3),
                ("min", C.c_double * 3), ("max", C.c_double * 3),
                ("start_waveform", C.c_uint64), ("start_evlr", C.c_uint64),
                ("n_evlr", C.c_uint32), ("n_points", C.c_uint64 * 16),
                ("zipped", C.c_int32), ("valid", C.c_int32)]


class LasWriteParams(C.Structure):
    _fields_ = [("name", C.c_char_p), ("separate_classes", C.c_int32),
                ("points_per_file", C.c_int32), ("point_format", C.c_int32),
                ("creation_day", C.c_int32), ("creation_year", C.c_int32)]


class LasWriteStats(C.Structure):
    _fields_ = [("class_totals", C.c_uint64 * 256), ("npoints", C.c_uint64),
                ("bytes", C.c_uint64), ("nfiles", C.c_uint32), ("point_format", C.c_int32),
                ("scale", C.c_double * 3), ("offset", C.c_double * 3),
                ("ms_classes", C.c_double), ("ms_plan", C.c_double),
                ("ms_encode", C.c_double), ("ms_copy", C.c_double), ("ms_disk", C.c_double)]


def default_params(tile_size=1.0, max_slope=1.0, thickness=0.0, min_hs=0.1,
                   fit_precision=0):
    """Reference defaults (mainwindow.cpp:403-410); fit_precision WB_FIT_FP64."""
    return ClassifyParams(tile_size, max_slope, thickness, min_hs, fit_precision)


_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)

# exported symbols of include/wb_gpu.h with their ctypes signatures
SIGNATURES = {
    "wb_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),

    "wb_ctx_destroy": (None, [_vp]),
    "wb_ctx_set_sharing": (C.c_int, [_vp, C.c_int]),
    "wb_host_alloc": (C.c_int, [C.c_uint64, C.POINTER(_vp)]),
    "wb_host_free": (None, [_vp]),
    "wb_device_memory": (C.c_int, [C.c_int, _u64p, _u64p]),
    "wb_last_error": (C.c_char_p, [_vp]),
    "wb_ctx_sync": (C.c_int, [_vp]),
    "wb_size_fit": (C.c_int, [_dp, C.c_int, _dp, _dp]),
    "wb_tree_build": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint64, C.POINTER(LasXform), _dp,
                                C.c_double, C.c_int, C.POINTER(_vp)]),
    "wb_tree_build_xyz": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint64, _dp, _dp, C.c_double,
                                    C.c_int, C.POINTER(_vp)]),
    "wb_tree_duplicates": (C.c_int, [_vp, _u64p, _u32p]),
    "wb_tree_free": (C.c_int, [_vp]),
    "wb_tree_info": (C.c_int, [_vp, _u64p, _u64p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "wb_tree_query_ms": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "wb_tree_leaves": (C.c_int, [_vp, C.POINTER(Leaf), C.c_uint64, _u64p]),
    "wb_tree_order": (C.c_int, [_vp, _u32p]),
    "wb_count_points_in": (C.c_int, [_vp, C.c_int, _dp, C.c_uint64, _u64p]),
    "wb_count_points_in_noalloc": (C.c_int, [_vp, C.c_int, _dp, C.c_uint64, _u64p]),
    "wb_points_in": (C.c_int, [_vp, C.c_int, _dp, C.c_uint64, _u64p, _u32p, C.c_uint64]),
    "wb_classify_batch": (C.c_int, [C.POINTER(_vp), C.c_int, _dp, _dp,
                                    C.POINTER(ClassifyParams), C.POINTER(_vp), C.c_int,
                                    C.POINTER(Stats)]),
    "wb_classify": (C.c_int, [_vp, _dp, _dp, C.POINTER(ClassifyParams), _vp, C.c_int,
                              C.POINTER(TileRecord), C.c_uint64, C.POINTER(Stats)]),
    "wb_tile_count": (C.c_int, [_dp, _dp, C.c_double, _u64p]),
    "wb_scan_tiles": (C.c_int, [_vp, _dp, _dp, C.POINTER(ClassifyParams),
                                C.POINTER(TileRecord), C.c_uint64, C.POINTER(Stats)]),
    "wb_classify_corners": (C.c_int, [_vp, _dp, C.c_int, C.POINTER(ClassifyParams), _vp,
                                      C.c_int, C.POINTER(TileRecord), C.c_uint64,
                                      C.POINTER(Stats)]),
    "wb_tile_count_corners": (C.c_int, [_dp, C.c_int, C.c_double, _u64p]),
    "wb_tree_dump": (C.c_int, [_vp, C.c_char_p, C.c_char_p, C.c_uint64, _u64p]),
    "wb_ldecimal": (C.c_int, [C.c_double, C.c_double, C.c_char_p, C.c_int]),
    "wb_las_read_header": (C.c_int, [C.c_char_p, C.POINTER(LasHeader)]),
    "wb_las_open": (C.c_int, [_vp, C.c_char_p, C.POINTER(_vp), C.POINTER(LasHeader)]),
    "wb_las_open_mem": (C.c_int, [_vp, _vp, C.c_uint64, C.POINTER(_vp),
                                  C.POINTER(LasHeader)]),
    "wb_las_close": (C.c_int, [_vp]),
    "wb_cloud_create": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.c_double, C.c_int,
                                  C.POINTER(_vp)]),
    "wb_cloud_info": (C.c_int, [_vp, _u64p, _dp, C.POINTER(C.c_int), _dp]),
    "wb_cloud_points": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "wb_cloud_read_points": (C.c_int, [_vp, _dp, _dp, _dp, _u32p]),
    "wb_cloud_decode_ms": (C.c_int, [_vp, _dp]),
    "wb_cloud_build": (C.c_int, [_vp, _dp, C.c_double, C.c_int, C.POINTER(_vp)]),
    "wb_cloud_classify": (C.c_int, [_vp, _vp, C.POINTER(ClassifyParams), _vp, C.c_int,
                                    C.POINTER(TileRecord), C.c_uint64, C.POINTER(Stats)]),
    "wb_cloud_free": (C.c_int, [_vp]),
    "wb_store_blocks": (C.c_int, [_vp, _vp, _vp, C.c_int, _vp, C.c_uint64, _u64p]),
    "wb_store_write_files": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_char_p, C.c_int]),
    "wb_store_read_blocks": (C.c_int, [_vp, _vp, C.c_uint64, _dp, _dp, _dp,
                                       C.POINTER(C.c_uint16), _u32p, _u64p]),
    "wb_cloud_write": (C.c_int, [_vp, _vp, _vp, C.c_int, C.POINTER(LasWriteParams),
                                 C.POINTER(LasWriteStats), C.c_char_p, C.c_uint64]),
}

_lib = None


def lib():
    """The HIP library; raises if it was not built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(GPU_SO):
            raise WbError("libwb_gpu.so is not built (run __graft_entry__.build())")
        L = C.CDLL(GPU_SO)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _arr(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _dptr(a):
    return a.ctypes.data_as(_dp)


def _on_device(a):
    """True for a tensor in device memory; a host tensor (pinned or not) is
    host memory, passed by its address without WB_DEVICE_INPUT."""
    return bool(getattr(a, "is_cuda", False))


def _host_arrays(arrs, dtype):
    """Host coordinate arrays as (keep-alive objects, addresses, length):
    numpy arrays are made contiguous in `dtype`; host tensors (e.g. pinned
    receive buffers) are passed as they are and must already be contiguous
    arrays of that type."""
    if all(hasattr(a, "data_ptr") for a in arrs):
        isz = np.dtype(dtype).itemsize
        for a in arrs:
            if not a.is_contiguous() or a.element_size() != isz or a.is_floating_point() != (
                    np.dtype(dtype).kind == "f"):
                raise ValueError(f"host tensor must be a contiguous {np.dtype(dtype)} array")
        return arrs, [a.data_ptr() for a in arrs], int(arrs[0].numel())
    arrs = [_arr(a, dtype) for a in arrs]
    return arrs, [a.ctypes.data for a in arrs], len(arrs[0])


class _PinnedBlock:
    """Owner of one wb_host_alloc block: freed when the last array viewing
    it goes."""

    def __init__(self, ptr):
        self.ptr = ptr

    def __del__(self):
        if self.ptr:
            lib().wb_host_free(_vp(self.ptr))
            self.ptr = None


def pinned_empty(n, dtype):
    """A numpy array of n elements in pinned host memory (hipHostMalloc):
    copies between it and the device run by DMA without staging."""
    dtype = np.dtype(dtype)
    nbytes = max(1, int(n) * dtype.itemsize)
    p = _vp()
    if lib().wb_host_alloc(nbytes, C.byref(p)) != 0:
        raise WbError(f"wb_host_alloc({nbytes}) failed")
    buf = (C.c_uint8 * nbytes).from_address(p.value)
    buf.owner = _PinnedBlock(p.value)
    return np.frombuffer(buf, dtype=dtype, count=int(n))


class Context:
    """wb_ctx: one HIP device + stream + device memory pool.  sharing: the
    long-running kernels leave room on every CU for other contexts' kernels
    (wb_ctx_set_sharing); results are the same."""

    def __init__(self, device=0, sharing=False):
        h = _vp()
        rc = lib().wb_ctx_create(device, C.byref(h))
        if rc != 0:
            raise WbError(f"wb_ctx_create({device}) failed with {rc}")
        self.h = h
        if sharing and lib().wb_ctx_set_sharing(h, 1) != 0:
            lib().wb_ctx_destroy(h)
            self.h = None
            raise WbError("wb_ctx_set_sharing failed")
        self.trees = weakref.WeakSet()
        self.las = weakref.WeakSet()  # LasFile and Cloud objects

    def check(self, rc, what):
        if rc != 0:
            raise WbError(f"{what}: {lib().wb_last_error(self.h).decode()} ({rc})")

    def sync(self):
        self.check(lib().wb_ctx_sync(self.h), "wb_ctx_sync")

    def store_read_blocks(self, blocks):
        """OctBuffer::read (octree.cpp:508-559) of store blocks (bytes or a
        uint8 array of 32768-byte blocks): (x, y, z, classification,
        block_count) of the records with a location, block by block."""
        buf = np.ascontiguousarray(np.frombuffer(blocks, dtype=np.uint8)
                                   if not isinstance(blocks, np.ndarray) else blocks,
                                   dtype=np.uint8)
        if buf.size % STORE_BLOCK:
            raise ValueError("not a whole number of 32768-byte blocks")
        nb = buf.size // STORE_BLOCK
        L = lib()
        counts = np.zeros(max(nb, 1), dtype=np.uint32)
        n = C.c_uint64()
        self.check(L.wb_store_read_blocks(self.h, buf.ctypes.data_as(_vp), nb, None, None, None,
                                          None, counts.ctypes.data_as(_u32p), C.byref(n)),
                   "wb_store_read_blocks")
        m = max(n.value, 1)
        x, y, z = np.zeros(m), np.zeros(m), np.zeros(m)
        cls = np.zeros(m, dtype=np.uint16)
        self.check(L.wb_store_read_blocks(self.h, buf.ctypes.data_as(_vp), nb,
                                          x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                                          z.ctypes.data_as(_dp),
                                          cls.ctypes.data_as(C.POINTER(C.c_uint16)),
                                          counts.ctypes.data_as(_u32p), C.byref(n)),
                   "wb_store_read_blocks")
        k = n.value
        return x[:k], y[:k], z[:k], cls[:k], counts[:nb]

    def close(self):
        """Frees the clouds, LAS files and trees still alive on this context,
        then the context."""
        if getattr(self, "h", None):
            objs = list(getattr(self, "las", ()))
            for o in sorted(objs, key=lambda o: not hasattr(o, "files")):
                o.close()
            for t in list(getattr(self, "trees", ())):
                t.free()
            lib().wb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- shapes (shape.h) -------------------------------------------------------
def Cylinder(cx, cy, r):
    """Cylinder(xy c, double r) -- shape.h:136-157."""
    return (WB_SHAPE_CYLINDER, (cx, cy, r, 0.0, 0.0))


def Sphere(cx, cy, cz, r):
    """Sphere(xyz c, double r) -- shape.h:120-134."""
    return (WB_SHAPE_SPHERE, (cx, cy, cz, r, 0.0))


def Hyperboloid(vx, vy, vz, r, slope):
    """Hyperboloid(xyz vertex, double radius of curvature, double slope) -- shape.h:96-118."""
    return (WB_SHAPE_HYPERBOLOID, (vx, vy, vz, r, slope))


def _shape_batch(shapes):
    """shapes: list of shape tuples of one kind, or (kind, params array Nx5)."""
    if isinstance(shapes, tuple) and len(shapes) == 2 and not isinstance(shapes[0], tuple):
        kind, params = shapes
        params = np.atleast_2d(_arr(params, np.float64))
        if params.shape[1] < WB_SHAPE_NPARAMS:
            params = np.pad(params, ((0, 0), (0, WB_SHAPE_NPARAMS - params.shape[1])))
        return kind, _arr(params, np.float64)
    kinds = {s[0] for s in shapes}
    if len(kinds) != 1:
        raise WbError("one shape kind per batch")
    return kinds.pop(), _arr([s[1] for s in shapes], np.float64)


class Octree:
    """Canonical octree resident in HBM (Octree + OctStore of octree.h)."""

    @staticmethod
    def size_fit_corners(corners):
        """Octree::sizeFit over a list of corners (octree.cpp:268-310)."""
        pts = _arr(np.asarray(corners).reshape(-1), np.float64)
        c = np.zeros(3)
        s = np.zeros(1)
        rc = lib().wb_size_fit(_dptr(pts), len(pts) // 3, _dptr(c), _dptr(s))
        if rc != 0:
            raise WbError("wb_size_fit failed")
        return c, float(s[0])

    @staticmethod
    def size_fit(bbmin, bbmax):
        """Octree::sizeFit over the two header corners (octree.cpp:268-310)."""
        pts = _arr(np.concatenate([bbmin, bbmax]), np.float64)
        c = np.zeros(3)
        s = np.zeros(1)
        rc = lib().wb_size_fit(_dptr(pts), 2, _dptr(c), _dptr(s))
        if rc != 0:
            raise WbError("wb_size_fit failed")
        return c, float(s[0])

    @classmethod
    def build(cls, ctx, xi, yi, zi, scale, offset, center, side, unit=1.0, sort_passes=None,
              dedupe=False):
        """Build from int32 LAS coordinates (numpy arrays, or device tensors
        exposing data_ptr() for input already resident in HBM).  dedupe merges
        points at one location as OctBuffer::put does (octree.cpp:620-660)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        xf = LasXform((C.c_double * 3)(*scale), (C.c_double * 3)(*offset), unit)
        cen = _arr(center, np.float64)
        dev = _on_device(xi)
        if dev:
            n = int(xi.numel())
            px, py, pz = xi.data_ptr(), yi.data_ptr(), zi.data_ptr()
        else:
            keep, (px, py, pz), n = _host_arrays((xi, yi, zi), np.int32)
        h = _vp()
        flags = (WB_DEVICE_INPUT if dev else 0) | (WB_BUILD_DEDUPE if dedupe else 0)
        rc = lib().wb_tree_build(ctx.h, px, py, pz, n, C.byref(xf), _dptr(cen), float(side),
                                 flags, C.byref(h))
        ctx.check(rc, "wb_tree_build")
        return self._adopt(ctx, h, n, cen, side)

    @classmethod
    def build_xyz(cls, ctx, x, y, z, resolution, center, side, dedupe=False):
        """Build from coordinates already decoded to FP64 (numpy arrays or
        device tensors).  resolution: per-axis spacing every nonzero
        coordinate difference is at least."""
        self = cls.__new__(cls)
        self.ctx = ctx
        cen = _arr(center, np.float64)
        res = _arr(resolution, np.float64)
        dev = _on_device(x)
        if dev:
            n = int(x.numel())
            px, py, pz = x.data_ptr(), y.data_ptr(), z.data_ptr()
        else:
            keep, (px, py, pz), n = _host_arrays((x, y, z), np.float64)
        h = _vp()
        flags = (WB_DEVICE_INPUT if dev else 0) | (WB_BUILD_DEDUPE if dedupe else 0)
        rc = lib().wb_tree_build_xyz(ctx.h, px, py, pz, n, _dptr(res), _dptr(cen), float(side),
                                     flags, C.byref(h))
        ctx.check(rc, "wb_tree_build_xyz")
        return self._adopt(ctx, h, n, cen, side)

    def _adopt(self, ctx, h, n, cen, side):
        self.h = h
        ctx.trees.add(self)
        nd = C.c_uint64()
        ctx.check(lib().wb_tree_duplicates(h, C.byref(nd), None), "wb_tree_duplicates")
        self.n_input = n
        self.ndup = nd.value
        self.n = n - nd.value
        self.center = cen
        self.side = float(side)
        return self

    def dump(self, path=None):
        """OctStore::dump (octree.cpp:360-376, 673-689, 888-891) as text;
        also written to path when given."""
        n = C.c_uint64()
        L = lib()
        self.ctx.check(L.wb_tree_dump(self.h, None, None, 0, C.byref(n)), "wb_tree_dump")
        buf = C.create_string_buffer(n.value + 1)
        self.ctx.check(L.wb_tree_dump(self.h, path.encode() if path else None, buf, n.value + 1,
                                      C.byref(n)), "wb_tree_dump")
        return buf.raw[:n.value].decode("utf-8")

    def store_blocks(self, cloud=None, cls=None, out=None):
        """The store's blocks (OctBuffer::write, octree.cpp:427-460) of every
        leaf in DFS order: nleaves * 32768 bytes into out (a numpy uint8
        array or a host tensor, e.g. pinned; allocated when None), returned.
        cloud: the Cloud the tree was built from (its records' fields);
        cls: class bytes in input order, host array or device tensor (a
        nonzero byte replaces the classification)."""
        L = lib()
        nb = C.c_uint64()
        ch = cloud.h if cloud is not None else None
        flags = 0
        if cls is not None and _on_device(cls):
            if cls.numel() != self.n_input:
                raise ValueError(f"class bytes: {cls.numel()} for {self.n_input} input points")
            cp, flags = _vp(cls.data_ptr()), WB_DEVICE_INPUT
        else:
            cp = _class_arg(cls, self.n_input)
        self.ctx.check(L.wb_store_blocks(self.h, ch, cp, flags, None, 0, C.byref(nb)),
                       "wb_store_blocks")
        need = nb.value * STORE_BLOCK
        if out is None:
            out = np.zeros(need, dtype=np.uint8)
        if hasattr(out, "data_ptr"):
            ptr, cap = out.data_ptr(), out.numel() * out.element_size()
        else:
            ptr, cap = out.ctypes.data, out.nbytes
        self.ctx.check(L.wb_store_blocks(self.h, ch, cp, flags, _vp(ptr), cap, C.byref(nb)),
                       "wb_store_blocks")
        return out

    def store_write_files(self, name, nfiles=1, cloud=None, cls=None):
        """The blocks as the store's files name0 .. name<nfiles-1>
        (OctStore::open, octree.cpp:812-830): block b in file b % nfiles at
        offset 32768 * (b // nfiles)."""
        ch = cloud.h if cloud is not None else None
        cp = _class_arg(cls, self.n_input)
        self.ctx.check(lib().wb_store_write_files(self.h, ch, cp, 0, name.encode(), int(nfiles)),
                       "wb_store_write_files")

    def duplicates(self):
        """(merged input index, input index of the first point at its
        location) pairs, shape (ndup, 2)."""
        out = np.zeros((max(self.ndup, 1), 2), dtype=np.uint32)
        nd = C.c_uint64()
        self.ctx.check(lib().wb_tree_duplicates(self.h, C.byref(nd), out.ctypes.data_as(_u32p)),
                       "wb_tree_duplicates")
        return out[:self.ndup]

    def free(self):
        # a closed context has already freed its trees
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().wb_tree_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def info(self):
        nl, nn = C.c_uint64(), C.c_uint64()
        md, sp = C.c_int(), C.c_int()
        self.ctx.check(lib().wb_tree_info(self.h, C.byref(nl), C.byref(nn), C.byref(md),
                                          C.byref(sp)), "wb_tree_info")
        return dict(nleaves=nl.value, nnodes=nn.value, maxdepth=md.value, sort_passes=sp.value)

    def leaves(self):
        """Leaves in DFS order: dict of arrays center (k,3), side, first, count, depth."""
        nl = self.info()["nleaves"]
        buf = (Leaf * max(nl, 1))()
        got = C.c_uint64()
        self.ctx.check(lib().wb_tree_leaves(self.h, buf, nl, C.byref(got)), "wb_tree_leaves")
        a = np.ctypeslib.as_array(buf)[:nl]
        return dict(center=np.array(a["center"], dtype=np.float64).reshape(nl, 3),
                    side=a["side"].copy(), first=a["first"].astype(np.int64),
                    count=a["count"].astype(np.int64), depth=a["depth"].astype(np.int32))

    def order(self):
        """Input indices in canonical order (leaf DFS order, input order in a leaf)."""
        out = np.zeros(self.n, dtype=np.uint32)
        self.ctx.check(lib().wb_tree_order(self.h, out.ctypes.data_as(_u32p)), "wb_tree_order")
        return out

    def count_points_in(self, shapes):
        """OctStore::countPointsIn for a batch of shapes (octree.cpp:1244-1264)."""
        kind, p = _shape_batch(shapes)
        out = np.zeros(len(p), dtype=np.uint64)
        self.ctx.check(lib().wb_count_points_in(self.h, kind, _dptr(p), len(p),
                                                out.ctypes.data_as(_u64p)), "wb_count_points_in")
        return out

    def count_points_in_noalloc(self, shapes):
        """count_points_in without device allocation: the shapes and counts
        are staged in the library's fixed per-GPU buffer (at most
        WB_NOALLOC_MAX_SHAPES shapes; calls on one GPU must not overlap)."""
        kind, p = _shape_batch(shapes)
        out = np.zeros(len(p), dtype=np.uint64)
        self.ctx.check(lib().wb_count_points_in_noalloc(self.h, kind, _dptr(p), len(p),
                                                        out.ctypes.data_as(_u64p)),
                       "wb_count_points_in_noalloc")
        return out

    def query_ms(self):
        """Device time (ms) of the kernels of the last count_points_in /
        points_in call on this tree: (counting kernel, index kernel)."""
        a, b = C.c_double(0), C.c_double(0)
        self.ctx.check(lib().wb_tree_query_ms(self.h, C.byref(a), C.byref(b)),
                       "wb_tree_query_ms")
        return a.value, b.value

    def points_in(self, shapes):
        """OctStore::pointsIn (unsorted) for a batch: list of input-index arrays
        in the reference's order (octree.cpp:1214-1242)."""
        kind, p = _shape_batch(shapes)
        ns = len(p)
        offs = np.zeros(ns + 1, dtype=np.uint64)
        L = lib()
        self.ctx.check(L.wb_points_in(self.h, kind, _dptr(p), ns, offs.ctypes.data_as(_u64p),
                                      None, 0), "wb_points_in")
        total = int(offs[-1])
        idx = np.zeros(max(total, 1), dtype=np.uint32)
        self.ctx.check(L.wb_points_in(self.h, kind, _dptr(p), ns, offs.ctypes.data_as(_u64p),
                                      idx.ctypes.data_as(_u32p), total), "wb_points_in")
        return [idx[int(offs[s]):int(offs[s + 1])].copy() for s in range(ns)]


def ldecimal(x, toler=0.0):
    """ldecimal (ldecimal.cpp:32-132)."""
    buf = C.create_string_buffer(64)
    n = lib().wb_ldecimal(float(x), float(toler), buf, 64)
    if n < 0:
        raise WbError("wb_ldecimal failed")
    return buf.value.decode()


def tile_count_corners(corners, tile_size=1.0):
    a = _arr(np.asarray(corners).reshape(-1), np.float64)
    n = C.c_uint64()
    if lib().wb_tile_count_corners(_dptr(a), len(a) // 3, float(tile_size), C.byref(n)) != 0:
        raise WbError("wb_tile_count_corners failed")
    return n.value


def device_memory(device=0):
    """(free, total) bytes of device memory."""
    f, t = C.c_uint64(), C.c_uint64()
    if lib().wb_device_memory(device, C.byref(f), C.byref(t)) != 0:
        raise WbError("wb_device_memory failed")
    return f.value, t.value


def tile_count(bbmin, bbmax, tile_size=1.0):
    a, b = _arr(bbmin, np.float64), _arr(bbmax, np.float64)
    n = C.c_uint64()
    if lib().wb_tile_count(_dptr(a), _dptr(b), float(tile_size), C.byref(n)) != 0:
        raise WbError("wb_tile_count failed")
    return n.value


def classify(tree, bbmin, bbmax, params=None, want_tiles=True, out=None):
    """SCAN + POSTSCAN + CLASSIFY of every flowsnake tile (scan.cpp:31-179,
    classify.cpp:96-173).  Returns (class bytes in input order, tile records
    or None, stats dict).  If `out` is a device tensor (uint8, n), the classes
    stay in HBM there; if it is a host buffer of n bytes (a numpy uint8 array
    or a host tensor, e.g. pinned), they are copied there; either way the
    first return value is `out`."""
    params = params or default_params()
    a, b = _arr(bbmin, np.float64), _arr(bbmax, np.float64)
    ntiles = tile_count(a, b, params.tile_size) if want_tiles else 0
    tiles = (TileRecord * max(ntiles, 1))() if want_tiles else None
    st = Stats()
    if out is not None and _on_device(out):
        cls = out
        ptr, flags = out.data_ptr(), WB_DEVICE_INPUT
    elif out is not None:
        # a host buffer (numpy or host tensor, e.g. pinned) of n bytes
        cls = out
        if hasattr(out, "data_ptr"):
            if out.numel() * out.element_size() != tree.n_input or not out.is_contiguous():
                raise ValueError("class output: a contiguous buffer of n bytes")
            ptr = out.data_ptr()
        else:
            if out.dtype != np.uint8 or out.size != tree.n_input or not out.flags.c_contiguous:
                raise ValueError("class output: a contiguous uint8 array of n bytes")
            ptr = out.ctypes.data
        flags = 0
    else:
        cls = np.zeros(tree.n_input, dtype=np.uint8)
        ptr, flags = cls.ctypes.data, 0
    rc = lib().wb_classify(tree.h, _dptr(a), _dptr(b), C.byref(params), ptr, flags,
                           tiles if want_tiles else None, ntiles, C.byref(st))
    tree.ctx.check(rc, "wb_classify")
    tarr = np.ctypeslib.as_array(tiles)[:ntiles].copy() if want_tiles else None
    return cls, tarr, st.as_dict()


def classify_batch(trees, bbmins, bbmaxs, params=None, outs=None):
    """classify() of several trees of one context in one pass, their ground
    tests in one kernel launch.  bbmins, bbmaxs: one header box per tree.
    outs: None (host class bytes are returned), one device tensor (uint8, n)
    per tree, or one host buffer of n bytes per tree (numpy or host tensor,
    e.g. pinned).  Returns (list of class bytes, list of stats dicts)."""
    params = params or default_params()
    k = len(trees)
    if k == 0:
        return [], []
    a = _arr(np.asarray(bbmins, dtype=np.float64).reshape(k, 3), np.float64)
    b = _arr(np.asarray(bbmaxs, dtype=np.float64).reshape(k, 3), np.float64)
    handles = (_vp * k)(*[t.h for t in trees])
    if outs is not None:
        cls = list(outs)
        if len(cls) != k:
            raise ValueError(f"{len(cls)} class outputs for {k} trees")
        dev = [_on_device(o) for o in cls]
        if any(dev) != all(dev):
            raise ValueError("class outputs: all device tensors or all host buffers")
        for o, t in zip(cls, trees):
            nb = o.numel() * o.element_size() if hasattr(o, "data_ptr") else o.nbytes
            if nb != t.n_input:
                raise ValueError(f"class output of {nb} bytes for {t.n_input} points")
        ptrs = (_vp * k)(*[o.data_ptr() if hasattr(o, "data_ptr") else o.ctypes.data
                           for o in cls])
        flags = WB_DEVICE_INPUT if all(dev) else 0
    else:
        cls = [np.zeros(t.n_input, dtype=np.uint8) for t in trees]
        ptrs = (_vp * k)(*[c.ctypes.data for c in cls])
        flags = 0
    st = (Stats * k)()
    rc = lib().wb_classify_batch(handles, k, _dptr(a), _dptr(b), C.byref(params), ptrs, flags, st)
    trees[0].ctx.check(rc, "wb_classify_batch")
    return cls, [st[j].as_dict() for j in range(k)]


def scan_tiles(tree, bbmin, bbmax, params=None):
    """SCAN + POSTSCAN of every flowsnake tile (scanCylinder and
    postscanCylinder, scan.cpp:31-179) without the classification.  Returns
    (tile records, stats dict)."""
    params = params or default_params()
    a, b = _arr(bbmin, np.float64), _arr(bbmax, np.float64)
    ntiles = tile_count(a, b, params.tile_size)
    tiles = (TileRecord * max(ntiles, 1))()
    st = Stats()
    rc = lib().wb_scan_tiles(tree.h, _dptr(a), _dptr(b), C.byref(params), tiles, ntiles,
                             C.byref(st))
    tree.ctx.check(rc, "wb_scan_tiles")
    return np.ctypeslib.as_array(tiles)[:ntiles].copy(), st.as_dict()
