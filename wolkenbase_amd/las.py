"""LAS files in and out of HBM: the reference's LasHeader (las.h:97-183),
the READ phase over several files (threads.cpp:479-545, wolkencli.cpp:
99-111) and CloudOutput (cloudoutput.h:29-47), over the C ABI of
include/wb_gpu.h.  No CPU fallback: without libwb_gpu.so every call raises."""
import ctypes as C

import numpy as np

from . import (LasHeader, LasWriteParams, LasWriteStats, Octree, Stats, TileRecord, WbError,
               WB_DEVICE_INPUT, WB_WRITE_DRY, WB_ZEROS_KEEP, WB_ZEROS_READ, _vp, default_params,
               lib)

ZEROS = {"read": WB_ZEROS_READ, "keep": WB_ZEROS_KEEP}


def header_dict(h):
    d = {}
    for f, _ in LasHeader._fields_:
        v = getattr(h, f)
        if isinstance(v, bytes):
            v = v.decode("latin-1")
        elif hasattr(v, "_length_"):
            v = list(v)
        d[f] = v
    return d


def read_header(path):
    """The header of a LAS file as LasHeader::openRead leaves it (host only)."""
    h = LasHeader()
    if lib().wb_las_read_header(str(path).encode(), C.byref(h)) != 0:
        raise WbError(f"{path}: not a LAS file")
    return header_dict(h)


class LasFile:
    """A LAS file opened for reading, its point records resident in HBM
    (LasHeader::openRead + readPoint for every point)."""

    def __init__(self, ctx, path=None, image=None):
        self.ctx = ctx
        h = LasHeader()
        p = _vp()
        if image is not None:
            self._image = bytes(image)
            rc = lib().wb_las_open_mem(ctx.h, self._image, len(self._image), C.byref(p),
                                       C.byref(h))
            self.name = "<memory>"
        else:
            rc = lib().wb_las_open(ctx.h, str(path).encode(), C.byref(p), C.byref(h))
            self.name = str(path)
        ctx.check(rc, "wb_las_open")
        self.h = p
        ctx.las.add(self)
        self.header = header_dict(h)

    # LasHeader accessors (las.h:143-183)
    def number_points(self, r=0):
        return self.header["n_points"][r]

    def get_scale(self, unit=1.0):
        return np.array(self.header["scale"]) * unit

    def get_offset(self, unit=1.0):
        return np.array(self.header["offset"]) * unit

    def min_corner(self, unit=1.0):
        return np.array(self.header["min"]) * unit

    def max_corner(self, unit=1.0):
        return np.array(self.header["max"]) * unit

    def get_version(self):
        return (self.header["version_major"] << 8) + self.header["version_minor"]

    def get_point_format(self):
        return self.header["point_format"]

    def is_valid(self):
        return bool(self.header["valid"])

    def is_zipped(self):
        return bool(self.header["zipped"])

    def close(self):
        # a closed context has already freed its files
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().wb_las_close(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cloud:
    """The points of one or more LasFiles decoded to FP64 in HBM, in file
    order, after the READ phase's rule for return number 0 ("read") or
    wolkencli's ("keep": every record)."""

    def __init__(self, ctx, files, unit=1.0, zeros="read"):
        self.ctx = ctx
        self.files = list(files)
        arr = (_vp * len(self.files))(*[f.h for f in self.files])
        p = _vp()
        ctx.check(lib().wb_cloud_create(ctx.h, arr, len(self.files), float(unit), ZEROS[zeros],
                                        C.byref(p)), "wb_cloud_create")
        self.h = p
        ctx.las.add(self)
        self.unit = float(unit)
        n = C.c_uint64()
        res = (C.c_double * 3)()
        jf = C.c_int()
        corners = (C.c_double * (6 * len(self.files)))()
        lib().wb_cloud_info(p, C.byref(n), res, C.byref(jf), corners)
        self.n = n.value
        ms = C.c_double()
        lib().wb_cloud_decode_ms(p, C.byref(ms))
        self.ms_decode = ms.value
        self.resolution = np.array(list(res))
        self.join_format = jf.value
        self.corners = np.array(list(corners)).reshape(-1, 3)

    def points(self):
        """Host copies (x, y, z, record index in its file) of the decoded points."""
        dp = C.POINTER(C.c_double)
        x, y, z = (np.zeros(max(self.n, 1)) for _ in range(3))
        rec = np.zeros(max(self.n, 1), dtype=np.uint32)
        self.ctx.check(lib().wb_cloud_read_points(self.h, x.ctypes.data_as(dp),
                                                  y.ctypes.data_as(dp), z.ctypes.data_as(dp),
                                                  rec.ctypes.data_as(C.POINTER(C.c_uint32))),
                       "wb_cloud_read_points")
        return x[:self.n], y[:self.n], z[:self.n], rec[:self.n]

    def size_fit(self):
        """Octree::sizeFit over the header corners (wolkencli.cpp:85-90)."""
        return Octree.size_fit_corners(self.corners)

    def build(self, center=None, side=None):
        """The octree of the cloud, duplicates merged (OctBuffer::put)."""
        if center is None:
            center, side = self.size_fit()
        cen = np.ascontiguousarray(center, dtype=np.float64)
        h = _vp()
        self.ctx.check(lib().wb_cloud_build(self.h, cen.ctypes.data_as(C.POINTER(C.c_double)),
                                            float(side), 0, C.byref(h)), "wb_cloud_build")
        t = Octree.__new__(Octree)
        t.ctx = self.ctx
        return t._adopt(self.ctx, h, self.n, cen, side)

    def classify(self, tree, params=None, want_tiles=True, out=None):
        """SCAN + POSTSCAN + CLASSIFY with the tiling of every file's header
        corners.  Returns (class bytes in cloud order, tiles or None, stats);
        with `out` a device tensor (uint8, n) the classes stay in HBM there."""
        from . import tile_count_corners
        params = params or default_params()
        ntiles = tile_count_corners(self.corners, params.tile_size) if want_tiles else 0
        tiles = (TileRecord * max(ntiles, 1))() if want_tiles else None
        st = Stats()
        if out is not None:
            cls, ptr, flags = out, out.data_ptr(), WB_DEVICE_INPUT
        else:
            cls = np.zeros(self.n, dtype=np.uint8)
            ptr, flags = cls.ctypes.data, 0
        rc = lib().wb_cloud_classify(self.h, tree.h, C.byref(params), ptr, flags,
                                     tiles if want_tiles else None, ntiles, C.byref(st))
        self.ctx.check(rc, "wb_cloud_classify")
        tarr = np.ctypeslib.as_array(tiles)[:ntiles].copy() if want_tiles else None
        return cls, tarr, st.as_dict()

    def write(self, tree, cls, name, separate_classes=True, points_per_file=10000000,
              point_format=-1, creation=(0, 0), dry=False):
        """CloudOutput::openFiles + writeFiles + closeFiles (cloudoutput.cpp:
        119-246).  cls: class bytes in cloud order (None: records keep their
        class).  Returns (stats dict, list of file names)."""
        prm = LasWriteParams(str(name).encode(), int(bool(separate_classes)),
                             int(points_per_file), int(point_format), int(creation[0]),
                             int(creation[1]))
        st = LasWriteStats()
        names = C.create_string_buffer(1 << 16)
        ptr, flags = None, WB_WRITE_DRY if dry else 0
        if cls is not None:
            if hasattr(cls, "data_ptr"):
                ptr, flags = cls.data_ptr(), flags | WB_DEVICE_INPUT
            else:
                cls = np.ascontiguousarray(cls, dtype=np.uint8)
                ptr = cls.ctypes.data
        rc = lib().wb_cloud_write(self.h, tree.h, ptr, flags, C.byref(prm), C.byref(st), names,
                                  1 << 16)
        self.ctx.check(rc, "wb_cloud_write")
        d = {f: getattr(st, f) for f, _ in LasWriteStats._fields_}
        d["class_totals"] = {k: v for k, v in enumerate(st.class_totals) if v}
        d["scale"], d["offset"] = list(st.scale), list(st.offset)
        return d, [s for s in names.value.decode().split("\n") if s]

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            lib().wb_cloud_free(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# CloudOutput::className (cloudoutput.cpp:39-117)
_CLASS_NAMES = ["raw", "nonground", "ground", "lowveg", "medveg", "highveg", "building",
                "lownoise", None, "water", "rail", "road", "overlap", "wireguard", "conductor",
                "tower", "insulator", "bridge", "highnoise", "overhead", "ignground", "snow",
                "temporal"]


def class_name(n):
    if 0 <= n < len(_CLASS_NAMES) and _CLASS_NAMES[n]:
        return _CLASS_NAMES[n]
    return f"reserved{n}"
