"""ctypes bindings to the LAS part of the CPU oracle (oracle/wb_oracle_las.c)
and the oracle's whole LAS pipeline: read files, READ-phase zero rule,
duplicate merge, build, scan/postscan/classify, class-separated output.

Test infrastructure only (the checker): imported by tests/ alone.
"""
import ctypes as C
import os

import numpy as np

import oracle_lib as O

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)
_ip = C.POINTER(C.c_int)


class Header(C.Structure):
    """wbo_las_header (same layout as wb_las_header)."""
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


class Point(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("intensity", C.c_uint16), ("ret", C.c_uint16), ("nret", C.c_uint16),
                ("scan_dir", C.c_uint8), ("edge", C.c_uint8), ("cls", C.c_uint16),
                ("cls_flags", C.c_uint16), ("channel", C.c_uint16), ("user", C.c_uint16),
                ("source", C.c_uint16), ("scan_angle", C.c_int32), ("gps", C.c_double),
                ("nir", C.c_uint16), ("red", C.c_uint16), ("green", C.c_uint16),
                ("blue", C.c_uint16)]


class Out(C.Structure):
    _fields_ = [("format", C.c_int), ("scale", C.c_double * 3), ("offset", C.c_double * 3),
                ("unit", C.c_double), ("min", C.c_double * 3), ("max", C.c_double * 3),
                ("n_points", C.c_uint64 * 16)]


POINT_LENGTH = [20, 28, 26, 34, 57, 63, 30, 36, 38, 59, 67]
_ready = False


def lib():
    global _ready
    L = O.lib()
    if not _ready:
        L.wbo_las_parse.argtypes = [C.c_char_p, C.c_size_t, C.POINTER(Header)]
        L.wbo_las_read_point.argtypes = [C.c_char_p, C.POINTER(Header), C.c_double,
                                         C.POINTER(Point)]
        L.wbo_las_out_init.argtypes = [C.POINTER(Out), C.c_int, _dp, _dp, C.c_double]
        L.wbo_las_write_point.argtypes = [C.c_char_p, C.POINTER(Out), C.POINTER(Point)]
        L.wbo_las_write_header.argtypes = [C.c_char_p, C.POINTER(Out), C.c_char_p, C.c_char_p,
                                           C.c_int, C.c_int]
        L.wbo_join_point_format.argtypes = [_ip, C.c_int]
        L.wbo_manygcd.argtypes = [_dp, C.c_int, C.c_double]
        L.wbo_manygcd.restype = C.c_double
        L.wbo_combine1_scale.argtypes = [_dp, _dp, C.c_int]
        L.wbo_combine1_scale.restype = C.c_double
        L.wbo_set_scale.argtypes = [C.c_double] * 4 + [_dp, _dp]
        L.wbo_dedupe.argtypes = [_dp, _dp, _dp, C.c_long, _lp, _lp]
        L.wbo_dedupe.restype = C.c_long
        L.wbo_class_name.argtypes = [C.c_int, C.c_char_p, C.c_int]
        L.wbo_output_plan.argtypes = [_lp, _u8p, C.c_long, _u64p, C.c_long, C.c_int, _ip, _ip,
                                      _lp]
        L.wbo_ldecimal.argtypes = [C.c_double, C.c_double, C.c_char_p, C.c_int]
        L.wbo_degtobin.argtypes = [C.c_double]
        L.wbo_degtobin.restype = C.c_int
        L.wbo_bintodeg.argtypes = [C.c_int]
        L.wbo_bintodeg.restype = C.c_double
        L.wbo_bound_rect.argtypes = [_dp, C.c_int, _dp]
        L.wbo_run_xyz.argtypes = [_dp, _dp, _dp, C.c_long, _dp, _dp, C.c_int,
                                  C.POINTER(O.Params), _u8p, C.POINTER(O.Tile),
                                  C.POINTER(O.Stats)]
        L.wbo_run_xyz.restype = C.c_long
        L.wbo_sizefit.argtypes = [_dp, C.c_int, _dp, _dp]
        L.wbo_store_record.argtypes = [C.POINTER(Point), C.c_char_p]
        L.wbo_store_block.argtypes = [C.POINTER(Point), C.c_int, C.c_char_p]
        L.wbo_store_block.restype = C.c_int
        L.wbo_store_read_record.argtypes = [C.c_char_p, C.POINTER(Point)]
        L.wbo_store_read_block.argtypes = [C.c_char_p, C.POINTER(Point)]
        L.wbo_store_read_block.restype = C.c_int
        _ready = True
    return L


def parse(image):
    h = Header()
    ok = lib().wbo_las_parse(bytes(image), len(image), C.byref(h))
    return h if ok else None


def read_points(image, h, unit=1.0):
    """Every record of the file decoded (LasHeader::readPoint)."""
    L = lib()
    n = h.n_points[0]
    out = []
    for i in range(n):
        off = h.point_offset + i * h.point_length
        p = Point()
        L.wbo_las_read_point(bytes(image[off:off + h.point_length]), C.byref(h), unit, C.byref(p))
        out.append(p)
    return out


def degtobin(a):
    return lib().wbo_degtobin(float(a))


def ldecimal(x, toler=0.0):
    b = C.create_string_buffer(128)
    lib().wbo_ldecimal(float(x), float(toler), b, 128)
    return b.value.decode()


def join_point_format(formats):
    f = (C.c_int * len(formats))(*formats)
    return lib().wbo_join_point_format(f, len(formats))


def combine1_scale(scales, offsets):
    s = np.ascontiguousarray(scales, dtype=np.float64)
    o = np.ascontiguousarray(offsets, dtype=np.float64)
    return lib().wbo_combine1_scale(s.ctypes.data_as(_dp), o.ctypes.data_as(_dp), len(s))


def bound_rect(corners):
    c = np.ascontiguousarray(corners, dtype=np.float64).reshape(-1)
    out = np.zeros(6)
    lib().wbo_bound_rect(c.ctypes.data_as(_dp), len(c) // 3, out.ctypes.data_as(_dp))
    return out


def class_name(k):
    b = C.create_string_buffer(32)
    lib().wbo_class_name(k, b, 32)
    return b.value.decode()


def dedupe(x, y, z):
    n = len(x)
    first = np.zeros(max(n, 1), dtype=np.int64)
    last = np.zeros(max(n, 1), dtype=np.int64)
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    nd = lib().wbo_dedupe(x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), z.ctypes.data_as(_dp), n,
                          first.ctypes.data_as(_lp), last.ctypes.data_as(_lp))
    return nd, first[:n], last[:n]


def output_plan(leaf_of, cls, totals, ppf, separate):
    n = len(leaf_of)
    lo = np.ascontiguousarray(leaf_of, dtype=np.int64)
    cl = np.ascontiguousarray(cls, dtype=np.uint8)
    tot = np.ascontiguousarray(totals, dtype=np.uint64)
    nfiles = np.zeros(256, dtype=np.int32)
    file_of = np.zeros(max(n, 1), dtype=np.int32)
    pos_of = np.zeros(max(n, 1), dtype=np.int64)
    ndig = lib().wbo_output_plan(lo.ctypes.data_as(_lp), cl.ctypes.data_as(_u8p), n,
                                 tot.ctypes.data_as(_u64p), int(ppf), int(bool(separate)),
                                 nfiles.ctypes.data_as(_ip), file_of.ctypes.data_as(_ip),
                                 pos_of.ctypes.data_as(_lp))
    return ndig, nfiles, file_of[:n], pos_of[:n]


def _ndecimal(i, dig):
    return f"{i:0{dig}d}" if dig else ""


def pipeline(images, params=None, unit=1.0, zeros="read", name="out", separate=True,
             ppf=10000000, point_format=-1, day=1, year=2020, classify=True):
    """The oracle's LAS pipeline.  Returns dict with the decoded cloud (x, y,
    z), the class bytes in cloud order, the number of merged duplicates, the
    canonical order of stored points, the dump text's inputs and the output
    files {file name: bytes}."""
    L = lib()
    params = params or O.default_params()
    hdrs = [parse(img) for img in images]
    pts = []       # (file, record index, Point) per cloud point
    for f, (img, h) in enumerate(zip(images, hdrs)):
        recs = read_points(img, h, unit)
        drop = zeros == "read" and len(recs) > 0 and recs[0].ret != 0
        fix = zeros == "read" and not drop
        for i, p in enumerate(recs):
            if drop and p.ret == 0:
                continue
            if fix and p.ret == 0:
                p.ret = 1
            pts.append((f, i, p))
    n = len(pts)
    x = np.array([p.x for _, _, p in pts])
    y = np.array([p.y for _, _, p in pts])
    z = np.array([p.z for _, _, p in pts])
    nd, first, last = dedupe(x, y, z) if n else (0, np.zeros(0, np.int64), np.zeros(0, np.int64))
    distinct = np.nonzero(first == np.arange(n))[0]
    corners = []
    for h in hdrs:
        corners += [h.min[a] * unit for a in range(3)] + [h.max[a] * unit for a in range(3)]
    corners = np.array(corners)
    res = np.array([combine1_scale([h.scale[a] * unit for h in hdrs],
                                   [h.offset[a] * unit for h in hdrs]) for a in range(3)])
    cen = np.zeros(3)
    side = np.zeros(1)
    L.wbo_sizefit(corners.ctypes.data_as(_dp), len(hdrs) * 2, cen.ctypes.data_as(_dp),
                  side.ctypes.data_as(_dp))
    xd, yd, zd = (np.ascontiguousarray(a[distinct]) for a in (x, y, z))
    cls_d = np.zeros(len(distinct), dtype=np.uint8)
    if classify and len(distinct):
        sn = O.Snake()
        O.lib().wbo_snake_setup_corners.argtypes = [_dp, C.c_int, C.c_double, C.POINTER(O.Snake)]
        O.lib().wbo_snake_setup_corners(corners.ctypes.data_as(_dp), len(hdrs) * 2,
                                        params.tile_size, C.byref(sn))
        tiles = (O.Tile * (sn.hi - sn.lo + 1))()
        st = O.Stats()
        rc = L.wbo_run_xyz(xd.ctypes.data_as(_dp), yd.ctypes.data_as(_dp),
                           zd.ctypes.data_as(_dp), len(xd), res.ctypes.data_as(_dp),
                           corners.ctypes.data_as(_dp), len(hdrs) * 2, C.byref(params),
                           cls_d.ctypes.data_as(_u8p), tiles, C.byref(st))
        if rc < 0:
            raise ValueError("more than 537 points in a level-21 cell")
    # class bytes in cloud order: a merged point takes its location's class
    rank = np.full(n, -1, dtype=np.int64)
    rank[distinct] = np.arange(len(distinct))
    cls = cls_d[rank[first]] if n else np.zeros(0, np.uint8)
    # the stored points in canonical order
    tree = O.Tree(xd, yd, zd, cen, float(side[0]))
    lv = tree.leaves()
    order_d = tree.order()
    leaf_of = np.repeat(np.arange(len(lv["count"])), lv["count"])
    canon = distinct[order_d]                 # cloud index (first) of each stored point
    recs_out = []
    ocls = np.zeros(len(canon), dtype=np.uint8)
    for k, c in enumerate(canon):
        f, i, p = pts[last[c]]
        q = Point.from_buffer_copy(p)
        if cls[c]:
            q.cls = int(cls[c])
        ocls[k] = q.cls & 255
        recs_out.append(q)
    totals = np.bincount(ocls, minlength=256).astype(np.uint64)
    ndig, nfiles, file_of, pos_of = output_plan(leaf_of, ocls, totals, ppf, separate)
    fo = point_format if point_format >= 0 else join_point_format([h.point_format for h in hdrs])
    br = bound_rect(corners)
    minc, maxc = (br[0], br[1], br[4]), (br[2], br[3], br[5])
    sc, of = np.zeros(3), np.zeros(3)
    for a in range(3):
        comb = combine1_scale([h.scale[a] * unit for h in hdrs], [h.offset[a] * unit for h in hdrs])
        s, o = C.c_double(), C.c_double()
        L.wbo_set_scale(minc[a], maxc[a], comb, unit, C.byref(s), C.byref(o))
        sc[a], of[a] = s.value, o.value
    sysid = b"EXTRACTION" if separate else (b"MERGE" if len(images) > 1 else b"MODIFICATION")
    files = {}
    LEN = POINT_LENGTH[fo]
    keys = [k for k in range(256) if nfiles[k]]
    for k in keys:
        for j in range(nfiles[k]):
            sel = np.nonzero(((ocls == k) if separate else np.ones(len(ocls), bool)) &
                             (file_of == j))[0]
            sel = sel[np.argsort(pos_of[sel], kind="stable")]
            out = Out()
            L.wbo_las_out_init(C.byref(out), fo, sc.ctypes.data_as(_dp), of.ctypes.data_as(_dp),
                               unit)
            body = bytearray(len(sel) * LEN)
            for m, idx in enumerate(sel):
                rec = C.create_string_buffer(LEN)
                L.wbo_las_write_point(rec, C.byref(out), C.byref(recs_out[idx]))
                body[m * LEN:(m + 1) * LEN] = rec.raw[:LEN]
            hd = C.create_string_buffer(375)
            L.wbo_las_write_header(hd, C.byref(out), sysid, b"Wolkenbase 0.1.2~alpha", day, year)
            fname = name + (("-" + class_name(k)) if separate else "") + \
                ("-" if ppf else "") + _ndecimal(j, ndig) + ".las"
            files[fname] = hd.raw[:375] + bytes(body)
    return dict(x=x, y=y, z=z, cls=cls, ndup=int(nd), canon=canon, totals=totals, files=files,
                leaves=lv, tree=tree, center=cen, side=float(side[0]), corners=corners,
                resolution=res, join_format=fo, stored=recs_out)


STORE_BLOCK = 32768


def store_record(p):
    """LasPoint::write (las.cpp:177-207) of one Point: 61 bytes."""
    out = C.create_string_buffer(61)
    lib().wbo_store_record(C.byref(p), out)
    return out.raw[:61]


def store_blocks(stored, counts):
    """OctBuffer::write (octree.cpp:427-460) of consecutive leaves: the
    stored Points in canonical order split by the leaves' counts."""
    out = bytearray()
    k = 0
    for c in counts:
        arr = (Point * max(int(c), 1))()
        for j in range(int(c)):
            arr[j] = stored[k + j]
        blk = C.create_string_buffer(STORE_BLOCK)
        assert lib().wbo_store_block(arr, int(c), blk) == 0
        out += blk.raw[:STORE_BLOCK]
        k += int(c)
    return bytes(out)


def store_read_block(blk):
    """OctBuffer::read (octree.cpp:508-559) of one block: list of Points."""
    arr = (Point * 537)()
    n = lib().wbo_store_read_block(bytes(blk), arr)
    return [arr[i] for i in range(n)]


def dump_text(tree, x, y, z):
    """OctStore::dump's text (octree.cpp:360-376, 673-689, 888-891) from the
    oracle tree, with the oracle's ldecimal."""
    lv = tree.leaves()
    order = tree.order()
    lines = []
    total = 0
    for l in range(len(lv["count"])):
        cx, cy, cz = lv["center"][l]
        h = lv["side"][l] / 2
        idx = order[lv["first"][l]:lv["first"][l] + lv["count"][l]]
        px, py, pz = x[idx], y[idx], z[idx]
        fin = int(np.sum(np.isfinite(px) & np.isfinite(py) & np.isfinite(pz)))
        inn = int(np.sum((np.abs(px - cx) <= h) & (np.abs(py - cy) <= h) & (np.abs(pz - cz) <= h)))
        s = "(%s,%s,%s)±%s %d points" % (ldecimal(cx), ldecimal(cy), ldecimal(cz),
                                             ldecimal(h), fin)
        if fin != inn:
            s += ", %d stray points" % (fin - inn)
        lines.append(s + "\n")
        total += fin
    return "".join(lines) + "%d total points\n" % total
