"""LAS file images packed field by field from the published ASPRS LAS 1.2
and 1.4 (R15) specifications: the public header block and point data
record formats 0-10.  Written independently of the oracle's writer, so a
mistake in either side's byte layout shows up as a mismatch.

Test infrastructure only.
"""
import struct

import numpy as np

# record lengths of formats 0-10 (LAS 1.4 R15, tables 7-20)
LENGTH = [20, 28, 26, 34, 57, 63, 30, 36, 38, 59, 67]
GPS = {1, 3, 4, 5, 6, 7, 8, 9, 10}
RGB = {2, 3, 5, 7, 8, 10}
NIR = {8, 10}
WAVE = {4, 5, 9, 10}


def record(fmt, f, extra=0):
    """One point record.  f: dict with xi, yi, zi, intensity, ret, nret,
    scan_dir, edge, cls, cls_flags (3 bits <6, 4 bits >=6), channel, user,
    scan_angle (rank in degrees <6, raw short >=6), source, gps, red, green,
    blue, nir, and for wave formats wave (descriptor index, byte offset,
    size, location, x, y, z)."""
    b = struct.pack("<iiiH", f["xi"], f["yi"], f["zi"], f["intensity"])
    if fmt < 6:
        b += struct.pack("<BBbBH",
                         (f["ret"] & 7) | ((f["nret"] & 7) << 3) | (f["scan_dir"] << 6) |
                         (f["edge"] << 7),
                         (f["cls"] & 31) | ((f["cls_flags"] & 7) << 5),
                         f["scan_angle"], f["user"], f["source"])
    else:
        b += struct.pack("<BBBBhH",
                         (f["ret"] & 15) | ((f["nret"] & 15) << 4),
                         (f["cls_flags"] & 15) | ((f["channel"] & 3) << 4) |
                         (f["scan_dir"] << 6) | (f["edge"] << 7),
                         f["cls"], f["user"], f["scan_angle"], f["source"])
    if fmt in GPS:
        b += struct.pack("<d", f["gps"])
    if fmt in RGB:
        b += struct.pack("<HHH", f["red"], f["green"], f["blue"])
    if fmt in NIR:
        b += struct.pack("<H", f["nir"])
    if fmt in WAVE:
        b += struct.pack("<BQIffff", *f.get("wave", (0, 0, 0, 0.0, 0.0, 0.0, 0.0)))
    assert len(b) == LENGTH[fmt]
    return b + bytes(extra)


def header(version, fmt, rec_len, n_legacy, n_by_ret_legacy, scale, offset, mins, maxs,
           n14=None, n14_by_ret=None, point_offset=None, n_vlr=0, system_id=b"TEST",
           software=b"las_spec", day=1, year=2020, source_id=7, encoding=0, guid=bytes(16)):
    """Public header block: 227 bytes (1.0-1.2), 235 (1.3) or 375 (1.4)."""
    size = {0: 227, 1: 227, 2: 227, 3: 235, 4: 375}[version[1]]
    if point_offset is None:
        point_offset = size
    h = b"LASF" + struct.pack("<HH", source_id, encoding) + guid
    h += struct.pack("<BB", *version)
    h += system_id.ljust(32, b"\0") + software.ljust(32, b"\0")
    h += struct.pack("<HHHIIBH", day, year, size, point_offset, n_vlr, fmt, rec_len)
    h += struct.pack("<I", n_legacy) + struct.pack("<5I", *n_by_ret_legacy)
    h += struct.pack("<3d", *scale) + struct.pack("<3d", *offset)
    h += struct.pack("<6d", maxs[0], mins[0], maxs[1], mins[1], maxs[2], mins[2])
    assert len(h) == 227
    if size >= 235:
        h += struct.pack("<Q", 0)  # start of waveform data
    if size == 375:
        h += struct.pack("<QI", 0, 0)  # start of EVLRs, number of EVLRs
        h += struct.pack("<Q", n14 if n14 is not None else 0)
        h += struct.pack("<15Q", *(n14_by_ret if n14_by_ret is not None else [0] * 15))
    assert len(h) == size
    return h + bytes(point_offset - size)


def random_fields(rng, n, fmt, xyz_range=200000):
    out = []
    for i in range(n):
        f = dict(xi=int(rng.integers(-xyz_range, xyz_range)),
                 yi=int(rng.integers(-xyz_range, xyz_range)),
                 zi=int(rng.integers(-20000, 20000)),
                 intensity=int(rng.integers(0, 65536)),
                 ret=int(rng.integers(0, 8 if fmt < 6 else 16)),
                 nret=int(rng.integers(0, 8 if fmt < 6 else 16)),
                 scan_dir=int(rng.integers(0, 2)), edge=int(rng.integers(0, 2)),
                 cls=int(rng.integers(0, 32 if fmt < 6 else 256)),
                 cls_flags=int(rng.integers(0, 8 if fmt < 6 else 16)),
                 channel=int(rng.integers(0, 4)) if fmt >= 6 else 0,
                 user=int(rng.integers(0, 256)),
                 scan_angle=int(rng.integers(-90, 91)) if fmt < 6 else
                 int(rng.integers(-30000, 30001)),
                 source=int(rng.integers(0, 65536)),
                 gps=float(rng.uniform(0, 1e6)),
                 red=int(rng.integers(0, 65536)), green=int(rng.integers(0, 65536)),
                 blue=int(rng.integers(0, 65536)), nir=int(rng.integers(0, 65536)),
                 wave=(int(rng.integers(0, 256)), int(rng.integers(0, 2**40)),
                       int(rng.integers(0, 2**31)), 1.5, 0.25, -0.5, 0.125))
        out.append(f)
    return out


def image(fmt, fields, version=(1, 4), scale=(0.01, 0.01, 0.01), offset=(1000.0, 2000.0, 0.0),
          extra=0, bounds=None, n_vlr=0, vlr_bytes=0, by_return=True):
    """A whole file: header (+ vlr_bytes of padding standing for VLRs) and
    records.  bounds: (mins, maxs); default the decoded extent.  by_return
    False writes the total with zero per-return counts (as photogrammetric
    files with return number 0 do)."""
    n = len(fields)
    rec = b"".join(record(fmt, f, extra) for f in fields)
    if bounds is None:
        if n:
            xs = [offset[0] + scale[0] * f["xi"] for f in fields]
            ys = [offset[1] + scale[1] * f["yi"] for f in fields]
            zs = [offset[2] + scale[2] * f["zi"] for f in fields]
            bounds = ((min(xs), min(ys), min(zs)), (max(xs), max(ys), max(zs)))
        else:
            bounds = ((0, 0, 0), (0, 0, 0))
    by_ret = [0] * 15
    for f in fields:
        if 1 <= f["ret"] <= 15 and by_return:
            by_ret[f["ret"] - 1] += 1
    size = {0: 227, 1: 227, 2: 227, 3: 235, 4: 375}[version[1]]
    legacy_ok = fmt < 6 and all(v == 0 for v in by_ret[5:])
    h = header(version, fmt, LENGTH[fmt] + extra, n if legacy_ok else 0,
               by_ret[:5] if legacy_ok else [0] * 5, scale, offset, bounds[0], bounds[1],
               n14=n, n14_by_ret=by_ret, point_offset=size + vlr_bytes, n_vlr=n_vlr)
    return h + rec
