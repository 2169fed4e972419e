"""The store's block format (octree.h:35-43, octree.cpp:427-559) in the
oracle: the 61-byte record of LasPoint::write against an independent layout
from the field list of las.cpp:177-207, the block layout of
OctBuffer::write, and the reading rules of OctBuffer::read."""
import math
import struct

import numpy as np

import oracle_las as OL
from wolkenbase_amd import cli

# LasPoint::write (las.cpp:177-207) without WAVEFORM: location (3 doubles),
# intensity, returnNum, nReturns (shorts), edgeLine*2+scanDirection (byte),
# classification, classificationFlags, scannerChannel, userData, pointSource
# (shorts), scanAngle (int), gpsTime (double), nir, red, green, blue (shorts)
LAYOUT = "<dddHHHBHHHHHidHHHH"


def point(**kw):
    p = OL.Point()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_record_layout_against_the_field_list():
    assert struct.calcsize(LAYOUT) == 61
    p = point(x=1.5, y=-2.25, z=1e6 + 0.125, intensity=0x1234, ret=3, nret=5, scan_dir=1, edge=1,
              cls=0x0207, cls_flags=9, channel=2, user=0xbeef, source=0x7777,
              scan_angle=-123456789, gps=3.75e8, nir=11, red=22, green=33, blue=44)
    want = struct.pack(LAYOUT, 1.5, -2.25, 1e6 + 0.125, 0x1234, 3, 5, 1 * 2 + 1, 0x0207, 9, 2,
                       0xbeef, 0x7777, -123456789, 3.75e8, 11, 22, 33, 44)
    assert OL.store_record(p) == want
    # the flags byte: edgeLine*2 + scanDirection
    assert OL.store_record(point(edge=1, scan_dir=0))[30] == 2
    assert OL.store_record(point(edge=0, scan_dir=1))[30] == 1


def test_record_matches_wolkencli_testfile():
    # the record wolkencli writes 537 times to "testfile" (wolkencli.cpp:
    # 112-118): (pi, e, sqrt 2) with LasPoint's default fields
    p = point(x=math.pi, y=math.exp(1), z=math.sqrt(2))
    assert OL.store_record(p) == cli.store_record(math.pi, math.exp(1), math.sqrt(2))


def test_block_layout():
    pts = [point(x=float(i), y=2.0 * i, z=-i, cls=i % 7, intensity=i) for i in range(5)]
    blk = OL.store_blocks(pts, [5])
    assert len(blk) == 32768
    for i in range(5):
        assert blk[61 * i:61 * (i + 1)] == OL.store_record(pts[i])
    # noPoint (las.cpp:122-132): NaN location, zero fields
    nan = struct.pack("<d", float("nan"))
    for r in (5, 100, 536):
        rec = blk[61 * r:61 * (r + 1)]
        assert rec[:24] == nan * 3 and rec[24:] == bytes(37)
    assert blk[537 * 61:] == bytes(32768 - 537 * 61) and 32768 - 537 * 61 == 11
    # a full leaf, and the empty tree's leaves
    full = OL.store_blocks([point(x=float(i)) for i in range(537)], [537])
    assert all(full[61 * r:61 * r + 8] == struct.pack("<d", float(r)) for r in range(537))
    assert OL.store_blocks([], []) == b""


def test_read_rules():
    rng = np.random.default_rng(3)
    pts = [point(x=float(v[0]), y=float(v[1]), z=float(v[2]), cls=int(c), gps=float(g),
                 scan_angle=int(a), red=int(r))
           for v, c, g, a, r in zip(rng.normal(size=(40, 3)) * 100, rng.integers(0, 30, 40),
                                    rng.normal(size=40) * 1e5, rng.integers(-2**31, 2**31, 40),
                                    rng.integers(0, 65536, 40))]
    blk = OL.store_blocks(pts, [40])
    back = OL.store_read_block(blk)
    assert len(back) == 40
    for a, b in zip(pts, back):
        assert OL.store_record(a) == OL.store_record(b)
    # a block never written (zeros): every location (0,0,0), so none
    assert OL.store_read_block(bytes(32768)) == []
    # a single point at the origin is kept; two equal first locations empty the block
    one = OL.store_blocks([point(x=0.0, y=0.0, z=0.0)], [1])
    assert len(OL.store_read_block(one)) == 1
    two = OL.store_blocks([point(x=1.0, y=2.0, z=3.0), point(x=1.0, y=2.0, z=3.0),
                           point(x=4.0)], [3])
    assert OL.store_read_block(two) == []
    # a NaN in any coordinate makes a record empty (point.cpp:252-255)
    mixed = OL.store_blocks([point(x=1.0), point(y=float("nan")), point(x=2.0)], [3])
    assert [p.x for p in OL.store_read_block(mixed)] == [1.0, 2.0]
