"""wolkencli (wolkencli.cpp:35-129) on the GPU path.

    python -m wolkenbase_amd.cli a.las [b.las ...]

Opens the LAS files and prints each one's version, point count and format;
fits the root cube to the header corners and prints it as (cx,cy,cz)±side;
puts every record into the octree (wolkencli keeps records with return
number 0), merging points at one location as the store does, and prints
the number of duplicates; writes the octree dump to "dumpfile" and
RECORDS copies of the 61-byte store record of the point (pi, e, sqrt 2) to
"testfile", both in the working directory.

With --classify NAME it goes on as the GUI's save does (wolkencanvas.cpp:
469-625): the READ phase's zero rule, SCAN, POSTSCAN, CLASSIFY, then
class-separated LAS files NAME-<class>-<k>.las of at most 10000000 points.
"""
import argparse
import math
import struct
import sys
import time

import wolkenbase_amd as W
from wolkenbase_amd import las as WL

RECORDS = 537  # octree.h:40-41


def store_record(x, y, z):
    """LasPoint::write (las.cpp:177-207) of a point with default fields:
    the 61-byte record of the store's block files (built without WAVEFORM)."""
    return (struct.pack("<ddd", x, y, z) + struct.pack("<HHH", 0, 0, 0) + b"\0" +
            struct.pack("<HHHH", 0, 0, 0, 0) + struct.pack("<H", 0) + struct.pack("<i", 0) +
            struct.pack("<d", 0.0) + struct.pack("<HHHH", 0, 0, 0, 0))


def main(argv=None):
    ap = argparse.ArgumentParser(description="wolkencli on the GPU path")
    ap.add_argument("input", nargs="*")
    ap.add_argument("--classify", metavar="NAME", help="classify and write LAS files NAME-*.las")
    ap.add_argument("--unit", type=float, default=1.0, help="length unit in metres")
    args = ap.parse_args(argv)
    ctx = W.Context(0)
    files = []
    for name in args.input:
        f = WL.LasFile(ctx, path=name)
        ver = f.get_version()
        print(f"Version {ver >> 8}.{ver & 255} {f.number_points()} points, "
              f"format {f.get_point_format()}")
        files.append(f)
    if not files:
        ctx.close()
        return 0
    cloud = WL.Cloud(ctx, files, unit=args.unit, zeros="read" if args.classify else "keep")
    center, side = cloud.size_fit()
    print(f"({W.ldecimal(center[0])},{W.ldecimal(center[1])},{W.ldecimal(center[2])})±{side:g}")
    with open("testfile", "wb") as tf:
        tf.write(store_record(math.pi, math.exp(1), math.sqrt(2)) * RECORDS)
    t0 = time.perf_counter()
    tree = cloud.build(center, side)
    info = tree.info()
    for f in files:
        print(f"{f.number_points()} points")
    print("All points in octree")
    print(f"{tree.ndup} duplicate points")
    print(f"{info['nleaves']} blocks")
    print("Dumping octree")
    tree.dump("dumpfile")
    if args.classify:
        cls, _, st = cloud.classify(tree, want_tiles=False)
        print("Classified points:")
        ws, names = cloud.write(tree, cls, args.classify, separate_classes=True,
                                points_per_file=10000000)
        for k, v in sorted(ws["class_totals"].items()):
            print(k, v)
        for n in names:
            print(n)
        print(f"{cloud.n} points in {1e3 * (time.perf_counter() - t0):.1f} ms "
              f"(build {st['ms_build']:.1f}, classify {st['ms_total']:.1f}, "
              f"write {ws['ms_encode'] + ws['ms_copy'] + ws['ms_disk']:.1f})")
    tree.free()
    cloud.close()
    for f in files:
        f.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
