"""Precision study of the scan's plane fit (scan.cpp:53-73).

For each seeded workload the octree is built once and the classification
pass runs three times:
  fp64        the reference fit (WB_FIT_FP64; bit-identical to the oracle,
              which the parity tests establish),
  fp32_naive  FP32 moments with plain sums and an FP32 solve (WB_FIT_FP32),
  fp32_comp   FP32 moments with compensated sums and an FP32 solve, tiles
              whose xy spread is near-collinear refitted in FP64
              (WB_FIT_FP32_COMP).
Reported per mode against fp64: the largest and mean angle between the plane
normals (-a, -b, 1) over non-empty tiles, the largest change of a tile's
untilted height range (the residual spread the scan derives from the fit)
and of its density, the tiles whose tree flag changed, the class bytes that
flipped, the FP64 fallback rate and the scan time.

    python -m wolkenbase_amd.precision_study [--workloads tile1m,tile20m_linear]
"""
import argparse
import json
import sys

import numpy as np

import wolkenbase_amd as W
from wolkenbase_amd import synth

MODES = (("fp32_naive", W.WB_FIT_FP32), ("fp32_comp", W.WB_FIT_FP32_COMP))


def normal_angle(ax, ay, bx, by):
    """Angle (rad) between the normals (-ax, -ay, 1) and (-bx, -by, 1)."""
    n1 = np.stack([-ax, -ay, np.ones_like(ax)], 1)
    n2 = np.stack([-bx, -by, np.ones_like(bx)], 1)
    cr = np.linalg.norm(np.cross(n1, n2), axis=1)
    return np.arctan2(cr, (n1 * n2).sum(1))


def study(ctx, name):
    cloud = synth.generate(**synth.PRESETS[name])
    bbmin, bbmax = cloud["bbmin"], cloud["bbmax"]
    center, side = W.Octree.size_fit(bbmin, bbmax)
    tree = W.Octree.build(ctx, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                          cloud["offset"], center, side)
    ref_cls, ref, ref_st = W.classify(tree, bbmin, bbmax, W.default_params())
    live = ref["npoints"] > 0
    out = {"workload": name, "n_points": int(tree.n), "nonempty_tiles": int(live.sum()),
           "fp64": {"ms_scan": round(ref_st["ms_scan"], 3)}}
    for mode, prec in MODES:
        cls, t, st = W.classify(tree, bbmin, bbmax, W.default_params(fit_precision=prec))
        ang = np.degrees(normal_angle(t["tilt_x"][live], t["tilt_y"][live],
                                      ref["tilt_x"][live], ref["tilt_y"][live]))
        dh = np.abs(t["height"][live] - ref["height"][live])
        dd = np.abs(t["density"][live] - ref["density"][live]) / ref["density"][live]
        flips = int((cls != ref_cls).sum())
        out[mode] = {
            "max_normal_angle_deg": float(ang.max()),
            "mean_normal_angle_deg": float(ang.mean()),
            "max_height_change_m": float(dh.max()),
            "max_density_rel_change": float(dd.max()),
            "tree_flag_changes": int((t["tree_flags"][live] != ref["tree_flags"][live]).sum()),
            "class_flips": flips,
            "class_flip_fraction": flips / max(1, int(tree.n)),
            "fp64_fallback_tiles": int(st["fit_fallbacks"]),
            "fp64_fallback_rate": st["fit_fallbacks"] / max(1, int(live.sum())),
            "ms_scan": round(st["ms_scan"], 3),
        }
    tree.free()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workloads", default="tile1m,tile20m,tile20m_linear")
    args = ap.parse_args(argv)
    ctx = W.Context(0)
    for name in args.workloads.split(","):
        print(json.dumps(study(ctx, name)), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
