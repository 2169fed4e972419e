"""Seeded synthetic LAS tiles for the benchmark and the tests (ctypes over
libwb_synth.so, built from csrc/wb_synth.c)."""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_SO = os.path.join(HERE, "libwb_synth.so")


class _Config(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("extent", C.c_double), ("density", C.c_double),
                ("buildings_per_ref", C.c_double), ("trees_per_ref", C.c_double),
                ("n_embankments", C.c_int), ("outlier_frac", C.c_double),
                ("linear_frac", C.c_double), ("origin_x", C.c_double), ("origin_y", C.c_double)]


class _Cloud(C.Structure):
    _fields_ = [("x", C.POINTER(C.c_int32)), ("y", C.POINTER(C.c_int32)),
                ("z", C.POINTER(C.c_int32)), ("intensity", C.POINTER(C.c_uint16)),
                ("returns", C.POINTER(C.c_uint8)), ("truth", C.POINTER(C.c_uint8)),
                ("n", C.c_long), ("scale", C.c_double * 3), ("offset", C.c_double * 3),
                ("bbmin", C.c_double * 3), ("bbmax", C.c_double * 3)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_SO):
            raise RuntimeError("libwb_synth.so is not built; run __graft_entry__.build()")
        L = C.CDLL(SYNTH_SO)
        L.wb_synth_generate.argtypes = [C.POINTER(_Config), C.POINTER(_Cloud)]
        L.wb_synth_free.argtypes = [C.POINTER(_Cloud)]
        _lib = L
    return _lib


# Workloads: the 1M-point CPU-runnable tile, the 20M-point tile (with slope
# breaks and outliers), the same tile with thin linear features for the
# precision study.  Feature counts are per 200 m x 200 m.
PRESETS = {
    "tile1m": dict(seed=1000, extent=200.0, density=25.0, buildings=40, trees=300,
                   embankments=0, outliers=0.0, linear=0.0),
    "tile20m": dict(seed=2000, extent=900.0, density=25.0, buildings=40, trees=300,
                    embankments=5, outliers=0.02, linear=0.0),
    "tile20m_linear": dict(seed=2000, extent=900.0, density=25.0, buildings=40, trees=300,
                           embankments=5, outliers=0.02, linear=0.05),
}


def generate(seed=1000, extent=200.0, density=25.0, buildings=40, trees=300, embankments=0,
             outliers=0.0, linear=0.0, origin=(500000.0, 4400000.0)):
    """Returns a dict: xi, yi, zi (int32), intensity, returns, truth, scale,
    offset, bbmin, bbmax (float64[3])."""
    L = _load()
    cfg = _Config(seed, extent, density, buildings, trees, embankments, outliers, linear,
                  origin[0], origin[1])
    cl = _Cloud()
    L.wb_synth_generate(C.byref(cfg), C.byref(cl))
    n = cl.n
    try:
        out = dict(
            xi=np.ctypeslib.as_array(cl.x, (n,)).copy() if n else np.zeros(0, np.int32),
            yi=np.ctypeslib.as_array(cl.y, (n,)).copy() if n else np.zeros(0, np.int32),
            zi=np.ctypeslib.as_array(cl.z, (n,)).copy() if n else np.zeros(0, np.int32),
            intensity=np.ctypeslib.as_array(cl.intensity, (n,)).copy() if n else np.zeros(0, np.uint16),
            returns=np.ctypeslib.as_array(cl.returns, (n,)).copy() if n else np.zeros(0, np.uint8),
            truth=np.ctypeslib.as_array(cl.truth, (n,)).copy() if n else np.zeros(0, np.uint8),
            scale=np.array(cl.scale[:]), offset=np.array(cl.offset[:]),
            bbmin=np.array(cl.bbmin[:]), bbmax=np.array(cl.bbmax[:]),
        )
    finally:
        L.wb_synth_free(C.byref(cl))
    return out


def preset(name, **over):
    p = dict(PRESETS[name])
    p.update(over)
    return generate(**p)


def serving_tile(k, rng):
    """Tile k of the serving stream: lognormal point count (median 2M, clipped
    to [50k, 40M]) at a density of 5-60 pts/m^2, seeded by 3,000,000+k."""
    n = float(np.clip(rng.lognormal(math.log(2e6), 1.0), 5e4, 4e7))
    dens = float(rng.uniform(5, 60))
    extent = math.sqrt(n / dens)
    return dict(seed=3000000 + k, extent=extent, density=dens, buildings=40, trees=300,
                embankments=0, outliers=0.0, linear=0.0)
