"""wb_count_points_in_noalloc, the batched count query that allocates no
device memory: through the C ABI it returns wb_count_points_in's counts on
consecutive calls (the second smaller, over the buffer the first left
behind), refuses batches above its maximum, and leaves free device memory
unchanged."""
import ctypes as C

import numpy as np
import pytest

import wolkenbase_amd as W
from wolkenbase_amd import synth

pytestmark = pytest.mark.gpu

KINDS = [W.WB_SHAPE_CYLINDER, W.WB_SHAPE_SPHERE, W.WB_SHAPE_HYPERBOLOID]


@pytest.fixture(scope="module")
def tree_and_points():
    ctx = W.Context(0)
    cloud = synth.generate(seed=77, extent=30.0, density=25.0)
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    tree = W.Octree.build(ctx, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                          cloud["offset"], center, side)
    pts = [cloud["offset"][a] + cloud["scale"][a] * cloud[k].astype(np.float64)
           for a, k in enumerate(("xi", "yi", "zi"))]
    yield ctx, tree, pts
    tree.free()
    ctx.close()


def shapes(rng, pts, kind, m):
    x, y, z = pts
    i = rng.integers(0, len(x), m)
    if kind == W.WB_SHAPE_CYLINDER:
        P = [x[i] + rng.normal(0, 0.3, m), y[i] + rng.normal(0, 0.3, m),
             rng.uniform(0.05, 3.0, m), np.zeros(m), np.zeros(m)]
    elif kind == W.WB_SHAPE_SPHERE:
        P = [x[i], y[i], z[i] + rng.normal(0, 0.5, m), rng.uniform(0.05, 4.0, m), np.zeros(m)]
    else:
        P = [x[i], y[i], z[i], rng.uniform(0.05, 1.0, m), rng.choice([1.0, 0.7, -1.0], m)]
    return np.ascontiguousarray(np.stack(P, 1))


def call_noalloc(tree, kind, P):
    """wb_count_points_in_noalloc through the C ABI: (return code, counts)."""
    out = np.full(len(P), 0xdeadbeef, dtype=np.uint64)
    rc = W.lib().wb_count_points_in_noalloc(
        tree.h, kind, P.ctypes.data_as(C.POINTER(C.c_double)), len(P),
        out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return rc, out


def call_alloc(tree, kind, P):
    out = np.zeros(len(P), dtype=np.uint64)
    rc = W.lib().wb_count_points_in(tree.h, kind, P.ctypes.data_as(C.POINTER(C.c_double)),
                                    len(P), out.ctypes.data_as(C.POINTER(C.c_uint64)))
    assert rc == 0
    return out


@pytest.mark.parametrize("kind", KINDS)
def test_counts_equal_the_allocating_entry_point(tree_and_points, kind):
    ctx, tree, pts = tree_and_points
    rng = np.random.default_rng(100 + kind)
    big, small = shapes(rng, pts, kind, 3000), shapes(rng, pts, kind, 257)
    for P in (big, small):  # the second call is smaller than the first
        rc, got = call_noalloc(tree, kind, P)
        assert rc == 0, W.lib().wb_last_error(ctx.h).decode()
        assert np.array_equal(got, call_alloc(tree, kind, P))
    assert int(got.sum()) > 0


def test_largest_batch_and_refusal_above_it(tree_and_points):
    ctx, tree, pts = tree_and_points
    rng = np.random.default_rng(5)
    P = shapes(rng, pts, W.WB_SHAPE_CYLINDER, W.WB_NOALLOC_MAX_SHAPES)
    rc, got = call_noalloc(tree, W.WB_SHAPE_CYLINDER, P)
    assert rc == 0
    assert np.array_equal(got, call_alloc(tree, W.WB_SHAPE_CYLINDER, P))
    P = shapes(rng, pts, W.WB_SHAPE_CYLINDER, W.WB_NOALLOC_MAX_SHAPES + 1)
    rc, got = call_noalloc(tree, W.WB_SHAPE_CYLINDER, P)
    assert rc == -4  # WB_ERANGE
    assert "65536" in W.lib().wb_last_error(ctx.h).decode()
    assert (got == 0xdeadbeef).all()  # nothing written


def test_free_device_memory_unchanged(tree_and_points):
    ctx, tree, pts = tree_and_points
    rng = np.random.default_rng(9)
    P = shapes(rng, pts, W.WB_SHAPE_SPHERE, 4000)
    assert call_noalloc(tree, W.WB_SHAPE_SPHERE, P)[0] == 0
    before, _ = W.device_memory(0)
    for m in (4000, 1000, 20000):
        assert call_noalloc(tree, W.WB_SHAPE_SPHERE, shapes(rng, pts, W.WB_SHAPE_SPHERE, m))[0] == 0
    after, _ = W.device_memory(0)
    assert after == before
