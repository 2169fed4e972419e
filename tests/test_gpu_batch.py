"""Several trees classified together (wb_classify_batch, one ground-kernel
launch for all) give each tree's class bytes and statistics exactly as
wb_classify of that tree alone, and the oracle's."""
import numpy as np
import pytest

import oracle_lib as O
import wolkenbase_amd as W
from wolkenbase_amd import synth

pytestmark = pytest.mark.gpu

CLOUDS = [dict(seed=71, extent=12.0, density=25.0),
          dict(seed=72, extent=30.0, density=10.0, embankments=1, outliers=0.02),
          dict(seed=73, extent=20.0, density=40.0, trees=900),
          dict(seed=74, extent=6.0, density=25.0, buildings=0, trees=0),
          dict(seed=75, extent=40.0, density=20.0, outliers=0.02, outlier_low=0.0)]


def build(ctx, cloud):
    center, side = W.Octree.size_fit(cloud["bbmin"], cloud["bbmax"])
    return W.Octree.build(ctx, cloud["xi"], cloud["yi"], cloud["zi"], cloud["scale"],
                          cloud["offset"], center, side)


def test_batch_equals_single_and_oracle(ctx):
    clouds = [synth.generate(**c) for c in CLOUDS]
    trees = [build(ctx, c) for c in clouds]
    try:
        got, stats = W.classify_batch(trees, [c["bbmin"] for c in clouds],
                                      [c["bbmax"] for c in clouds])
        for k, (c, t) in enumerate(zip(clouds, trees)):
            single, _, st = W.classify(t, c["bbmin"], c["bbmax"], want_tiles=False)
            assert np.array_equal(got[k], single), k
            for f in ("n_class1", "n_class2", "n_unclassified", "heavy_points",
                      "exact_gap_points", "ntiles", "nonempty_tiles"):
                assert stats[k][f] == st[f], (k, f)
        ref, _, _ = O.run(clouds[1])
        assert np.array_equal(got[1], ref)
        assert sum(s["n_class2"] for s in stats) > 0
    finally:
        for t in trees:
            t.free()


def test_batch_of_one_and_repeated(ctx):
    c = synth.generate(seed=76, extent=15.0, density=25.0)
    t = build(ctx, c)
    try:
        single, _, _ = W.classify(t, c["bbmin"], c["bbmax"], want_tiles=False)
        for _ in range(2):
            got, _ = W.classify_batch([t], [c["bbmin"]], [c["bbmax"]])
            assert np.array_equal(got[0], single)
        # the same tree twice in one batch
        got, _ = W.classify_batch([t, t], [c["bbmin"]] * 2, [c["bbmax"]] * 2)
        assert np.array_equal(got[0], single) and np.array_equal(got[1], single)
    finally:
        t.free()


def test_batch_refuses_trees_of_two_contexts(ctx):
    c = synth.generate(seed=77, extent=8.0, density=25.0)
    other = W.Context(0)
    t1, t2 = build(ctx, c), build(other, c)
    try:
        with pytest.raises(W.WbError, match="different contexts"):
            W.classify_batch([t1, t2], [c["bbmin"]] * 2, [c["bbmax"]] * 2)
    finally:
        t1.free()
        t2.free()
        other.close()
