"""The serving loop: tiles served while others are on the GPU, alone or in
batches of small tiles, from pinned or pageable host buffers, give the same
class bytes as the same tiles classified alone."""
import numpy as np
import pytest

import wolkenbase_amd as W
from wolkenbase_amd import serve, synth

pytestmark = pytest.mark.gpu


def _pool():
    return [synth.generate(seed=500 + k, extent=e, density=d)
            for k, (e, d) in enumerate([(12.0, 25.0), (40.0, 10.0), (25.0, 40.0),
                                        (60.0, 8.0), (18.0, 60.0), (30.0, 25.0)])]


def test_served_results_equal_isolated():
    pool = _pool()
    # small_points above every pool tile: requests queued together are batched
    res = serve.load_test(serve.pin_tiles(pool), slots=3, load=0.9, nreq=30, seed=7,
                          small_points=10_000_000, batch_points=200_000)
    print(res)
    assert res["mismatched_results"] == 0
    assert res["requests"] == 30


def test_batch_of_tiles_equals_each_alone():
    pool = _pool()
    pinned = serve.pin_tiles(pool)
    ctx = W.Context(0)
    try:
        params = W.default_params()
        alone = [serve.classify_tile(ctx, t, params) for t in pool]
        for idx, src in (([0, 1, 2, 3, 4, 5], pool), ([5, 0, 3], pinned), ([2], pool)):
            got = serve.classify_tiles(ctx, [src[k] for k in idx], params)
            assert len(got) == len(idx)
            for g, k in zip(got, idx):
                assert np.array_equal(g, alone[k]), f"tile {k} in a batch of {len(idx)}"
    finally:
        ctx.close()


def test_pinned_input_equals_pageable():
    pool = _pool()[:3]
    pinned = serve.pin_tiles(pool)
    ctx = W.Context(0)
    try:
        params = W.default_params()
        for t, p in zip(pool, pinned):
            a = serve.classify_tile(ctx, t, params)
            b = serve.classify_tile(ctx, p, params)
            assert np.array_equal(a, b)
            assert set(np.unique(a)) <= {1, 2}
    finally:
        ctx.close()


def test_admission_waits_for_memory():
    # a budget that fits one request at a time still serves every request
    pool = [synth.generate(seed=600 + k, extent=15.0, density=25.0) for k in range(3)]
    srv = serve.Server(slots=3, params=W.default_params())
    srv.budget = max(serve.Request(t).need for t in pool)
    try:
        reqs = [srv.submit(serve.Request(t, k)) for k, t in enumerate(pool)]
        for r in reqs:
            assert r.done.wait(120)
    finally:
        srv.close()
    ctx = W.Context(0)
    for r in reqs:
        assert r.error is None
        assert np.array_equal(r.cls, serve.classify_tile(ctx, pool[r.tag], W.default_params()))
    ctx.close()


def test_pinned_buffers():
    a = W.pinned_empty(1000, np.int32)
    assert a.dtype == np.int32 and a.shape == (1000,) and a.flags.writeable
    a[:] = np.arange(1000)
    b = a[10:20]
    del a
    assert b.sum() == sum(range(10, 20))
    with pytest.raises(W.WbError):
        W.pinned_empty(1 << 62, np.uint8)
