"""The serving loop: tiles served while others are on the GPU give the same
class bytes as the same tiles served alone."""
import numpy as np
import pytest

import wolkenbase_amd as W
from wolkenbase_amd import serve, synth

pytestmark = pytest.mark.gpu


def test_served_results_equal_isolated():
    pool = [synth.generate(seed=500 + k, extent=e, density=d)
            for k, (e, d) in enumerate([(12.0, 25.0), (40.0, 10.0), (25.0, 40.0),
                                        (60.0, 8.0), (18.0, 60.0), (30.0, 25.0)])]
    res = serve.load_test(pool, slots=3, load=0.9, nreq=30, seed=7)
    print(res)
    assert res["mismatched_results"] == 0
    assert res["requests"] == 30


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
