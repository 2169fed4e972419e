"""Serving loop on the host: failure handling (no request may wait forever
when a device context cannot be opened or a worker dies), small-tile
batching with per-request errors, the scheduling order, and the serving
stream's point-count strata (CPU only; the device calls are replaced by
fakes)."""
import math
import statistics
import threading
import time

import numpy as np
import pytest

from wolkenbase_amd import serve


class _Ctx:
    opened = 0

    def __init__(self, device=0, sharing=False):
        _Ctx.opened += 1

    def close(self):
        _Ctx.opened -= 1


def _tile(n=5, bad=False):
    t = {"xi": np.zeros(n, np.int32)}
    if bad:
        t["bad"] = True
    return t


@pytest.fixture
def fake_device(monkeypatch):
    monkeypatch.setattr(serve.W, "device_memory", lambda device=0: (1 << 30, 1 << 30))
    monkeypatch.setattr(serve.W, "Context", _Ctx)
    _Ctx.opened = 0
    yield
    assert _Ctx.opened == 0


def _fake_classify(calls):
    def classify_tile(ctx, tile, params, times=None):
        calls.append(("one", len(tile["xi"])))
        if tile.get("bad"):
            raise ValueError("bad tile")
        return np.ones(len(tile["xi"]), np.uint8)

    def classify_tiles(ctx, tiles, params):
        calls.append(("batch", len(tiles)))
        if any(t.get("bad") for t in tiles):
            raise ValueError("bad tile in batch")
        return [np.full(len(t["xi"]), 2, np.uint8) for t in tiles]

    return classify_tile, classify_tiles


def test_context_failure_raises_in_constructor(monkeypatch, fake_device):
    made = []

    class Bad(_Ctx):
        def __init__(self, device=0, sharing=False):
            if len(made) == 2:
                raise RuntimeError("device lost")
            made.append(1)
            super().__init__(device)

    monkeypatch.setattr(serve.W, "Context", Bad)
    with pytest.raises(RuntimeError, match="device lost"):
        serve.Server(slots=4)


def test_dead_worker_fails_pending_and_later_requests(monkeypatch, fake_device):
    gate = threading.Event()

    class Fatal(BaseException):
        pass

    def classify_tile(ctx, tile, params, times=None):
        gate.wait(5)
        raise Fatal("worker killed")

    def classify_tiles(ctx, tiles, params):
        gate.wait(5)
        raise Fatal("worker killed")

    monkeypatch.setattr(serve, "classify_tile", classify_tile)
    monkeypatch.setattr(serve, "classify_tiles", classify_tiles)
    srv = serve.Server(slots=1, params=object())
    first = srv.submit(serve.Request(_tile()))
    queued = [srv.submit(serve.Request(_tile())) for _ in range(3)]
    gate.set()
    for r in [first] + queued:
        assert r.done.wait(5), "a request was left waiting"
        assert r.error is not None
    late = srv.submit(serve.Request(_tile()))
    assert late.done.wait(1) and isinstance(late.error, RuntimeError)
    srv.close()


def test_request_error_keeps_worker_serving(monkeypatch, fake_device):
    calls = []
    one, batch = _fake_classify(calls)
    monkeypatch.setattr(serve, "classify_tile", one)
    monkeypatch.setattr(serve, "classify_tiles", batch)
    srv = serve.Server(slots=1, params=object())
    a = srv.submit(serve.Request(_tile(bad=True)))
    b = srv.submit(serve.Request(_tile()))
    assert a.done.wait(5) and b.done.wait(5)
    assert isinstance(a.error, ValueError) and b.error is None and len(b.cls) == 5
    srv.close()


def _hold_worker(monkeypatch, calls):
    """Fakes whose first call blocks until the returned event is set, so
    requests queue up behind it."""
    gate = threading.Event()
    one, batch = _fake_classify(calls)

    def one_gated(ctx, tile, params, times=None):
        if len(calls) == 0:
            calls.append(("hold", len(tile["xi"])))
            gate.wait(5)
            return np.ones(len(tile["xi"]), np.uint8)
        return one(ctx, tile, params, times)

    monkeypatch.setattr(serve, "classify_tile", one_gated)
    monkeypatch.setattr(serve, "classify_tiles", batch)
    return gate


def _wait_started(calls):
    for _ in range(500):
        if calls:
            return
        time.sleep(0.01)
    raise AssertionError("worker did not start")


def test_small_requests_batched_and_bad_batch_split(monkeypatch, fake_device):
    calls = []
    gate = _hold_worker(monkeypatch, calls)
    srv = serve.Server(slots=1, params=object(), small_points=100, batch_points=1000)
    hold = srv.submit(serve.Request(_tile(50)))
    _wait_started(calls)
    good = [srv.submit(serve.Request(_tile(10 + k))) for k in range(4)]
    gate.set()
    for r in [hold] + good:
        assert r.done.wait(5) and r.error is None
    assert ("batch", 4) in calls
    assert all(r.batch == 4 and (r.cls == 2).all() for r in good)
    # a batch holding a bad tile is retried tile by tile: only it fails
    calls.clear()
    gate = _hold_worker(monkeypatch, calls)
    hold = srv.submit(serve.Request(_tile(50)))
    _wait_started(calls)
    mixed = [srv.submit(serve.Request(_tile(10, bad=(k == 1)))) for k in range(3)]
    gate.set()
    for r in [hold] + mixed:
        assert r.done.wait(5)
    assert ("batch", 3) in calls
    assert [r.error is None for r in mixed] == [True, False, True]
    assert isinstance(mixed[1].error, ValueError)
    srv.close()


def test_batch_limits_and_large_requests_alone(monkeypatch, fake_device):
    calls = []
    gate = _hold_worker(monkeypatch, calls)
    srv = serve.Server(slots=1, params=object(), small_points=100, batch_points=250,
                       policy="fifo")
    hold = srv.submit(serve.Request(_tile(50)))
    _wait_started(calls)
    reqs = [srv.submit(serve.Request(_tile(n))) for n in (100, 100, 100, 500, 100)]
    gate.set()
    for r in [hold] + reqs:
        assert r.done.wait(5) and r.error is None
    # fifo: the first two small ones fit 250 points, the third waits; the
    # large one runs alone
    assert calls[1] == ("batch", 2)
    assert reqs[3].batch == 1
    srv.close()


def test_ratio_policy_prefers_small_waiting_requests(monkeypatch, fake_device):
    srv = serve.Server(slots=1, params=object())
    srv.close()
    big = serve.Request(_tile(10))
    big.n, big.est = 30_000_000, serve.service_estimate(30_000_000)
    small = serve.Request(_tile(10))
    small.n, small.est = 100_000, serve.service_estimate(100_000)
    now = 100.0
    big.t_arrive, small.t_arrive = now - 0.010, now - 0.005
    srv.q.extend([big, small])
    assert srv._order([0, 1], now) == [1, 0]   # the small one's ratio is higher
    srv.policy = "fifo"
    assert srv._order([0, 1], now) == [0, 1]
    # a large request that waited long enough comes first
    srv.policy = "ratio"
    big.t_arrive = now - 10.0
    assert srv._order([0, 1], now) == [0, 1]


def test_small_only_workers_leave_large_requests(monkeypatch, fake_device):
    srv = serve.Server(slots=2, small_slots=1, params=object(), small_points=100)
    srv.close()  # the workers are gone; _next is called directly below
    srv.stopping = False
    big = serve.Request(_tile(500))
    big.t_arrive = time.perf_counter()
    srv.q.append(big)
    # a small-only worker finds nothing to take (it would wait; with the
    # server stopping and the request left to others it returns nothing)
    srv.stopping = True
    srv.q.append(serve.Request(_tile(50)))
    srv.q[-1].t_arrive = time.perf_counter()
    got = srv._next(small_only=True)
    assert [r.n for r in got] == [50]
    got = srv._next(small_only=False)
    assert [r.n for r in got] == [500]
    srv.used = 0


def test_stream_strata_follow_the_lognormal():
    specs = serve.stream_specs(200, seed=3000, max_points=4e7)
    n = np.array([s["extent"] ** 2 * s["density"] for s in specs])
    nd = statistics.NormalDist(math.log(2e6), 1.0)
    # quantile k of the lognormal, within rounding
    for k in (0, 50, 100, 150, 199):
        assert n[k] == pytest.approx(math.exp(nd.inv_cdf((k + 0.5) / 200)), rel=1e-9)
    assert np.all(np.diff(n) > 0)
    assert np.median(n) == pytest.approx(2e6, rel=0.02)
    dens = np.array([s["density"] for s in specs])
    assert dens.min() >= 5 and dens.max() <= 60
    assert len({s["seed"] for s in specs}) == 200
    # clipping at both ends
    specs = serve.stream_specs(10000, seed=1, max_points=1e7)
    n = np.array([s["extent"] ** 2 * s["density"] for s in specs])
    assert n.min() == pytest.approx(5e4, rel=1e-9) and n.max() == pytest.approx(1e7, rel=1e-9)
