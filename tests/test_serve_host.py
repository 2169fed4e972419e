"""Serving loop failure handling on the host: no request may wait forever
when a device context cannot be opened or a worker dies (CPU only; the
device calls are replaced by fakes)."""
import threading

import numpy as np
import pytest

from wolkenbase_amd import serve


class _Ctx:
    opened = 0

    def __init__(self, device=0):
        _Ctx.opened += 1

    def close(self):
        _Ctx.opened -= 1


def _tile(n=5):
    return {"xi": np.zeros(n, np.int32)}


@pytest.fixture
def fake_device(monkeypatch):
    monkeypatch.setattr(serve.W, "device_memory", lambda device=0: (1 << 30, 1 << 30))
    monkeypatch.setattr(serve.W, "Context", _Ctx)
    _Ctx.opened = 0
    yield
    assert _Ctx.opened == 0


def test_context_failure_raises_in_constructor(monkeypatch, fake_device):
    made = []

    class Bad(_Ctx):
        def __init__(self, device=0):
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

    def classify_tile(ctx, tile, params):
        gate.wait(5)
        raise Fatal("worker killed")

    monkeypatch.setattr(serve, "classify_tile", classify_tile)
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

    def classify_tile(ctx, tile, params):
        calls.append(1)
        if len(calls) == 1:
            raise ValueError("bad tile")
        return np.ones(len(tile["xi"]), np.uint8)

    monkeypatch.setattr(serve, "classify_tile", classify_tile)
    srv = serve.Server(slots=1, params=object())
    a = srv.submit(serve.Request(_tile()))
    b = srv.submit(serve.Request(_tile()))
    assert a.done.wait(5) and b.done.wait(5)
    assert isinstance(a.error, ValueError) and b.error is None and b.cls.sum() == 5
    srv.close()
