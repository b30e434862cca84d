"""Host logic of the continuous-batching server (CPU only): a batch that
fails on a worker ends run() with that error instead of leaving it waiting for
latencies that never come, and the admission budget is released."""
import contextlib
import threading

import numpy as np
import pytest
import torch

from differentiable_gaussian_mixture_renderer_amd import serving
from differentiable_gaussian_mixture_renderer_amd._native import DgmrError


class _FakeStream:
    def synchronize(self):
        pass


@pytest.fixture
def host_only(monkeypatch):
    """Streams and the batched splatter replaced by host stand-ins."""
    monkeypatch.setattr(serving, "_worker_stream", lambda device, k: _FakeStream())
    monkeypatch.setattr(serving, "_inline_stream", lambda device: _FakeStream())
    monkeypatch.setattr(serving.torch.cuda, "stream", lambda s: contextlib.nullcontext())
    monkeypatch.setattr(serving, "BatchSplatter", lambda **kw: None)


def _server(n_requests=12):
    resident = [serving.ResidentScene([torch.zeros(10, 1, 3), torch.zeros(10, 1), torch.zeros(10, 3), torch.zeros(10, 3),
                                       torch.zeros(10, 4)], 0)]
    trace = serving.make_trace(n_requests, rate_hz=2000.0, n_scenes=1, seed=1, min_res=(16, 16), max_res=(32, 32))
    return serving.ContinuousBatchingServer(resident, hbm_budget_bytes=1 << 30, workers=2, device="cpu"), trace


def test_worker_failure_ends_run_with_the_error(host_only, monkeypatch):
    # the first dispatch fails, whichever path (one request alone or a batch)
    # it takes: how many dispatches the trace makes, and down which path,
    # depends on thread timing (under load every request may have arrived
    # before the first is scheduled, and all of them form one batch)
    calls = {"n": 0}
    lock = threading.Lock()

    def first_fails():
        with lock:
            calls["n"] += 1
            k = calls["n"]
        if k == 1:
            raise DgmrError(2, "dgmr_splat_forward", "pair capacity kept growing")

    def flaky_single(self, sc, r):
        first_fails()
        return [None], [None]

    def flaky_batched(sp, sc, batch, n_bwd):
        first_fails()
        return ([None] * len(batch),) * 2

    monkeypatch.setattr(serving.ContinuousBatchingServer, "_single", flaky_single)
    monkeypatch.setattr(serving.ContinuousBatchingServer, "_batched", staticmethod(flaky_batched))
    srv, trace = _server()
    done = {}

    def go():
        try:
            srv.run(trace)
            done["result"] = "returned"
        except RuntimeError as exc:
            done["result"] = exc

    th = threading.Thread(target=go, daemon=True)
    th.start()
    th.join(timeout=30.0)
    assert not th.is_alive(), "run() kept waiting after a worker failed"
    assert isinstance(done["result"], RuntimeError)
    assert isinstance(done["result"].__cause__, DgmrError)
    assert srv._in_flight == 0


def test_all_batches_complete_without_failures(host_only, monkeypatch):
    monkeypatch.setattr(serving.ContinuousBatchingServer, "_single", lambda self, sc, r: ([None], [None]))
    monkeypatch.setattr(serving.ContinuousBatchingServer, "_batched", staticmethod(lambda sp, sc, batch, n_bwd: ([None] * len(batch),) * 2))
    srv, trace = _server()
    stats = srv.run(trace)
    assert len(stats.latencies) == len(trace)
    assert srv._in_flight == 0
    assert np.all(np.array(stats.latencies) >= 0)


def test_idle_server_runs_a_lone_request_without_a_hand_off(host_only, monkeypatch):
    """Spaced arrivals (each finds the server idle) run on the scheduling
    thread itself; a burst over several scenes goes to the worker threads."""
    threads = []

    def single(self, sc, r):
        threads.append(threading.current_thread())
        return [None], [None]

    def batched(sp, sc, batch, n_bwd):
        threads.append(threading.current_thread())
        return ([None] * len(batch),) * 2

    monkeypatch.setattr(serving.ContinuousBatchingServer, "_single", single)
    monkeypatch.setattr(serving.ContinuousBatchingServer, "_batched", staticmethod(batched))
    resident = [serving.ResidentScene([torch.zeros(10, 1, 3), torch.zeros(10, 1), torch.zeros(10, 3), torch.zeros(10, 3),
                                       torch.zeros(10, 4)], 0) for _ in range(3)]
    spaced = serving.make_trace(6, rate_hz=50.0, n_scenes=3, seed=2, min_res=(16, 16), max_res=(32, 32))
    srv = serving.ContinuousBatchingServer(resident, hbm_budget_bytes=1 << 30, workers=2, device="cpu")
    srv.run(spaced)
    assert threads and all(t is threading.current_thread() for t in threads)
    threads.clear()
    burst = serving.make_trace(24, rate_hz=50.0, n_scenes=3, seed=3, min_res=(16, 16), max_res=(32, 32))
    for r in burst:
        r.arrival = 0.0
    srv = serving.ContinuousBatchingServer(resident, hbm_budget_bytes=1 << 30, workers=2, device="cpu")
    stats = srv.run(burst)
    assert len(stats.latencies) == len(burst)
    assert any(t is not threading.current_thread() for t in threads)


def test_batches_overlap_only_with_a_backlog(host_only, monkeypatch):
    """Below overlap_backlog waiting requests one batch runs at a time (a
    request does not share the GPU with another at low load); a burst over
    several scenes fills every worker."""
    import time
    state = {"running": 0, "most": 0}
    lock = threading.Lock()

    def work(n):
        with lock:
            state["running"] += 1
            state["most"] = max(state["most"], state["running"])
        time.sleep(0.004)
        with lock:
            state["running"] -= 1
        return ([None] * n,) * 2

    monkeypatch.setattr(serving.ContinuousBatchingServer, "_single", lambda self, sc, r: work(1))
    monkeypatch.setattr(serving.ContinuousBatchingServer, "_batched", staticmethod(lambda sp, sc, batch, n_bwd: work(len(batch))))
    resident = [serving.ResidentScene([torch.zeros(10, 1, 3), torch.zeros(10, 1), torch.zeros(10, 3), torch.zeros(10, 3),
                                       torch.zeros(10, 4)], 0) for _ in range(4)]
    # pairs of requests of different scenes 1 ms apart, every 30 ms: the second waits for the first
    trace = serving.make_trace(10, rate_hz=50.0, n_scenes=4, seed=4, min_res=(16, 16), max_res=(32, 32))
    for k, r in enumerate(trace):
        r.scene = k % 2
        r.arrival = 0.03 * (k // 2) + 0.001 * (k % 2)
    srv = serving.ContinuousBatchingServer(resident, hbm_budget_bytes=1 << 30, workers=3, device="cpu", overlap_backlog=3)
    stats = srv.run(trace)
    assert len(stats.latencies) == len(trace) and state["most"] == 1
    burst = serving.make_trace(40, rate_hz=50.0, n_scenes=4, seed=5, min_res=(16, 16), max_res=(32, 32))
    for r in burst:
        r.arrival = 0.0
    state["most"] = 0
    srv = serving.ContinuousBatchingServer(resident, hbm_budget_bytes=1 << 30, workers=3, device="cpu", overlap_backlog=3)
    stats = srv.run(burst)
    assert len(stats.latencies) == len(burst) and state["most"] == 3
    assert srv._batches == 0 and srv._in_flight == 0


def test_resident_sizes_span_the_range():
    sizes = serving.resident_sizes(8)
    assert sizes[0] == 50_000 and sizes[-1] == 3_000_000
    assert all(a < b for a, b in zip(sizes, sizes[1:]))
    ratios = [b / a for a, b in zip(sizes, sizes[1:])]
    assert max(ratios) / min(ratios) < 1.001   # evenly spaced in log
