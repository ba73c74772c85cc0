"""The serving loop's host logic on the CPU: size classes, slot sizing, Poisson
arrivals, batching order and error handling, with stand-in batch objects in
place of device batches (no GPU is touched)."""
from __future__ import annotations

import collections
import threading

import numpy as np
import pytest


class FakeStream:
    cuda_stream = 0


class FakeBatch:
    """Answers each slot with a vector filled with the query's first word.
    fail_poll: poll() raises from the k-th completed launch on (a device error)."""

    def __init__(self, fail_load: bool = False, fail_after: int | None = None, fail_poll: int | None = None):
        self.fail_load = fail_load
        self.fail_after = fail_after
        self.fail_poll = fail_poll
        self.loaded = []
        self.launches = []
        self.tuned = []
        self.forms = []
        self.asynchronous = []
        self.polled = 0
        self.pending_out = None

    def load_many(self, first, queries, keys, v_folding, asynchronous=False):
        if self.fail_load or (self.fail_after is not None and len(self.launches) >= self.fail_after):
            raise RuntimeError("injected load failure")
        assert first == 0 and len(queries) == len(keys) == len(v_folding)
        self.loaded = [int(q[0]) for q in queries]
        self.forms.append(("vectors", len(queries)))
        self.asynchronous.append(asynchronous)

    def load_packed(self, first, queries, keys, asynchronous=False):
        if self.fail_load:
            raise RuntimeError("injected load failure")
        assert first == 0 and len(queries) == len(keys)
        self.loaded = [int(q[0]) for q in queries]
        self.forms.append(("packed", len(queries)))
        self.asynchronous.append(asynchronous)

    def run(self, n, stages, stream=None):
        pass

    def tune(self, n):
        self.tuned.append(n)
        return {"chosen": "valu"}

    def set_dbmul(self, n, variant):
        self.tuned.append(("set", n, variant))

    def graph_launch(self, n, stream=None):
        self.launches.append((n, len(self.loaded)))

    def read_responses_async(self, n, out, stream=None):
        self.pending_out = (out, list(self.loaded[:n]))

    def synchronize(self):
        pass

    def poll(self):
        if self.fail_poll is not None and self.polled >= self.fail_poll:
            raise RuntimeError("injected device error")
        self.polled += 1
        if self.pending_out is not None:  # the queued copy lands when the run is seen done
            out, vals = self.pending_out
            for i, v in enumerate(vals):
                out[i] = v
            self.pending_out = None
        return True


def _server(batches_per_lane: dict[str, list[FakeBatch]], capacity: int = 4):
    from spiral_pir.serving import Server, _Lane

    srv = Server.__new__(Server)
    srv.lanes = {}
    for name, batches in batches_per_lane.items():
        ln = _Lane.__new__(_Lane)
        ln.name, ln.db, ln.capacity = name, None, capacity
        ln.batches = batches
        ln.out = [np.zeros((capacity, 4), np.uint32) for _ in batches]
        ln.streams = [FakeStream() for _ in batches]
        ln.free = collections.deque(range(len(batches)))
        ln.inflight, ln.pending, ln.ready, ln.choice = {}, collections.deque(), set(), {}
        ln.max_inflight = len(batches)
        ln.keys = {"c": object(), "p": type("PackedKeys", (), {"packed": True})()}
        ln.launched_batches = ln.launched_queries = 0
        srv.lanes[name] = ln
    srv.admission = {}
    srv.poll_interval = 1e-4
    srv.pack_q2_bits = 0
    srv.max_inflight_total = None
    srv._queue = collections.deque()
    srv._lock = threading.Lock()
    srv._wake = threading.Event()
    srv._stop = False
    srv._thread = None
    srv._error = None
    return srv


def _q(v):
    return np.array([v, 0, 0, 0], np.uint32)


def test_size_classes():
    from spiral_pir.serving import size_class

    assert [size_class(n, 128) for n in (1, 2, 3, 5, 9, 33, 100, 128)] == [1, 2, 3, 5, 10, 40, 112, 128]
    assert size_class(5, 6) == 5 and size_class(6, 6) == 6 and size_class(7, 6) == 6
    # classes past 8 are at most 1.25x apart, so a run wastes at most a fifth of its slots
    from spiral_pir.serving import SIZE_CLASSES

    big = [c for c in SIZE_CLASSES if c >= 8]
    assert all(b <= 1.25 * a for a, b in zip(big, big[1:]))


def test_slot_bytes_grow_with_shape():
    from spiral_pir import Shape
    from spiral_pir.serving import slot_device_bytes

    a, b = slot_device_bytes(Shape(9, 6)), slot_device_bytes(Shape(10, 7))
    assert 60 * 2**20 < a < b < 200 * 2**20


def test_poisson_rate():
    from spiral_pir.serving import poisson_times

    t = poisson_times(1000.0, 5.0, np.random.default_rng(3))
    assert abs(len(t) - 5000) < 300 and np.all(np.diff(t) > 0) and t[-1] < 5.0


def test_requests_are_batched_and_answered_in_order():
    srv = _server({"x": [FakeBatch(), FakeBatch()]}, capacity=4)
    reqs = [srv.submit("x", "c", _q(i), np.zeros(2, np.uint32), tag=i) for i in range(10)]
    srv.start()
    for r in reqs:
        assert int(r.wait(5)[0]) == r.tag
    srv.stop()
    ln = srv.lanes["x"]
    assert ln.launched_queries == 10 and ln.launched_batches == 3  # 4 + 4 + 2
    assert sorted(n for b in ln.batches for n, _ in b.launches) == [2, 4, 4]


def test_load_failure_fails_its_requests_only():
    bad, good = FakeBatch(fail_load=True), FakeBatch()
    srv = _server({"bad": [bad], "good": [good]}, capacity=8)
    r_bad = srv.submit("bad", "c", _q(1), np.zeros(2, np.uint32))
    r_good = srv.submit("good", "c", _q(2), np.zeros(2, np.uint32))
    srv.start()
    with pytest.raises(RuntimeError, match="injected"):
        r_bad.wait(5)
    assert int(r_good.wait(5)[0]) == 2
    # the lane recovers its buffer and keeps serving
    r_again = srv.submit("good", "c", _q(3), np.zeros(2, np.uint32))
    assert int(r_again.wait(5)[0]) == 3
    srv.stop()
    assert srv.lanes["bad"].free and not srv.lanes["bad"].inflight


def test_unknown_lane_or_client_is_rejected():
    srv = _server({"x": [FakeBatch()]})
    with pytest.raises(KeyError):
        srv.submit("y", "c", _q(0), np.zeros(2, np.uint32))
    with pytest.raises(KeyError):
        srv.submit("x", "nobody", _q(0), np.zeros(2, np.uint32))


def test_wait_times_out_when_never_answered():
    from spiral_pir.serving import Request

    r = Request("x", "c", _q(0), np.zeros(2, np.uint32))
    with pytest.raises(TimeoutError):
        r.wait(0.01)


def test_stop_drains_pending_requests():
    srv = _server({"x": [FakeBatch()]}, capacity=2)
    srv.start()
    reqs = [srv.submit("x", "c", _q(i), np.zeros(2, np.uint32), tag=i) for i in range(7)]
    srv.stop()
    assert all(r.done and int(r.response[0]) == r.tag for r in reqs)


def test_packed_and_vector_requests_batch_separately():
    """Consecutive requests of one form share a batch; a change of form starts a new one."""
    b = FakeBatch()
    srv = _server({"x": [b]}, capacity=8)
    forms = ["v", "v", "p", "p", "p", "v"]
    reqs = [srv.submit("x", "p" if f == "p" else "c", _q(i), None if f == "p" else np.zeros(2, np.uint32), tag=i)
            for i, f in enumerate(forms)]
    srv.start()
    for r in reqs:
        assert int(r.wait(5)[0]) == r.tag
    srv.stop()
    assert b.forms == [("vectors", 2), ("packed", 3), ("vectors", 1)]


def test_packed_request_needs_conversion_key():
    srv = _server({"x": [FakeBatch()]})
    with pytest.raises(ValueError, match="conversion key"):
        srv.submit("x", "c", _q(0), None)


def test_loads_are_queued_asynchronously():
    b = FakeBatch()
    srv = _server({"x": [b]}, capacity=4)
    srv.start()
    assert int(srv.submit("x", "c", _q(5), np.zeros(2, np.uint32)).wait(5)[0]) == 5
    srv.stop()
    assert b.asynchronous == [True]


def test_device_error_fails_every_request_and_the_server():
    """A readout that raises fails its own batch, every pending and queued request
    (none is left waiting forever), and later submissions raise."""
    bad = FakeBatch(fail_poll=0)
    srv = _server({"x": [bad]}, capacity=2)
    reqs = [srv.submit("x", "c", _q(i), np.zeros(2, np.uint32), tag=i) for i in range(5)]
    srv.start()
    for r in reqs:
        with pytest.raises(RuntimeError, match="injected device error"):
            r.wait(5)
    with pytest.raises(RuntimeError, match="injected device error"):
        srv.stop()
    with pytest.raises(RuntimeError, match="failed"):
        srv.submit("x", "c", _q(9), np.zeros(2, np.uint32))


def test_device_error_after_good_batches():
    """Batches answered before the error keep their responses; the rest fail."""
    b = FakeBatch(fail_poll=2)
    srv = _server({"x": [b]}, capacity=1)
    srv.start()
    first = [srv.submit("x", "c", _q(i), np.zeros(2, np.uint32), tag=i) for i in range(2)]
    for r in first:
        assert int(r.wait(5)[0]) == r.tag
    later = [srv.submit("x", "c", _q(10 + i), np.zeros(2, np.uint32)) for i in range(3)]
    for r in later:
        with pytest.raises(RuntimeError):
            r.wait(5)
    with pytest.raises(RuntimeError):
        srv.stop()


def test_latency_stats_split_queue_and_device_time():
    from spiral_pir.serving import Request, latency_stats

    reqs = []
    for k in range(4):
        r = Request("x", "c", _q(k), None)
        r.t_submit, r.t_launch, r.t_done, r.batch_size = 1.0 + k, 1.001 + k, 1.003 + k, 2
        reqs.append(r)
    st = latency_stats(reqs)
    assert abs(st["p50_ms"] - 3.0) < 1e-6 and abs(st["queue_ms"]["p50"] - 1.0) < 1e-6
    assert abs(st["device_ms"]["p99"] - 2.0) < 1e-6 and st["mean_batch"] == 2.0


def test_tuning_is_shared_between_buffers():
    """A size class is tuned once per lane; the other buffers get the same choice."""
    a, b = FakeBatch(), FakeBatch()
    srv = _server({"x": [a, b]}, capacity=4)
    ln = srv.lanes["x"]
    srv._prepare(ln, 0, 4)
    srv._prepare(ln, 1, 4)
    srv._prepare(ln, 1, 4)
    assert a.tuned == [4] and b.tuned == [("set", 4, "valu")]


def test_max_inflight_limits_concurrent_batches():
    b0, b1 = FakeBatch(), FakeBatch()
    srv = _server({"x": [b0, b1]}, capacity=2)
    srv.lanes["x"].max_inflight = 1
    ln = srv.lanes["x"]
    for i in range(5):
        ln.pending.append(__import__("spiral_pir.serving", fromlist=["Request"]).Request("x", "c", _q(i), np.zeros(2, np.uint32)))
    while ln.free and ln.pending and len(ln.inflight) < ln.max_inflight:
        srv._launch(ln)
    assert len(ln.inflight) == 1 and len(ln.pending) == 3


def test_global_limit_runs_oldest_lane_first():
    """With one batch on the GPU over all lanes, the lane holding the oldest
    waiting request launches first."""
    from spiral_pir.serving import Request

    a, b = FakeBatch(), FakeBatch()
    srv = _server({"a": [a], "b": [b]}, capacity=4)
    srv.max_inflight_total = 1
    old = Request("b", "c", _q(1), np.zeros(2, np.uint32))
    old.t_submit = 1.0
    new = Request("a", "c", _q(2), np.zeros(2, np.uint32))
    new.t_submit = 2.0
    srv.lanes["a"].pending.append(new)
    srv.lanes["b"].pending.append(old)
    srv._stop = True
    srv._thread = None
    srv.start()
    srv.stop()
    assert b.launches and a.launches
    assert old.t_launch <= new.t_launch
