"""Continuous-batching PIR server over databases resident in HBM.

Requests (a query ciphertext and its GSW fold vectors, or a packed query
whose fold vectors the device converts from the query itself, from a client
whose keys are resident on the device) arrive at any time.  A scheduler
thread keeps, per database ("lane"), a few batch buffers of `capacity` slots,
each with its own stream.  Whenever a buffer is free and requests are waiting,
it takes up to `capacity` of them (of one form: packed or with fold vectors)
and queues, without waiting on the host, their upload (one bulk call on the
buffer's io stream), the replay of the captured stage chain of the smallest
size class that holds them (which waits for the upload on the device), and
the copy of the responses into the buffer's pinned host array.  Requests
complete when the buffer's event reports that copy done.  Buffers of one lane
and of different lanes run on the device at the same time, so queries that
arrive while others are being answered are admitted as soon as a buffer frees
up, and their transfers and kernels overlap on the GPU.

With `pack_q2_bits` set, every response is packed and modulus-switched on the
device (Spiral's compact response) and requests complete with its bytes.

Admission is by free device memory: the slot capacity of each lane is sized at
start-up from the free HBM (hipMemGetInfo, after the int8-limb copies of the
databases that the MFMA multiply reads are built) and the device bytes one
slot needs; requests beyond the free slots wait in the lane's queue.

The reference has no serving layer (it answers one query from test files);
this module is the host-side loop over the library's batch, keys and graph
entry points.
"""
from __future__ import annotations

import collections
import gc
import threading
import time

import numpy as np

from .params import POLY_LEN, Shape
from .server import STAGE_EXPAND, Batch, Database, Keys

# Run sizes a batch is rounded up to (one captured graph each); the slots past
# the requests of a run are answered and discarded, so the classes are spaced
# by at most 1.25x past 8.
SIZE_CLASSES = (1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 56, 64, 80, 96, 112, 128)


def slot_device_bytes(shape: Shape) -> int:
    """Device bytes one batch slot holds: query, expanded v, copy of the client's
    keys, fold vectors, accumulators, upload and response staging, and its share
    of the int8-limb multiply workspaces (selector limbs and products of its two
    rows).  The database's own limb copy is counted once per database, not here."""
    n = POLY_LEN
    v = (1 << shape.g) * 4 * n * 4
    keys = (shape.left_words + shape.right_words) * 4
    fold = shape.fold_words * 4
    acc = shape.n2 * shape.cols * 4 * n * 4
    query = 4 * n * 4
    staging = (4 * n + shape.fold_words) * 4 + shape.response_words * 4
    mfma = 2 * n * shape.rows * 4 * 2 + 2 * n * shape.n2 * 2 * shape.cols * 4
    return v + keys + fold + acc + query + staging + mfma


def size_class(n: int, capacity: int) -> int:
    for c in SIZE_CLASSES:
        if c >= n:
            return min(c, capacity)
    return capacity


class Request:
    """One query in flight; `wait()` returns its n2 response ciphertexts (or, on a
    server that packs responses, the compact response bytes)."""

    __slots__ = ("lane", "client", "query", "v_folding", "tag", "t_submit", "t_launch", "t_done", "response",
                 "error", "batch_size", "_ev")

    def __init__(self, lane: str, client, query: np.ndarray, v_folding: np.ndarray, tag=None):
        self.lane = lane
        self.client = client
        self.query = query
        self.v_folding = v_folding
        self.tag = tag
        self.t_submit = time.perf_counter()
        self.t_launch = None
        self.t_done = None
        self.response = None
        self.error = None
        self.batch_size = 0
        self._ev = threading.Event()

    def fail(self, e: BaseException) -> None:
        if not self._ev.is_set():
            self.error = e
            self._ev.set()

    def wait(self, timeout: float | None = None) -> np.ndarray:
        if not self._ev.wait(timeout):
            raise TimeoutError("request not answered in time")
        if self.error is not None:
            raise self.error
        return self.response

    @property
    def done(self) -> bool:
        return self._ev.is_set()

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit

    @property
    def queue_time(self) -> float:
        """Submission to launch: waiting for a free batch buffer."""
        return self.t_launch - self.t_submit

    @property
    def device_time(self) -> float:
        """Launch to completion seen by the scheduler: upload, the stage chain, the
        response copy (and the batches ahead of it on the device)."""
        return self.t_done - self.t_launch


class _Lane:
    def __init__(self, name: str, db: Database, capacity: int, buffers: int, pack_q2_bits: int = 0):
        import torch

        self.name = name
        self.db = db
        self.capacity = capacity
        self.batches = [Batch(db, capacity) for _ in range(buffers)]
        if pack_q2_bits:
            for b in self.batches:
                b.set_packing(pack_q2_bits)
        # pinned host arrays the responses of each buffer are copied into
        if pack_q2_bits:
            nb = db.shape.packed_response_bytes(pack_q2_bits)
            self.out = [pinned_u8(capacity * nb).reshape(capacity, nb) for _ in range(buffers)]
        else:
            rw = db.shape.response_words
            self.out = [pinned_u32(capacity * rw).reshape(capacity, db.shape.n2, 2, 2, POLY_LEN)
                        for _ in range(buffers)]
        self.streams = [torch.cuda.Stream() for _ in range(buffers)]
        self.free = collections.deque(range(buffers))
        self.inflight: dict[int, list[Request]] = {}
        self.pending: collections.deque[Request] = collections.deque()
        self.keys: dict = {}
        self.ready: set = set()  # (buffer, size class) with its multiply variant set
        self.choice: dict = {}   # size class -> multiply variant, tuned on one buffer
        self.max_inflight = buffers
        self.launched_batches = 0
        self.launched_queries = 0

    def close(self) -> None:
        for b in self.batches:
            b.close()
        for k in self.keys.values():
            k.close()


class Server:
    """Continuous batching over resident databases `dbs` (name -> Database).

    capacity per lane = min(max_batch, share of `hbm_fraction` of the free HBM
    / (buffers x slot_device_bytes)); `weights` splits the memory between lanes
    (default: equally).  `max_inflight` batches per lane and
    `max_inflight_total` batches over all lanes run on the device at once
    (None: no limit); the defaults, one and one, run the batches one after
    the other, the lane whose oldest waiting request is oldest first.  Each
    next batch takes every request that arrived meanwhile, so under load
    batches grow instead of multiplying, and each pays its database scan once;
    and a batch running alone ends sooner than two sharing the GPU (serial
    service beats processor sharing for jobs of similar size).  At 2000
    queries/s over (9, 6) / (10, 7), 70/30, p50 / p99 were 4.7 / 9.9 ms with
    one batch on the GPU, 6.3 / 11.5 ms with one per lane and 11.7 / 21 ms
    with two per lane.  The second buffer of a lane still lets the uploads of
    its next batch be queued early."""

    def __init__(self, dbs: dict[str, Database], max_batch: int = 128, buffers: int = 2,
                 hbm_fraction: float = 0.8, weights: dict[str, float] | None = None,
                 poll_interval: float = 1e-4, pack_q2_bits: int = 0, max_inflight: int | None = 1,
                 max_inflight_total: int | None = 1):
        import torch

        for db in dbs.values():  # the MFMA multiply's limb copies count against the free HBM
            db.build_limbs()
        torch.cuda.synchronize()
        free, total = torch.cuda.mem_get_info()
        weights = weights or {name: 1.0 for name in dbs}
        wsum = sum(weights[name] for name in dbs)
        self.admission = {"free_bytes_at_start": int(free), "total_bytes": int(total), "hbm_fraction": hbm_fraction,
                          "lanes": {}}
        self.lanes: dict[str, _Lane] = {}
        for name, db in dbs.items():
            per_slot = slot_device_bytes(db.shape)
            share = free * hbm_fraction * weights[name] / wsum
            cap = min(max_batch, int(share // (buffers * per_slot)))
            if cap < 1:
                raise MemoryError(f"lane {name}: {share / 2**30:.1f} GiB of free HBM does not hold one slot "
                                  f"({per_slot / 2**20:.0f} MiB) per buffer")
            self.lanes[name] = _Lane(name, db, cap, buffers, pack_q2_bits)
            if max_inflight is not None:
                self.lanes[name].max_inflight = max(1, min(buffers, max_inflight))
            self.admission["lanes"][name] = {"capacity": cap, "buffers": buffers, "slot_bytes": per_slot}
        self.poll_interval = poll_interval
        self.pack_q2_bits = pack_q2_bits
        self.max_inflight_total = max_inflight_total
        self._queue: collections.deque[Request] = collections.deque()
        self._lock = threading.Lock()
        self._wake = threading.Event()
        self._stop = False
        self._thread: threading.Thread | None = None
        self._error: BaseException | None = None

    # ---- set-up ---------------------------------------------------------------
    def register(self, lane: str, client, left: np.ndarray, right: np.ndarray, conv: np.ndarray | None = None,
                 packing: np.ndarray | None = None) -> None:
        """Make a client's expansion keys (and conversion key, for packed queries,
        and packing keys, for compact responses) resident for `lane`."""
        ln = self.lanes[lane]
        if self.pack_q2_bits and packing is None:
            raise ValueError("this server packs responses: register the client's packing keys")
        ln.keys[client] = Keys(ln.db, left, right, conv, packing)

    def warmup(self, lane: str, client, query: np.ndarray, v_folding: np.ndarray | None = None) -> None:
        """Fill every buffer of `lane` with copies of one query (packed when
        v_folding is None) and autotune and capture every size class of that
        form, so that serving starts without first-use costs."""
        ln = self.lanes[lane]
        keys = ln.keys[client]
        for buf, batch in enumerate(ln.batches):
            stream = ln.streams[buf].cuda_stream
            if v_folding is None:
                batch.load_packed(0, [query] * ln.capacity, [keys] * ln.capacity)
            else:
                batch.load_many(0, [query] * ln.capacity, [keys] * ln.capacity, [v_folding] * ln.capacity)
            for c in sorted({size_class(n, ln.capacity) for n in range(1, ln.capacity + 1)}):
                self._prepare(ln, buf, c)
                batch.graph_launch(c, stream)
                batch.synchronize()

    # ---- requests -------------------------------------------------------------
    def submit(self, lane: str, client, query: np.ndarray, v_folding: np.ndarray | None = None, tag=None) -> Request:
        """Queue one query; v_folding None means a packed query (the client's
        keys must carry a conversion key)."""
        if self._error is not None:
            raise RuntimeError("the serving loop has failed") from self._error
        if self._thread is not None and not self._thread.is_alive():
            raise RuntimeError("the serving loop is not running")
        if lane not in self.lanes:
            raise KeyError(f"no database lane {lane!r}")
        if client not in self.lanes[lane].keys:
            raise KeyError(f"client {client!r} has no resident keys for lane {lane!r}")
        if v_folding is None and not getattr(self.lanes[lane].keys[client], "packed", False):
            raise ValueError(f"client {client!r} has no conversion key for packed queries")
        r = Request(lane, client, query, v_folding, tag)
        with self._lock:
            self._queue.append(r)
        self._wake.set()
        return r

    # ---- the loop -------------------------------------------------------------
    def start(self) -> None:
        """Start the scheduler thread.  Everything allocated so far (torch, the
        databases, the keys, the warm-up) is moved out of the garbage
        collector's reach first (gc.freeze): a full collection over a process
        with torch loaded walks about a million objects and stopped the
        scheduler for 40-55 ms at a time, which showed as the latency tail."""
        if self._thread is not None:
            raise RuntimeError("server already running")
        gc.collect()
        gc.freeze()
        self._stop = False
        self._thread = threading.Thread(target=self._run, name="pir-serving", daemon=True)
        self._thread.start()

    def stop(self) -> None:
        """Answer everything submitted so far, then stop the loop."""
        self._stop = True
        self._wake.set()
        if self._thread is not None:
            self._thread.join()
            self._thread = None
        if self._error is not None:
            raise self._error

    def close(self) -> None:
        if self._thread is not None:
            self.stop()
        for ln in self.lanes.values():
            ln.close()

    def _run(self) -> None:
        try:
            while True:
                with self._lock:
                    arrivals = list(self._queue)
                    self._queue.clear()
                for r in arrivals:
                    self.lanes[r.lane].pending.append(r)
                progressed = bool(arrivals)
                busy = False
                for ln in self.lanes.values():
                    progressed |= self._complete(ln)
                if self.max_inflight_total is None:
                    for ln in self.lanes.values():
                        while ln.free and ln.pending and len(ln.inflight) < ln.max_inflight:
                            self._launch(ln)
                            progressed = True
                else:
                    # GPU-wide limit: the lane whose oldest waiting request is oldest goes first
                    while sum(len(ln.inflight) for ln in self.lanes.values()) < self.max_inflight_total:
                        ready = [ln for ln in self.lanes.values()
                                 if ln.free and ln.pending and len(ln.inflight) < ln.max_inflight]
                        if not ready:
                            break
                        self._launch(min(ready, key=lambda ln: ln.pending[0].t_submit))
                        progressed = True
                for ln in self.lanes.values():
                    busy |= bool(ln.inflight or ln.pending)
                if self._stop and not busy:
                    with self._lock:
                        if not self._queue:
                            return
                if not progressed:
                    self._wake.wait(self.poll_interval)
                    self._wake.clear()
        except BaseException as e:  # fail every request not answered, surface the error on stop()
            self._error = e
            with self._lock:
                queued = list(self._queue)
                self._queue.clear()
            for ln in self.lanes.values():
                for reqs in ln.inflight.values():
                    for r in reqs:
                        r.fail(e)
                ln.inflight.clear()
                for r in ln.pending:
                    r.fail(e)
                ln.pending.clear()
            for r in queued:
                r.fail(e)

    def _prepare(self, ln: _Lane, buf: int, c: int) -> None:
        """Multiply variant of size class c on buffer buf: tuned on the first
        buffer that needs it, set on the others (same database, same choice)."""
        if (buf, c) in ln.ready:
            return
        batch = ln.batches[buf]
        if c in ln.choice:
            batch.set_dbmul(c, ln.choice[c])
        else:
            batch.run(c, STAGE_EXPAND, ln.streams[buf].cuda_stream)
            ln.choice[c] = batch.tune(c)["chosen"]
        ln.ready.add((buf, c))

    def _launch(self, ln: _Lane) -> None:
        """Queue one batch on a free buffer: upload, stage chain, response copy.
        Nothing here waits for the device."""
        buf = ln.free.popleft()
        packed = ln.pending[0].v_folding is None
        reqs = []
        while ln.pending and len(reqs) < ln.capacity and (ln.pending[0].v_folding is None) == packed:
            reqs.append(ln.pending.popleft())
        n = len(reqs)
        batch = ln.batches[buf]
        stream = ln.streams[buf].cuda_stream
        t = time.perf_counter()
        try:
            if packed:
                batch.load_packed(0, [r.query for r in reqs], [ln.keys[r.client] for r in reqs], asynchronous=True)
            else:
                batch.load_many(0, [r.query for r in reqs], [ln.keys[r.client] for r in reqs],
                                [r.v_folding for r in reqs], asynchronous=True)
            c = size_class(n, ln.capacity)
            self._prepare(ln, buf, c)
            batch.graph_launch(c, stream)
            if self.pack_q2_bits:
                batch.read_packed_async(n, ln.out[buf], stream)
            else:
                batch.read_responses_async(n, ln.out[buf], stream)
        except Exception as e:
            for r in reqs:
                r.fail(e)
            batch.synchronize()  # whatever was queued ends before the buffer is reused (raises if the device failed)
            ln.free.append(buf)
            return
        for r in reqs:
            r.t_launch = t
            r.batch_size = n
        ln.inflight[buf] = reqs
        ln.launched_batches += 1
        ln.launched_queries += n

    def _complete(self, ln: _Lane) -> bool:
        done_any = False
        for buf in list(ln.inflight):
            batch = ln.batches[buf]
            reqs = ln.inflight[buf]
            try:
                if not batch.poll():
                    continue
                del ln.inflight[buf]
                out = ln.out[buf]
                t = time.perf_counter()
                for i, r in enumerate(reqs):
                    r.response = out[i].copy()
                    r.t_done = t
                    r._ev.set()
            except BaseException as e:  # a device error: this batch fails, then the loop
                ln.inflight.pop(buf, None)
                for r in reqs:
                    r.fail(e)
                raise
            ln.free.append(buf)
            done_any = True
        return done_any


def pinned_u32(n_words: int) -> np.ndarray:
    """A host array in pinned memory (asynchronous DMA for bulk loads)."""
    import torch

    return torch.empty(n_words, dtype=torch.int32, pin_memory=True).numpy().view(np.uint32)


def pinned_u8(n_bytes: int) -> np.ndarray:
    """A host byte array in pinned memory (asynchronous DMA of compact responses)."""
    import torch

    return torch.empty(n_bytes, dtype=torch.uint8, pin_memory=True).numpy()


def poisson_times(rate: float, duration: float, rng: np.random.Generator) -> np.ndarray:
    """Arrival offsets (s) of a Poisson process of `rate` per second over `duration` s."""
    n = max(1, int(rate * duration * 1.5) + 10)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    return t[t < duration]


def drive(server: Server, arrivals: list[tuple[float, str, object, np.ndarray, np.ndarray, object]]) -> list[Request]:
    """Submit (offset_s, lane, client, query, v_folding, tag) at their offsets from now
    and return the requests.  The wait is a sleep, never a spin: a spinning
    Python thread holds the interpreter lock, and the scheduler thread would
    then wait up to a switch interval (5 ms) after each library call to run
    again.  Requests are submitted at most ~0.1 ms late (sleep granularity);
    their latency is counted from the actual submission."""
    t0 = time.perf_counter()
    out = []
    for off, lane, client, q, vf, tag in arrivals:
        dt = t0 + off - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        out.append(server.submit(lane, client, q, vf, tag))
    return out


def latency_stats(reqs: list[Request]) -> dict:
    """Latency percentiles (submission to completion), split into queue time
    (submission to launch) and device time (launch to completion), the mean
    batch size, and the sustained rate over the span of the requests."""
    lat = np.array([r.latency for r in reqs]) * 1e3
    que = np.array([r.queue_time for r in reqs]) * 1e3
    dev = np.array([r.device_time for r in reqs]) * 1e3
    t_first = min(r.t_submit for r in reqs)
    t_last = max(r.t_done for r in reqs)
    return {"completed": len(reqs), "sustained_qps": len(reqs) / (t_last - t_first),
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()), "max_ms": float(lat.max()),
            "queue_ms": {"p50": float(np.percentile(que, 50)), "p99": float(np.percentile(que, 99))},
            "device_ms": {"p50": float(np.percentile(dev, 50)), "p99": float(np.percentile(dev, 99))},
            "mean_batch": float(np.mean([r.batch_size for r in reqs]))}
