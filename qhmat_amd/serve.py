"""Serving loop: resident H-matrices, HBM-budget admission, continuous
batching of apply requests, p50/p99 latency.

* Stores are admitted against an HBM budget (their resident device bytes);
  a store that does not fit raises AdmissionError.
* Requests (store name, FP64 right-hand sides n x k) are queued.  The
  dispatcher takes whatever is queued when a pass slot is free, merges the
  requests that target the same store into one k-wide apply (requests whose
  operands would overflow the budget wait for memory to free up), and
  launches it on one of several streams, so passes overlap on the device and
  requests that arrive meanwhile join the next pass.
* A completion thread waits for each pass and hands every request its own
  columns of the result.

A right-hand-side column's arithmetic does not depend on which other columns
share its pass, only on the pass's arithmetic path (vector ALUs for narrow
passes, matrix cores for 16-wide ones): a request's result is bitwise the same
alone and in a mix that keeps it on its path, and within the precision's
tolerance otherwise (tests/test_gpu_serve.py).
"""

from __future__ import annotations

import collections
import queue
import sys
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from .device import apply as device_apply
from .device import precision_code
from .device import workspace_bytes


class AdmissionError(RuntimeError):
    """The store does not fit in the serving HBM budget."""


@dataclass
class _Request:
    store: str
    x: object  # torch tensor (n, k) on the device
    ready: object  # torch.cuda.Event recorded after x was written
    k: int
    arrival: float
    finite: object = None  # device bool for a device-tensor x, checked on completion
    future: Future = field(default_factory=Future)


@dataclass
class _Resident:
    native: object
    order: int
    device_bytes: int


class Server:
    """Continuous-batching apply server on one device."""

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, precision: str = "fp64",
                 slots: int = 3, max_rhs: int = 256):
        import torch

        self.torch = torch
        self.device = device
        self.budget = int(hbm_budget)
        self.precision = precision
        self.max_rhs = int(max_rhs)
        self._stores: dict[str, _Resident] = {}
        self._resident_bytes = 0
        self._inflight_bytes = 0
        self._queue: collections.deque[_Request] = collections.deque()
        self._cv = threading.Condition()
        self._slots = threading.Semaphore(slots)
        self._streams = [torch.cuda.Stream(device=device) for _ in range(slots)]
        self._next_stream = 0
        self._busy: dict[str, int] = {}  # passes in flight per store
        self._done: queue.Queue = queue.Queue()
        self._running = False
        self.latencies: list[float] = []
        self.passes = 0
        self.completed = 0
        self.completed_rhs = 0

    # -- resident stores ----------------------------------------------------------------
    def admit(self, name: str, native) -> int:
        """Uploads `native` and keeps it resident; returns its device bytes."""
        with self._cv:
            if name in self._stores:
                raise ValueError(f"store {name!r} is already resident")
        native.device_store(self.device)
        size = int(native.device_stats(self.device)["device_bytes"])
        with self._cv:
            if self._resident_bytes + self._inflight_bytes + size > self.budget:
                native.release_device(self.device)
                raise AdmissionError(f"store {name!r} ({size} bytes) exceeds the HBM budget")
            self._stores[name] = _Resident(native, native.order, size)
            self._resident_bytes += size
        return size

    def evict(self, name: str) -> None:
        """Frees a resident store.  Requests for it still queued fail with
        KeyError; passes already launched on it finish first."""
        with self._cv:
            res = self._stores.pop(name)
            self._resident_bytes -= res.device_bytes
            kept = collections.deque()
            for req in self._queue:
                if req.store == name:
                    req.future.set_exception(KeyError(f"store {name!r} was evicted"))
                else:
                    kept.append(req)
            self._queue = kept
            # wait for the passes in flight on this store to complete
            while self._busy.get(name, 0):
                self._cv.wait()
        res.native.release_device(self.device)

    @property
    def resident_bytes(self) -> int:
        return self._resident_bytes

    # -- requests -------------------------------------------------------------------------
    def submit(self, name: str, x) -> Future:
        """Queues y = H_name x; the future resolves to a device tensor (n, k)."""
        torch = self.torch
        res = self._stores.get(name)
        if res is None:
            raise KeyError(f"store {name!r} is not resident")
        finite = None
        if isinstance(x, np.ndarray):
            x = np.ascontiguousarray(x, dtype=np.float64)
            if not np.isfinite(x).all():
                raise ValueError("nonfinite right-hand side")
            x = torch.from_numpy(x).to(device=f"cuda:{self.device}")
        else:
            x = x.to(device=f"cuda:{self.device}", dtype=torch.float64)
            # checked on the device, without a host round trip; the request's
            # future raises when its pass completes
            finite = torch.isfinite(x).all()
        x = x.reshape(res.order, -1)
        if x.shape[1] > self.max_rhs:
            raise ValueError(f"request of {x.shape[1]} right-hand sides exceeds the server's "
                             f"max_rhs ({self.max_rhs}); split it")
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(self.device))
        req = _Request(name, x, ready, int(x.shape[1]), time.perf_counter(), finite)
        with self._cv:
            self._queue.append(req)
            self._cv.notify()
        return req.future

    # -- loop -------------------------------------------------------------------------------
    def start(self) -> "Server":
        self._running = True
        self._threads = [threading.Thread(target=self._dispatch_loop, daemon=True),
                         threading.Thread(target=self._complete_loop, daemon=True)]
        for t in self._threads:
            t.start()
        return self

    def stop(self) -> None:
        with self._cv:
            self._running = False
            self._cv.notify_all()
        self._threads[0].join()
        self._done.put(None)
        self._threads[1].join()

    def __enter__(self):
        return self.start()

    def __exit__(self, *exc):
        self.stop()
        return False

    def _take(self) -> dict:
        """Everything queued that fits the memory budget, grouped by store
        (name -> (resident store, requests)); called with the lock held."""
        groups: dict[str, list[_Request]] = collections.OrderedDict()
        rhs: dict[str, int] = {}
        kept = collections.deque()
        while self._queue:
            req = self._queue.popleft()
            res = self._stores.get(req.store)
            if res is None:  # evicted meanwhile
                req.future.set_exception(KeyError(f"store {req.store!r} is not resident"))
                continue
            k_now = rhs.get(req.store, 0)
            cost = 16 * res.order * req.k
            over_budget = (self._resident_bytes + self._inflight_bytes + cost > self.budget
                           and (groups or self._inflight_bytes))
            # a lone request always fits a pass (submit bounds its width)
            if (k_now and k_now + req.k > self.max_rhs) or over_budget:
                kept.append(req)  # waits for the next pass
                continue
            if req.store not in groups:
                # the store stays resident until this pass completes
                self._busy[req.store] = self._busy.get(req.store, 0) + 1
                groups[req.store] = (res, [])
            groups[req.store][1].append(req)
            rhs[req.store] = k_now + req.k
            self._inflight_bytes += cost
        self._queue = kept
        return groups

    def _dispatch_loop(self) -> None:
        torch = self.torch
        while True:
            self._slots.acquire()
            with self._cv:
                while self._running and not self._queue:
                    self._cv.wait()
                if not self._running and not self._queue:
                    self._slots.release()
                    return
                groups = self._take()
            if not groups:
                # everything queued waits for memory: until a pass completes
                self._slots.release()
                with self._cv:
                    self._cv.wait(timeout=0.001)
                continue
            stream = self._streams[self._next_stream]
            self._next_stream = (self._next_stream + 1) % len(self._streams)
            work = []
            with torch.cuda.stream(stream):
                for name, (res, reqs) in groups.items():
                    for r in reqs:
                        stream.wait_event(r.ready)
                    x = reqs[0].x if len(reqs) == 1 else torch.cat([r.x for r in reqs], dim=1)
                    k = x.shape[1]
                    ws = torch.empty(max(workspace_bytes(res.native, k, self.device), 8),
                                     dtype=torch.uint8, device=x.device)
                    y = device_apply(res.native, x, precision=self.precision, stream=stream,
                                     workspace=ws)
                    work.append((reqs, y, 16 * res.order * k))
                done = torch.cuda.Event()
                done.record(stream)
            self.passes += 1
            self._done.put((done, work))

    def _complete_loop(self) -> None:
        while True:
            item = self._done.get()
            if item is None:
                return
            done, work = item
            done.synchronize()
            now = time.perf_counter()
            freed = 0
            for reqs, y, cost in work:
                col = 0
                for r in reqs:
                    # statistics first, so that they include a request once its
                    # result is visible
                    self.latencies.append(now - r.arrival)
                    self.completed += 1
                    self.completed_rhs += r.k
                    if r.finite is not None and not bool(r.finite):
                        r.future.set_exception(ValueError("nonfinite right-hand side"))
                    else:
                        r.future.set_result(y[:, col:col + r.k])
                    col += r.k
                freed += cost
            with self._cv:
                self._inflight_bytes -= freed
                for reqs, _, _ in work:
                    self._busy[reqs[0].store] -= 1
                self._cv.notify_all()
            self._slots.release()

    # -- statistics ---------------------------------------------------------------------------
    def reset_stats(self) -> None:
        """Forgets the latencies and counts recorded so far (e.g. of a warm-up)."""
        self.latencies = []
        self.passes = 0
        self.completed = 0
        self.completed_rhs = 0

    def latency_ms(self) -> dict:
        lat = np.asarray(self.latencies) * 1e3
        if lat.size == 0:
            return {"p50": None, "p99": None, "mean": None}
        return {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                "mean": float(lat.mean())}


def poisson_load(server: Server, names: list[str], orders: dict[str, int], rate: float,
                 duration: float, ks=(1, 4, 16), session: int = 50, seed: int = 0) -> dict:
    """Open-loop Poisson arrivals at `rate` req/s for `duration` seconds.

    Requests come in sessions of `session` consecutive requests (a GMRES
    client's iterations) that target one store with one k.  One request per
    (store, k) is served first and its statistics dropped, so the figures are
    of the steady state (first launches of each kernel shape, workspace
    growth).  Returns throughput and latency figures.
    """
    import torch

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=int(rate * duration * 1.2) + 10)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < duration]
    # pre-made right-hand sides per (store, k), already on the device
    xs = {(nm, k): torch.randn(orders[nm], k, dtype=torch.float64, device=f"cuda:{server.device}")
          for nm in names for k in ks}
    torch.cuda.synchronize()
    for key in xs:
        server.submit(key[0], xs[key]).result()
    server.reset_stats()
    futures = []
    # The generator sleeps until each arrival (a busy-wait would hold the
    # interpreter lock from the dispatcher), with a short switch interval so
    # the server threads run promptly.
    interval = sys.getswitchinterval()
    sys.setswitchinterval(5e-5)
    try:
        t0 = time.perf_counter()
        s_name, s_k = None, None
        for i, at in enumerate(arrivals):
            if i % session == 0:
                s_name = names[int(rng.integers(len(names)))]
                s_k = int(ks[int(rng.integers(len(ks)))])
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            futures.append(server.submit(s_name, xs[(s_name, s_k)]))
        for f in futures:
            f.result()
        elapsed = time.perf_counter() - t0
    finally:
        sys.setswitchinterval(interval)
    out = {"offered_rate": rate, "requests": len(futures), "seconds": elapsed,
           "throughput_rps": len(futures) / elapsed,
           "passes": server.passes}
    out.update({f"latency_{k}_ms": v for k, v in server.latency_ms().items()})
    return out


def closed_loop_load(server: Server, names: list[str], orders: dict[str, int], rate: float,
                     duration: float, ks=(1, 4, 16), session: int = 50, seed: int = 0) -> dict:
    """Closed-loop GMRES-like clients: sessions start at Poisson times (rate /
    session per second, so `rate` requests per second are offered while the
    server keeps up); each session targets one store with one k and issues
    its `session` requests one after another, the next as soon as the
    previous result is back.  Returns the offered and achieved request rates,
    the session count and the latency figures of the requests.
    """
    import torch

    rng = np.random.default_rng(seed)
    xs = {(nm, k): torch.randn(orders[nm], k, dtype=torch.float64, device=f"cuda:{server.device}")
          for nm in names for k in ks}
    torch.cuda.synchronize()
    for key in xs:  # first launch of every shape, dropped from the statistics
        server.submit(key[0], xs[key]).result()
    server.reset_stats()
    gaps = rng.exponential(session / rate, size=int(rate / session * duration * 1.5) + 4)
    starts = np.cumsum(gaps)
    starts = starts[starts < duration]
    plans = [(names[int(rng.integers(len(names)))], int(ks[int(rng.integers(len(ks)))]))
             for _ in starts]
    lock = threading.Lock()
    finished = threading.Event()
    state = {"open": len(plans), "errors": []}

    def run_session(name: str, k: int):
        left = [session]

        def step(fut=None):
            if fut is not None:
                err = fut.exception()
                if err is not None:
                    with lock:
                        state["errors"].append(err)
                    left[0] = 0
            if left[0] == 0:
                with lock:
                    state["open"] -= 1
                    if state["open"] == 0:
                        finished.set()
                return
            left[0] -= 1
            try:
                fut = server.submit(name, xs[(name, k)])
            except Exception as e:  # e.g. the store was evicted: the session ends
                with lock:
                    state["errors"].append(e)
                left[0] = 0
                step()
                return
            fut.add_done_callback(step)

        step()

    interval = sys.getswitchinterval()
    sys.setswitchinterval(5e-5)
    try:
        t0 = time.perf_counter()
        if not plans:
            finished.set()
        for at, (name, k) in zip(starts, plans):
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            run_session(name, k)
        finished.wait()
        elapsed = time.perf_counter() - t0
    finally:
        sys.setswitchinterval(interval)
    if state["errors"]:
        raise state["errors"][0]
    done = server.completed
    # the Poisson draw of session starts offers len(plans) * session
    # requests over `duration` seconds: the rate the run actually offered
    out = {"mode": "closed-loop sessions", "offered_rate": rate, "sessions": len(plans),
           "session_requests": session,
           "offered_rate_drawn": len(plans) * session / duration if duration > 0 else 0.0,
           "requests": done, "seconds": round(elapsed, 4),
           "throughput_rps": done / elapsed, "passes": server.passes}
    out.update({f"latency_{k}_ms": v for k, v in server.latency_ms().items()})
    return out


# ---- the library's serving loop (qh_server_*) ------------------------------------------


def _server_sigs():
    import ctypes as C

    from . import _native as N

    if getattr(_server_sigs, "done", False):
        return N
    _p, _u64, _i64, _i32, _dbl = C.c_void_p, C.c_uint64, C.c_int64, C.c_int32, C.c_double
    N._sig("qh_server_create", C.POINTER(ServerConfig), C.POINTER(_p))
    N._sig("qh_server_destroy", _p)
    N._sig("qh_server_admit", _p, _p, C.POINTER(_i32))
    N._sig("qh_server_evict", _p, _i32)
    N._sig("qh_server_submit", _p, _i32, _p, _i64, _p, C.POINTER(_u64))
    N._sig("qh_server_wait", _p, _u64, _dbl, C.POINTER(_i32))
    N._sig("qh_server_get_stats", _p, C.POINTER(ServerStats))
    N._sig("qh_server_reset_stats", _p)
    N._sig("qh_server_closed_loop", _p, C.POINTER(_i32), _u64, C.POINTER(_i64), _u64, _dbl, _dbl,
           _i32, _u64, C.POINTER(LoadResult))
    _server_sigs.done = True
    return N


import ctypes as _C  # noqa: E402


class ServerConfig(_C.Structure):
    _fields_ = [("device", _C.c_int32), ("precision", _C.c_int32), ("slots", _C.c_int32),
                ("max_rhs", _C.c_int64), ("hbm_budget", _C.c_uint64)]


class ServerStats(_C.Structure):
    _fields_ = [("completed", _C.c_uint64), ("completed_rhs", _C.c_uint64), ("passes", _C.c_uint64),
                ("resident_bytes", _C.c_uint64), ("buffer_bytes", _C.c_uint64),
                ("p50_ms", _C.c_double), ("p99_ms", _C.c_double), ("mean_ms", _C.c_double)]


class LoadResult(_C.Structure):
    _fields_ = [("offered_rate", _C.c_double), ("offered_rate_drawn", _C.c_double),
                ("sessions", _C.c_uint64), ("requests", _C.c_uint64), ("passes", _C.c_uint64),
                ("errors", _C.c_uint64), ("seconds", _C.c_double), ("throughput_rps", _C.c_double),
                ("p50_ms", _C.c_double), ("p99_ms", _C.c_double), ("mean_ms", _C.c_double)]


_TICKET_PENDING, _TICKET_DONE, _TICKET_FAILED = 0, 1, 2


@dataclass
class NativeRequest:
    """A request queued on a NativeServer: wait() for its y."""

    ticket: int
    x: object
    y: object


class NativeServer:
    """The serving loop of the library (qh_server_*): the same admission and
    continuous batching as Server, with the dispatcher and the completion
    thread native, so no interpreter lock is on a request's path.

    Requests name a store admitted under a name; x is an FP64 device tensor
    (n, k) whose values are complete when submitted (the server's passes run
    on its own streams; submit synchronises the current stream when asked
    to).  A pass merges every queued request of the store at the head of the
    queue up to max_rhs columns (gathered into one apply and scattered back);
    a request's result is bitwise qh_apply's for it alone on the same
    arithmetic path, within the precision's tolerance otherwise.
    """

    def __init__(self, device: int = 0, hbm_budget: int = 64 << 30, precision: str = "fp64",
                 slots: int = 3, max_rhs: int = 64):
        N = _server_sigs()
        self._N = N
        self.device = device
        self.precision = precision
        self.max_rhs = int(max_rhs)
        cfg = ServerConfig(device, precision_code(precision), slots, self.max_rhs, int(hbm_budget))
        handle = _C.c_void_p()
        N.check(N.lib().qh_server_create(_C.byref(cfg), _C.byref(handle)))
        self._h = handle
        self._ids: dict[str, int] = {}
        self._natives: dict[str, object] = {}

    def close(self) -> None:
        """Stops the server; the admitted stores are released from it."""
        if self._h:
            self._N.check(self._N.lib().qh_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, name: str, native) -> None:
        if name in self._ids:
            raise ValueError(f"store {name!r} is already resident")
        store = native.device_store(self.device)
        sid = _C.c_int32()
        self._N.check(self._N.lib().qh_server_admit(self._h, store, _C.byref(sid)))
        self._ids[name] = int(sid.value)
        self._natives[name] = native

    def evict(self, name: str) -> None:
        sid = self._ids.pop(name)
        self._natives.pop(name)
        self._N.check(self._N.lib().qh_server_evict(self._h, sid))

    def submit(self, name: str, x, sync: bool = True) -> NativeRequest:
        """Queues y = H_name x for an FP64 device tensor x (n,) or (n, k)."""
        torch = __import__("torch")
        native = self._natives[name]
        x = x.reshape(native.order, -1)
        if x.dtype != torch.float64 or not x.is_cuda or not x.is_contiguous():
            raise ValueError("x must be a contiguous FP64 device tensor")
        y = torch.empty_like(x)
        if sync:
            torch.cuda.current_stream(x.device).synchronize()
        ticket = _C.c_uint64()
        self._N.check(self._N.lib().qh_server_submit(self._h, self._ids[name], x.data_ptr(),
                                                     x.shape[1], y.data_ptr(), _C.byref(ticket)))
        return NativeRequest(int(ticket.value), x, y)

    def wait(self, req: NativeRequest, timeout: float | None = None):
        state = _C.c_int32()
        self._N.check(self._N.lib().qh_server_wait(self._h, req.ticket,
                                                   -1.0 if timeout is None else float(timeout),
                                                   _C.byref(state)))
        if state.value == _TICKET_PENDING:
            raise TimeoutError("the request is not done")
        if state.value == _TICKET_FAILED:
            raise RuntimeError("the request failed (its store was evicted or its pass failed)")
        return req.y

    def stats(self) -> dict:
        st = ServerStats()
        self._N.check(self._N.lib().qh_server_get_stats(self._h, _C.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def reset_stats(self) -> None:
        self._N.check(self._N.lib().qh_server_reset_stats(self._h))

    def closed_loop(self, names: list[str], ks=(1, 4, 16), rate: float = 2000.0,
                    duration: float = 3.0, session: int = 50, seed: int = 0) -> dict:
        """Closed-loop GMRES-like sessions driven from the library's own
        threads (qh_server_closed_loop); the figures closed_loop_load gives
        for Server."""
        ids = (_C.c_int32 * len(names))(*[self._ids[n] for n in names])
        kk = (_C.c_int64 * len(ks))(*[int(k) for k in ks])
        out = LoadResult()
        self._N.check(self._N.lib().qh_server_closed_loop(self._h, ids, len(names), kk, len(ks),
                                                          float(rate), float(duration),
                                                          int(session), int(seed),
                                                          _C.byref(out)))
        r = {name: getattr(out, name) for name, _ in out._fields_}
        return {"mode": "closed-loop sessions, native server", "offered_rate": r["offered_rate"],
                "sessions": r["sessions"], "session_requests": session,
                "offered_rate_drawn": r["offered_rate_drawn"], "requests": r["requests"],
                "seconds": r["seconds"], "throughput_rps": r["throughput_rps"],
                "passes": r["passes"], "errors": r["errors"], "latency_p50_ms": r["p50_ms"],
                "latency_p99_ms": r["p99_ms"], "latency_mean_ms": r["mean_ms"]}
