"""Launch-shape autotuner for the resident apply.

The apply's one free launch parameter per call is how many right-hand sides a
pass carries: 1, 2 or 4 on the vector ALUs (the narrow passes) or 16, 32 or
64 on the matrix cores (the wide passes); more columns take several passes.
Fewer per pass means fewer registers and LDS and more waves per CU but more
passes over the codes; more per pass the reverse.  The result is bitwise the same for every
choice (each column's arithmetic is independent of the others), so the tuner
only times: it applies the store with every candidate on device tensors,
keeps the fastest per (store, k, precision) and records it in the store
(qh_store_set_rhs_block), where later applies and batches pick it up.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
from dataclasses import dataclass, field

from . import _native as N
from .device import precision_code, workspace_bytes

CANDIDATES = (1, 2, 4, 16, 32, 64)


@dataclass
class TuneResult:
    k: int
    precision: str
    best: int
    us_per_apply: dict = field(default_factory=dict)  # candidate -> microseconds
    cached: bool = False  # taken from a TuneCache without timing


def store_fingerprint(native) -> str:
    """What a launch-shape choice depends on, hashed: the order, every
    block's rectangle, kind and rank and the store's bit totals (the codes'
    values do not change the timing, their shapes do)."""
    h = hashlib.sha1()
    h.update(f"{native.order}:{native.logical_bits}:{native.allocated_bytes}".encode())
    for b in native.blocks:
        r = b.rectangle
        h.update(f"{r.row0},{r.row1},{r.col0},{r.col1},{int(b.dense)},{b.rank};".encode())
    return h.hexdigest()


_BUILD = None


def library_build() -> str:
    """The native library's version and a hash of its file: a rebuilt
    library (kernels changed) invalidates cached shapes."""
    global _BUILD
    if _BUILD is None:
        with open(N.library_path(), "rb") as f:
            _BUILD = f"{N.version()}-{hashlib.sha1(f.read()).hexdigest()[:12]}"
    return _BUILD


class TuneCache:
    """Tuned launch shapes kept across processes in a JSON file, keyed by
    store fingerprint, device name, library version, precision and k."""

    def __init__(self, path: str | os.PathLike):
        self.path = os.fspath(path)
        self.entries: dict = {}
        if os.path.exists(self.path):
            with open(self.path) as f:
                self.entries = json.load(f)

    @staticmethod
    def key(fingerprint: str, device_name: str, precision: str, k: int) -> str:
        return f"{fingerprint}|{device_name}|{library_build()}|{precision}|{int(k)}"

    def lookup(self, key: str):
        e = self.entries.get(key)
        return None if e is None else int(e["best"])

    def record(self, key: str, result: "TuneResult") -> None:
        self.entries[key] = {"best": int(result.best),
                             "us_per_apply": {str(a): round(b, 3) for a, b in
                                              result.us_per_apply.items()}}

    def save(self) -> None:
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.entries, f, indent=1, sort_keys=True)
        os.replace(tmp, self.path)


def candidates(k: int) -> list[int]:
    """Blocks worth trying for k columns: every block up to the first that
    covers k in one pass."""
    out = []
    for kb in CANDIDATES:
        out.append(kb)
        if kb >= k:
            break
    return out


def tune_rhs_block(native, k: int, *, precision: str = "fp32split", device: int = 0,
                   repeats: int = 20, warmup: int = 3, cache: TuneCache | None = None) -> TuneResult:
    """Times every candidate block for applies of k columns and keeps the
    fastest in the store.  With a cache, a shape already tuned for this
    store, device, library version, precision and k is applied without
    timing, and a new one is recorded and saved.  Needs a HIP device."""
    import torch

    key = None
    if cache is not None:
        key = TuneCache.key(store_fingerprint(native), torch.cuda.get_device_name(device),
                            precision, k)
        best = cache.lookup(key)
        if best is not None:
            native.set_rhs_block(k, best, device)
            return TuneResult(k=k, precision=precision, best=best, cached=True)

    n = native.order
    gen = torch.Generator(device="cpu").manual_seed(k)
    x = torch.randn(n, k, dtype=torch.float64, generator=gen).to(f"cuda:{device}")
    y = torch.empty_like(x)
    ws = torch.empty(max(workspace_bytes(native, k), 8), dtype=torch.uint8, device=x.device)
    lib = N.lib()
    store = native.device_store(device)
    stream = torch.cuda.current_stream(x.device)
    args = (C.c_void_p(x.data_ptr()), k, k, C.c_void_p(y.data_ptr()), k,
            precision_code(precision), C.c_void_p(stream.cuda_stream),
            C.c_void_p(ws.data_ptr()), ws.numel())
    result = TuneResult(k=k, precision=precision, best=0)
    for kb in candidates(k):
        native.set_rhs_block(k, kb, device)
        for _ in range(warmup):
            N.check(lib.qh_apply(store, *args))
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(repeats):
            N.check(lib.qh_apply(store, *args))
        end.record(stream)
        end.synchronize()
        result.us_per_apply[kb] = start.elapsed_time(end) * 1e3 / repeats
    result.best = min(result.us_per_apply, key=result.us_per_apply.get)
    native.set_rhs_block(k, result.best, device)
    if cache is not None:
        cache.record(key, result)
        cache.save()
    return result


def tune_batch(natives, ks, *, precision: str = "fp32split", device: int = 0,
               repeats: int = 10, cache: TuneCache | None = None) -> list[TuneResult]:
    """Tunes every distinct (store, k) of a batch before the batch is built,
    so that the batch groups its members by the tuned blocks."""
    done = {}
    for native, k in zip(natives, ks):
        key = (id(native), int(k))
        if key not in done:
            done[key] = tune_rhs_block(native, int(k), precision=precision, device=device,
                                       repeats=repeats, cache=cache)
    return list(done.values())
