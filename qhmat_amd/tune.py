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
                   repeats: int = 20, warmup: int = 3) -> TuneResult:
    """Times every candidate block for applies of k columns and keeps the
    fastest in the store.  Needs a HIP device."""
    import torch

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
    return result


def tune_batch(natives, ks, *, precision: str = "fp32split", device: int = 0,
               repeats: int = 10) -> list[TuneResult]:
    """Tunes every distinct (store, k) of a batch before the batch is built,
    so that the batch groups its members by the tuned blocks."""
    done = {}
    for native, k in zip(natives, ks):
        key = (id(native), int(k))
        if key not in done:
            done[key] = tune_rhs_block(native, int(k), precision=precision, device=device,
                                       repeats=repeats)
    return list(done.values())
