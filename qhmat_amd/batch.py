"""Batched apply of many resident stores (qh_batch_* / qh_graph_*).

Members of uneven order, rank and right-hand-side count are applied by one
launch pair per right-hand-side class; the whole batch can be captured once
into a HIP graph and replayed with a single launch.  Each member's result is
bitwise the result of applying it alone.
"""

from __future__ import annotations

import ctypes as C
from typing import Sequence

from . import _native as N
from .device import precision_code

_lib = N.lib()
_p = C.c_void_p
N._sig("qh_batch_create", C.POINTER(_p), C.POINTER(C.c_int64), C.c_uint64, C.POINTER(_p))
N._sig("qh_batch_destroy", _p)
N._sig("qh_batch_layout", _p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
       C.POINTER(C.c_uint64))
N._sig("qh_batch_apply", _p, _p, _p, C.c_int, _p, _p, C.c_uint64)
N._sig("qh_batch_graph_create", _p, _p, _p, C.c_int, _p, C.c_uint64, C.POINTER(_p))
N._sig("qh_graph_launch", _p, _p)
N._sig("qh_graph_destroy", _p)


def _stream_handle(stream):
    if stream is None:
        import torch

        stream = torch.cuda.current_stream()
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


class Graph:
    """An instantiated HIP graph of one batch apply."""

    def __init__(self, handle, keep):
        self._h = handle
        self._keep = keep  # tensors the graph reads and writes

    def launch(self, stream=None) -> None:
        N.check(_lib.qh_graph_launch(self._h, _p(_stream_handle(stream))))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.qh_graph_destroy(self._h)
            self._h = None


class DeviceBatch:
    """A fixed set of resident stores applied together on one device."""

    def __init__(self, natives: Sequence, ks: Sequence[int], device: int = 0):
        if len(natives) != len(ks) or not natives:
            raise ValueError("one right-hand-side count per store, at least one store")
        self.natives = list(natives)
        self.ks = [int(k) for k in ks]
        self.device = device
        stores = (_p * len(natives))(*[n.device_store(device) for n in natives])
        karr = (C.c_int64 * len(ks))(*self.ks)
        h = _p()
        N.check(_lib.qh_batch_create(stores, karr, len(natives), C.byref(h)))
        self._h = h
        offs = (C.c_uint64 * len(natives))()
        total, ws = C.c_uint64(), C.c_uint64()
        N.check(_lib.qh_batch_layout(h, offs, C.byref(total), C.byref(ws)))
        self.offsets = list(offs)
        self.total = int(total.value)
        self.workspace_bytes = int(ws.value)
        self.orders = [n.order for n in natives]

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.qh_batch_destroy(self._h)
            self._h = None

    # -- buffers ------------------------------------------------------------------
    def empty(self):
        import torch

        return torch.empty(self.total, dtype=torch.float64, device=f"cuda:{self.device}")

    def workspace(self):
        import torch

        return torch.empty(max(self.workspace_bytes, 8), dtype=torch.uint8,
                           device=f"cuda:{self.device}")

    def pack(self, xs, out=None):
        """Concatenates per-member (n_i, k_i) tensors into one batch buffer."""
        out = self.empty() if out is None else out
        for x, off, n, k in zip(xs, self.offsets, self.orders, self.ks):
            out[off:off + n * k].view(n, k).copy_(x.reshape(n, k))
        return out

    def views(self, buf):
        """Per-member (n_i, k_i) views of a batch buffer."""
        return [buf[off:off + n * k].view(n, k)
                for off, n, k in zip(self.offsets, self.orders, self.ks)]

    # -- apply -----------------------------------------------------------------------
    def apply(self, x, y=None, *, precision: str = "fp64", stream=None, workspace=None):
        y = self.empty() if y is None else y
        ws = self.workspace() if workspace is None else workspace
        N.check(_lib.qh_batch_apply(self._h, _p(x.data_ptr()), _p(y.data_ptr()),
                                    precision_code(precision), _p(_stream_handle(stream)),
                                    _p(ws.data_ptr()), ws.numel()))
        return y

    def graph(self, x, y, *, precision: str = "fp64", workspace=None) -> Graph:
        """Captures one apply with these buffers into a replayable graph."""
        ws = self.workspace() if workspace is None else workspace
        h = _p()
        N.check(_lib.qh_batch_graph_create(self._h, _p(x.data_ptr()), _p(y.data_ptr()),
                                           precision_code(precision), _p(ws.data_ptr()),
                                           ws.numel(), C.byref(h)))
        return Graph(h, (x, y, ws, self))
