"""Device-side apply on PyTorch-ROCm tensors.

PyTorch supplies device memory and the current HIP stream; the arithmetic is
the library's HIP kernels (qh_apply).  Nothing here computes on the host.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native

_PRECISIONS = {
    "fp64": _native.PRECISION_FP64,
    "fp32": _native.PRECISION_FP32SPLIT,
    "fp32split": _native.PRECISION_FP32SPLIT,
}


def precision_code(precision: str) -> int:
    try:
        return _PRECISIONS[precision]
    except KeyError:
        raise ValueError("precision must be 'fp64' or 'fp32split'") from None


def _torch():
    import torch

    return torch


def require_device() -> None:
    """Raises the reference's error when no GPU is visible (cuda.cu:115-117)."""
    torch = _torch()
    if not torch.cuda.is_available() or _native.device_count() == 0:
        raise RuntimeError("no CUDA device is available")


def workspace_bytes(native, k: int, device: int = 0) -> int:
    out = C.c_uint64()
    _native.check(_native.lib().qh_apply_workspace_bytes(native.device_store(device), int(k),
                                                         C.byref(out)))
    return int(out.value)


def apply(native, x, y=None, *, precision: str = "fp64", stream=None, workspace=None):
    """y = H x for an FP64 device tensor x of shape (n,) or (n, k).

    Rows of x and y must be contiguous (stride 1 along k); any row stride is
    accepted.  The apply is ordered on `stream` (a torch stream or a raw
    hipStream_t; default: the current stream of x's device) after the work
    already queued on the current stream, and does not synchronise: the
    result is ready on `stream`.  Copies and allocations made here are made
    on `stream`, and tensors it reads or writes that come from elsewhere are
    recorded on it for the caching allocator.  `workspace`: optional uint8
    device tensor of at least workspace_bytes(native, k).
    """
    torch = _torch()
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise TypeError("x must be a device tensor")
    if x.dtype != torch.float64:
        raise TypeError("x must be float64")
    vec = x.dim() == 1
    x2 = x.unsqueeze(1) if vec else x
    if x2.dim() != 2:
        raise ValueError("right-hand side must be one- or two-dimensional")
    n = native.order
    if x2.shape[0] != n:
        raise ValueError("right-hand side has the wrong row count")
    k = x2.shape[1]
    if k == 0:
        raise ValueError("right-hand-side shape mismatch")
    y2 = None
    if y is not None:
        y2 = y.unsqueeze(1) if y.dim() == 1 else y
        if y2.shape != (n, k) or y2.dtype != torch.float64 or y2.device != x2.device:
            raise ValueError("result tensor has the wrong shape, dtype or device")
        if y2.stride(1) != 1 and k > 1:
            raise ValueError("result rows must be contiguous")
    device = x2.device.index if x2.device.index is not None else torch.cuda.current_device()
    current = torch.cuda.current_stream(x2.device)
    if stream is None:
        st = current
    elif hasattr(stream, "cuda_stream"):
        st = stream
    else:
        st = torch.cuda.ExternalStream(int(stream), device=x2.device)
    other = st.cuda_stream != current.cuda_stream
    if other:
        st.wait_stream(current)  # x and y as the caller's current stream leaves them
    store = native.device_store(device)
    with torch.cuda.stream(st):
        if x2.stride(1) != 1 and k > 1:
            x2 = x2.contiguous()
        if y2 is None:
            y2 = torch.empty((n, k), dtype=torch.float64, device=x2.device)
        ws_ptr, ws_bytes = None, 0
        if workspace is None:
            need = workspace_bytes(native, k, device)
            if need:
                workspace = torch.empty(need, dtype=torch.uint8, device=x2.device)
        if workspace is not None:
            ws_ptr, ws_bytes = workspace.data_ptr(), workspace.numel() * workspace.element_size()
        ldx = x2.stride(0) if k > 1 else max(x2.stride(0), 1)
        ldy = y2.stride(0) if k > 1 else max(y2.stride(0), 1)
        _native.check(_native.lib().qh_apply(
            store, C.c_void_p(x2.data_ptr()), ldx, k, C.c_void_p(y2.data_ptr()), ldy,
            precision_code(precision), C.c_void_p(st.cuda_stream), C.c_void_p(ws_ptr), ws_bytes))
    if other:
        # memory made on another stream, used on st: not reused before st is done
        for t in (x2, y2, workspace):
            if t is not None:
                t.record_stream(st)
    if y is not None:
        return y
    return y2.squeeze(1) if vec else y2


def apply_numpy(native, a: np.ndarray, precision: str = "fp64") -> np.ndarray:
    """numpy (n, k) in, numpy (n, k) out, through the device (blocking)."""
    require_device()
    torch = _torch()
    x = torch.from_numpy(np.ascontiguousarray(a)).to("cuda")
    y = apply(native, x, precision=precision)
    torch.cuda.current_stream().synchronize()
    return y.cpu().numpy()


def reconstruct(native, out=None, *, stream=None):
    """Dense reconstruction on the device (qh_reconstruct); see
    NativeHMatrix.reconstruct_device."""
    require_device()
    torch = _torch()
    n = native.order
    if out is None:
        out = torch.empty((n, n), dtype=torch.float64, device="cuda")
    elif (out.dtype != torch.float64 or out.device.type != "cuda" or out.dim() != 2
          or tuple(out.shape) != (n, n) or out.stride(1) != 1):
        raise ValueError("out must be a float64 device tensor of shape (n, n) with unit column stride")
    store = native.device_store(out.device.index if out.device.index is not None else 0)
    st = stream if stream is not None else torch.cuda.current_stream(out.device)
    handle = C.c_void_p(st.cuda_stream if hasattr(st, "cuda_stream") else st)
    _native.check(_native.lib().qh_reconstruct(store, C.c_void_p(out.data_ptr()), out.stride(0),
                                               handle))
    return out


def decode_codes(native, block: int, side: int) -> np.ndarray:
    """Codes of one block side, (rank, rows) int32, decoded by the device."""
    require_device()
    torch = _torch()
    info = native.block(block)
    if info.dense:
        raise ValueError("dense blocks carry no codes")
    r = info.rectangle
    rows = (r.row1 - r.row0) if side == 0 else (r.col1 - r.col0)
    out = torch.zeros((max(info.rank, 1), rows), dtype=torch.int32, device="cuda")
    store = native.device_store(torch.cuda.current_device())
    stream = torch.cuda.current_stream()
    _native.check(_native.lib().qh_decode_codes(store, int(block), int(side),
                                                C.c_void_p(out.data_ptr()),
                                                C.c_void_p(stream.cuda_stream)))
    stream.synchronize()
    return out.cpu().numpy()[: info.rank]
