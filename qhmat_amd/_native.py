"""ctypes binding of libqhmat_hip.so, shaped like the reference's ``_core``
module (src/qhmat/bindings.cpp:54-204): ``NativeHMatrix``, ``Rectangle``,
``Partition``, ``BlockInfo``, ``Allocation``, ``coefficient``, ``allocate``,
``spatial_partition`` and ``cuda_built``.

The library must be present: importing this module without it raises
ImportError naming the build command.  There is no Python or CPU stand-in for
the device path.
"""

from __future__ import annotations

import ctypes as C
import threading
from dataclasses import dataclass
from pathlib import Path
from typing import Sequence

import numpy as np

try:
    # PyTorch-ROCm ships its own HIP and HSA runtimes (SONAME libamdhip64.so.7,
    # like /opt/rocm's).  Loading torch first makes the dynamic loader bind
    # this library to that same runtime, so device pointers, streams and the
    # device itself are shared with torch instead of opened a second time.
    import torch  # noqa: F401
except ImportError:  # standalone use: the library binds to /opt/rocm's runtime
    pass

_LIB_PATH = Path(__file__).resolve().parent / "libqhmat_hip.so"
if not _LIB_PATH.exists():
    raise ImportError(
        f"{_LIB_PATH} is missing: build it with `python qhmat_amd/build_lib.py` "
        "(hipcc, --offload-arch=gfx950)"
    )
_lib = C.CDLL(str(_LIB_PATH))

# -- status codes (include/qhmat_hip.h) -------------------------------------
QH_OK, QH_EINVAL, QH_ERUNTIME, QH_EOVERFLOW, QH_ENOMEM, QH_ENODEV = range(6)
PRECISION_FP64 = 0
PRECISION_FP32SPLIT = 1

_u64, _i64, _i32, _dbl = C.c_uint64, C.c_int64, C.c_int32, C.c_double
_p = C.c_void_p


class StoreDesc(C.Structure):
    _fields_ = [
        ("order", _u64), ("perm", _p), ("nblocks", _u64), ("rect", _p),
        ("is_dense", _p), ("rank", _p), ("dense", _p), ("widths", _p),
        ("lscale", _p), ("rscale", _p), ("lwords", _p), ("rwords", _p),
    ]


class FactorPlan:
    """Decodes listed factor sides of a resident store into one FP32 device
    buffer per call (qh_factor_plan_*): side 0 of a block as rows x rank,
    side 1 as rank x cols (the right factor transposed), row-major, each at
    its offset (in floats)."""

    def __init__(self, native, blocks, sides, offsets, out_floats: int, device: int = 0):
        b = np.ascontiguousarray(blocks, dtype=np.uint64)
        sd = np.ascontiguousarray(sides, dtype=np.int32)
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        if not (b.shape == sd.shape == off.shape) or b.ndim != 1:
            raise ValueError("blocks, sides and offsets must be one-dimensional and alike")
        self._native = native  # the store outlives the plan
        self.out_floats = int(out_floats)
        h = C.c_void_p()
        check(_lib.qh_factor_plan_create(native.device_store(device), _ptr(b), _ptr(sd), _ptr(off),
                                         b.size, self.out_floats, C.byref(h)))
        self._h = h

    def decode(self, out, stream=None) -> None:
        """Writes the factors into `out`, an FP32 device tensor of at least
        out_floats elements, on `stream` (default: torch's current stream)."""
        import torch

        if out.dtype != torch.float32 or not out.is_cuda or out.numel() < self.out_floats \
                or not out.is_contiguous():
            raise ValueError("output must be a contiguous FP32 device tensor of out_floats elements")
        st = torch.cuda.current_stream() if stream is None else stream
        check(_lib.qh_factor_plan_decode(self._h, C.c_void_p(out.data_ptr()),
                                         C.c_void_p(st.cuda_stream)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.qh_factor_plan_destroy(self._h)
            self._h = None


class StoreStats(C.Structure):
    _fields_ = [
        ("order", _u64), ("nblocks", _u64), ("ndense", _u64), ("nlowrank", _u64),
        ("ncomponents", _u64), ("code_words", _u64), ("left_code_words", _u64),
        ("right_code_words", _u64), ("codes", _u64),
        ("dense_values", _u64), ("device_bytes", _u64),
        ("lowrank_flops_per_rhs", _u64), ("dense_flops_per_rhs", _u64),
        ("tiles", _u64), ("items", _u64), ("right_waves", _u64),
        ("pass_code_words", _u64 * 2), ("fused_components", _u64),
        ("pass_fused_components", _u64 * 2), ("pass_dense_values", _u64 * 2), ("device", _i32),
    ]

    def as_dict(self) -> dict:
        out = {name: getattr(self, name) for name, _ in self._fields_}
        for name in ("pass_code_words", "pass_fused_components", "pass_dense_values"):
            out[name] = list(getattr(self, name))
        return out


class Profile(C.Structure):
    _fields_ = [("applies", _u64), ("pass_ms", _dbl * 2)]


class _HostInfo(C.Structure):
    _fields_ = [
        ("order", _u64), ("nblocks", _u64), ("finalized", _i32),
        ("logical_bits", _u64), ("allocated_bytes", _u64), ("resident_bytes", _u64),
        ("reconstruction_bound", _dbl),
    ]


class _BlockInfo(C.Structure):
    _fields_ = [
        ("row0", _u64), ("row1", _u64), ("col0", _u64), ("col1", _u64),
        ("admissible", _i32), ("dense", _i32), ("rank", _u64),
        ("truncation_error", _dbl), ("quantization_bound", _dbl),
        ("logical_bits", _u64), ("allocated_bytes", _u64),
    ]


def _sig(name, *argtypes):
    fn = getattr(_lib, name)
    fn.argtypes = list(argtypes)
    fn.restype = C.c_int
    return fn


_lib.qh_last_error.restype = C.c_char_p
_lib.qh_last_error.argtypes = []
_lib.qh_version.restype = C.c_char_p
_sig("qh_device_count", C.POINTER(C.c_int))
_sig("qh_store_create", C.POINTER(StoreDesc), C.c_int, C.POINTER(_p))
_sig("qh_store_destroy", _p)
_sig("qh_store_get_stats", _p, C.POINTER(StoreStats))
_sig("qh_apply_workspace_bytes", _p, _i64, C.POINTER(_u64))
_sig("qh_apply", _p, _p, _i64, _i64, _p, _i64, C.c_int, _p, _p, _u64)
_sig("qh_apply_noalloc", _p, _p, _i64, _i64, _p, _i64, C.c_int, _p)
_sig("qh_noalloc_scratch_bytes", C.POINTER(_u64))
_sig("qh_decode_codes", _p, _u64, C.c_int, _p, _p)
_sig("qh_reconstruct", _p, _p, _i64, _p)
_sig("qh_factor_plan_create", _p, _p, _p, _p, _u64, _u64, C.POINTER(_p))
_sig("qh_factor_plan_decode", _p, _p, _p)
_sig("qh_factor_plan_destroy", _p)
_sig("qh_plan_emulate", C.POINTER(StoreDesc), _p, _i64, _p)
_sig("qh_plan_emulate_wide", C.POINTER(StoreDesc), _p, _i64, _p, C.c_int)
_sig("qh_store_set_rhs_block", _p, _i64, C.c_int)
_sig("qh_profile_begin", _p, _u64)
_sig("qh_profile_end", _p, C.POINTER(Profile))
_sig("qh_host_create", _u64, _p, C.POINTER(_p))
_sig("qh_host_destroy", _p)
_sig("qh_host_add_dense", _p, _u64, _u64, _u64, _u64, C.c_int, _p, _u64)
_sig("qh_host_add_low_rank", _p, _u64, _u64, _u64, _u64, C.c_int, _p, _p, _u64, _p, _p,
     _u64, _u64, _dbl)
_sig("qh_host_finalize", _p)
_sig("qh_host_matmul", _p, _p, _u64, _p, C.c_int)
_sig("qh_host_reconstruct", _p, _p)
_sig("qh_host_get_info", _p, C.POINTER(_HostInfo))
_sig("qh_host_block_info", _p, _u64, C.POINTER(_BlockInfo))
_sig("qh_host_permutation", _p, _p)
_sig("qh_host_block_words", _p, _u64, C.c_int, _p, C.POINTER(_u64), _p, _p)
_sig("qh_host_save", _p, C.c_char_p)
_sig("qh_host_load", C.c_char_p, C.POINTER(_p))
_sig("qh_host_describe", _p, C.POINTER(StoreDesc))
_sig("qh_coefficient", C.c_int, C.POINTER(_dbl))
_sig("qh_allocate", _p, _u64, _p, _u64, _dbl, _u64, _u64, _p, C.POINTER(_dbl),
     C.POINTER(_u64), C.POINTER(_i32))
_sig("qh_partition_create", _p, _p, _p, _u64, _u64, _u64, _dbl, C.POINTER(_p))
_sig("qh_partition_sizes", _p, C.POINTER(_u64), C.POINTER(_u64))
_sig("qh_partition_copy", _p, _p, _p)
_sig("qh_partition_destroy", _p)


def check(status: int) -> None:
    """Raises the exception the reference would raise for this status."""
    if status == QH_OK:
        return
    message = _lib.qh_last_error().decode("utf-8", "replace")
    if status == QH_EINVAL:
        raise ValueError(message)
    if status == QH_EOVERFLOW:
        raise OverflowError(message)
    if status == QH_ENOMEM:
        raise MemoryError(message)
    raise RuntimeError(message)


def version() -> str:
    """The native library's version string (qh_version)."""
    return _lib.qh_version().decode()


def lib() -> C.CDLL:
    return _lib


def library_path() -> Path:
    return _LIB_PATH


def _ptr(a: np.ndarray | None):
    return None if a is None or a.size == 0 else a.ctypes.data_as(_p)


def _f64(a, name: str) -> np.ndarray:
    try:
        return np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    except (TypeError, ValueError) as exc:  # mirror pybind11's forcecast failure
        raise TypeError(f"{name} must be convertible to a float64 array") from exc


def device_count() -> int:
    n = C.c_int(0)
    check(_lib.qh_device_count(C.byref(n)))
    return int(n.value)


# -- value types (bindings.cpp:57-97) ----------------------------------------
@dataclass(frozen=True)
class Rectangle:
    """Block bounds [row0:row1, col0:col1] in clustered index coordinates."""

    row0: int
    row1: int
    col0: int
    col1: int
    admissible: bool = False

    def __post_init__(self):
        for f in ("row0", "row1", "col0", "col1"):
            v = getattr(self, f)
            if isinstance(v, (bool, np.bool_)) or not isinstance(v, (int, np.integer)) or v < 0:
                raise TypeError(f"{f} must be a nonnegative integer")
            object.__setattr__(self, f, int(v))
        object.__setattr__(self, "admissible", bool(self.admissible))


@dataclass(frozen=True)
class Partition:
    """Clustered permutation and disjoint block rectangles."""

    permutation: list
    blocks: list


@dataclass(frozen=True)
class BlockInfo:
    """Read-only geometry, rank, error, and storage for a stored block."""

    rectangle: Rectangle
    dense: bool
    rank: int
    truncation_error: float
    quantization_bound: float
    logical_bits: int
    allocated_bytes: int


@dataclass(frozen=True)
class Allocation:
    """Result of discrete bit allocation for one factorized block."""

    bits: list
    bound: float
    payload_bits: int
    feasible: bool


def coefficient(bits: int) -> float:
    """Return c(b) = 1/q + 1/(4*q*q), where q = 2**(b-1) - 1."""
    out = _dbl()
    check(_lib.qh_coefficient(int(bits), C.byref(out)))
    return float(out.value)


def allocate(singular_values, bit_widths: Sequence[int], budget: float, rows: int,
             cols: int) -> Allocation:
    """Choose per-component widths satisfying the singular-value surrogate budget."""
    s = _f64(singular_values, "singular_values")
    if s.ndim != 1:
        raise ValueError("singular_values must be one-dimensional")
    menu = np.ascontiguousarray(np.asarray(list(bit_widths), dtype=np.int32))
    bits = np.zeros(s.size, dtype=np.int32)
    bound, payload, feasible = _dbl(), _u64(), _i32()
    check(_lib.qh_allocate(_ptr(s), s.size, _ptr(menu), menu.size, float(budget), int(rows),
                           int(cols), _ptr(bits), C.byref(bound), C.byref(payload),
                           C.byref(feasible)))
    return Allocation([int(b) for b in bits], float(bound.value), int(payload.value),
                      bool(feasible.value))


def spatial_partition(points, support_lower, support_upper, leaf_size: int,
                      admissibility: float) -> Partition:
    """Median-split points and form a disjoint block partition using box separation."""
    p = _f64(points, "points")
    if p.ndim != 2:
        raise ValueError("points must have shape (n, d)")
    n, d = p.shape
    lo = _f64(support_lower, "support_lower")
    hi = _f64(support_upper, "support_upper")
    for name, a in (("support_lower", lo), ("support_upper", hi)):
        if a.shape != (n, d):
            raise ValueError(f"{name} has the wrong shape")
    handle = _p()
    check(_lib.qh_partition_create(_ptr(p), _ptr(lo), _ptr(hi), n, d, int(leaf_size),
                                   float(admissibility), C.byref(handle)))
    try:
        count, nblocks = _u64(), _u64()
        check(_lib.qh_partition_sizes(handle, C.byref(count), C.byref(nblocks)))
        perm = np.zeros(count.value, dtype=np.uint64)
        rects = np.zeros(5 * nblocks.value, dtype=np.uint64)
        check(_lib.qh_partition_copy(handle, _ptr(perm), _ptr(rects)))
    finally:
        _lib.qh_partition_destroy(handle)
    r = rects.reshape(-1, 5)
    blocks = [Rectangle(int(a), int(b), int(c), int(e), bool(f)) for a, b, c, e, f in r]
    return Partition([int(i) for i in perm], blocks)


class NativeHMatrix:
    """Low-level packed H-matrix store (reference bindings.cpp:124-198).

    ``matmul(x, use_cuda=True)`` runs on the MI355X through a store uploaded
    once per device and kept resident; ``use_cuda=False`` runs the host
    backend.  Device-resident inputs (torch tensors) go through
    :meth:`apply_device` without a host round trip.
    """

    def __init__(self, order: int, permutation: Sequence[int], *, _handle=None):
        self._dev = {}
        self._dev_lock = threading.Lock()
        if _handle is not None:
            self._h = _handle
            return
        perm = np.ascontiguousarray(np.asarray(list(permutation), dtype=np.uint64))
        if perm.ndim != 1:
            raise ValueError("permutation must be one-dimensional")
        handle = _p()
        check(_lib.qh_host_create(int(order), _ptr(perm), C.byref(handle)))
        self._h = handle

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            for store in getattr(self, "_dev", {}).values():
                _lib.qh_store_destroy(store)
            self._dev = {}
            _lib.qh_host_destroy(h)
            self._h = None

    # -- construction ------------------------------------------------------
    def add_dense(self, rectangle: Rectangle, data) -> None:
        """Add a dense FP64 block; its row and column clusters must be identical."""
        m, n = rectangle.row1 - rectangle.row0, rectangle.col1 - rectangle.col0
        a = _f64(data, "dense block")
        if a.ndim != 2 or a.shape != (m, n):
            raise ValueError("dense block has the wrong shape")
        check(_lib.qh_host_add_dense(self._h, rectangle.row0, rectangle.row1, rectangle.col0,
                                     rectangle.col1, int(rectangle.admissible), _ptr(a), a.size))

    def add_low_rank(self, rectangle: Rectangle, u, singular_values, v, bits,
                     menu_size: int, truncation_error: float) -> None:
        """Add an off-diagonal SVD block; balanced quantized packing is internal."""
        s = _f64(singular_values, "singular_values")
        if s.ndim != 1:
            raise ValueError("singular_values must be one-dimensional")
        m, n = rectangle.row1 - rectangle.row0, rectangle.col1 - rectangle.col0
        uu, vv = _f64(u, "left singular vectors"), _f64(v, "right singular vectors")
        if uu.ndim != 2 or uu.shape != (m, s.size):
            raise ValueError("left singular vectors has the wrong shape")
        if vv.ndim != 2 or vv.shape != (n, s.size):
            raise ValueError("right singular vectors has the wrong shape")
        ucol = np.ascontiguousarray(uu.T)  # column-major m x r
        vcol = np.ascontiguousarray(vv.T)
        w = np.ascontiguousarray(np.asarray(list(bits), dtype=np.int32))
        check(_lib.qh_host_add_low_rank(self._h, rectangle.row0, rectangle.row1, rectangle.col0,
                                        rectangle.col1, int(rectangle.admissible), _ptr(ucol),
                                        _ptr(s), s.size, _ptr(vcol), _ptr(w), w.size,
                                        int(menu_size), float(truncation_error)))

    def finalize(self) -> None:
        """Validate that added blocks form a disjoint complete partition."""
        check(_lib.qh_host_finalize(self._h))

    def save(self, path: str) -> None:
        """Write a checksummed file; refuse to overwrite an existing file."""
        check(_lib.qh_host_save(self._h, str(path).encode()))

    @staticmethod
    def load(path: str) -> "NativeHMatrix":
        """Load and validate a saved native matrix."""
        handle = _p()
        check(_lib.qh_host_load(str(path).encode(), C.byref(handle)))
        return NativeHMatrix(0, [], _handle=handle)

    # -- properties ----------------------------------------------------------
    def _info(self) -> _HostInfo:
        info = _HostInfo()
        check(_lib.qh_host_get_info(self._h, C.byref(info)))
        return info

    @property
    def order(self) -> int:
        # fixed at creation; cached because the info query walks every block
        n = self.__dict__.get("_order")
        if n is None:
            n = self._order = int(self._info().order)
        return n

    @property
    def logical_bits(self) -> int:
        return int(self._info().logical_bits)

    @property
    def allocated_bytes(self) -> int:
        return int(self._info().allocated_bytes)

    @property
    def resident_bytes(self) -> int:
        return int(self._info().resident_bytes)

    @property
    def reconstruction_bound(self) -> float:
        return float(self._info().reconstruction_bound)

    @property
    def finalized(self) -> bool:
        return bool(self._info().finalized)

    def block(self, index: int) -> BlockInfo:
        """Metadata of one block (insertion order)."""
        info = _BlockInfo()
        check(_lib.qh_host_block_info(self._h, int(index), C.byref(info)))
        return BlockInfo(
            Rectangle(info.row0, info.row1, info.col0, info.col1, bool(info.admissible)),
            bool(info.dense), int(info.rank), float(info.truncation_error),
            float(info.quantization_bound), int(info.logical_bits), int(info.allocated_bytes))

    @property
    def blocks(self) -> list:
        return [self.block(i) for i in range(int(self._info().nblocks))]

    @property
    def permutation(self) -> np.ndarray:
        out = np.zeros(self.order, dtype=np.uint64)
        check(_lib.qh_host_permutation(self._h, _ptr(out)))
        return out

    def block_words(self, index: int, side: int):
        """(words, scales, widths) of one side (0 left, 1 right) of a block."""
        count = _u64()
        check(_lib.qh_host_block_words(self._h, index, side, None, C.byref(count), None, None))
        info = _BlockInfo()
        check(_lib.qh_host_block_info(self._h, index, C.byref(info)))
        rank = int(info.rank)
        words = np.zeros(count.value, dtype=np.uint32)
        scales = np.zeros(rank, dtype=np.float64)
        widths = np.zeros(rank, dtype=np.int32)
        check(_lib.qh_host_block_words(self._h, index, side, _ptr(words), C.byref(count),
                                       _ptr(scales), _ptr(widths)))
        return words, scales, widths

    # -- apply -----------------------------------------------------------------
    def matmul(self, x, use_cuda: bool = False, *, precision: str = "fp64"):
        """Apply to a vector or multiple right-hand sides (bindings.cpp:152-179)."""
        a = _f64(x, "right-hand side")
        if a.ndim not in (1, 2):
            raise ValueError("right-hand side must be one- or two-dimensional")
        n = self.order
        if a.shape[0] != n:
            raise ValueError("right-hand side has the wrong row count")
        rhs = 1 if a.ndim == 1 else a.shape[1]
        if use_cuda:
            if not np.isfinite(a).all():
                raise ValueError("nonfinite right-hand side")
            return self._matmul_device(a, rhs, precision)
        y = np.empty((n, rhs), dtype=np.float64)
        if a.size:
            check(_lib.qh_host_matmul(self._h, _ptr(a), rhs, _ptr(y), 0))
        else:
            check(_lib.qh_host_matmul(self._h, None, rhs, None, 0))
        return y.reshape(n) if a.ndim == 1 else y

    def reconstruct(self) -> np.ndarray:
        """Reconstruct a dense FP64 matrix in source index order."""
        n = self.order
        out = np.zeros((n, n), dtype=np.float64)
        check(_lib.qh_host_reconstruct(self._h, _ptr(out)))
        return out

    # -- device store ---------------------------------------------------------
    def device_store(self, device: int = 0):
        """Handle of the store resident on `device`, uploaded on first use."""
        with self._dev_lock:
            store = self._dev.get(device)
            if store is None:
                desc = StoreDesc()
                check(_lib.qh_host_describe(self._h, C.byref(desc)))
                handle = _p()
                check(_lib.qh_store_create(C.byref(desc), int(device), C.byref(handle)))
                store = self._dev[device] = handle
            return store

    def set_rhs_block(self, k: int, kb: int, device: int = 0) -> None:
        """Applies of k right-hand sides on `device` run in passes of kb
        right-hand sides: 1, 2 or 4 (vector-ALU passes) or 16, 32 or 64
        (matrix-core passes); 0 restores the default.  The result is bitwise
        the same for every kb of one of the two groups; see qhmat_amd.tune for
        choosing it."""
        check(_lib.qh_store_set_rhs_block(self.device_store(device), int(k), int(kb)))

    def upload_copy(self, device: int = 0):
        """A further, independent device copy (the caller frees it with
        qh_store_destroy); used to rotate copies past the last-level cache."""
        desc = StoreDesc()
        check(_lib.qh_host_describe(self._h, C.byref(desc)))
        handle = _p()
        check(_lib.qh_store_create(C.byref(desc), int(device), C.byref(handle)))
        return handle

    def plan_emulate(self, x, wide: bool = False, precision: str = "fp64") -> np.ndarray:
        """Host check of the device layout (qh_plan_emulate, or for the wide
        passes qh_plan_emulate_wide in `precision`'s virtual-row order): y = H x
        computed by replaying the kernels' index arithmetic with bounds checks."""
        a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(self.order, -1))
        y = np.full_like(a, np.nan)
        desc = StoreDesc()
        check(_lib.qh_host_describe(self._h, C.byref(desc)))
        if wide:
            prec = PRECISION_FP32SPLIT if precision == "fp32split" else PRECISION_FP64
            check(_lib.qh_plan_emulate_wide(C.byref(desc), _ptr(a), a.shape[1], _ptr(y), prec))
        else:
            check(_lib.qh_plan_emulate(C.byref(desc), _ptr(a), a.shape[1], _ptr(y)))
        return y

    def device_stats(self, device: int = 0) -> dict:
        stats = StoreStats()
        check(_lib.qh_store_get_stats(self.device_store(device), C.byref(stats)))
        return stats.as_dict()

    def release_device(self, device: int | None = None) -> None:
        """Frees the resident copy on one device (or all).  Raises ValueError,
        keeping the copy, while a batch built on it exists."""
        with self._dev_lock:
            for d in list(self._dev):
                if device is None or d == device:
                    check(_lib.qh_store_destroy(self._dev[d]))
                    del self._dev[d]

    def _matmul_device(self, a: np.ndarray, rhs: int, precision: str) -> np.ndarray:
        from . import device as _device

        y = _device.apply_numpy(self, a.reshape(a.shape[0], rhs), precision)
        return y.reshape(a.shape[0]) if a.ndim == 1 else y

    def apply_device(self, x, y=None, *, precision: str = "fp64", stream=None,
                     workspace=None):
        """y = H x on device tensors (see qhmat_amd.device.apply)."""
        from . import device as _device

        return _device.apply(self, x, y, precision=precision, stream=stream,
                             workspace=workspace)

    def reconstruct_device(self, out=None, *, stream=None):
        """Dense FP64 matrix in source index order, formed on the device
        (qh_reconstruct): a torch tensor (order, order) on the current
        device, or `out` (row-major, unit column stride) filled in place."""
        from . import device as _device

        return _device.reconstruct(self, out, stream=stream)

    def decode_codes(self, block: int, side: int):
        """Integer codes of one block side decoded on the device (test hook)."""
        from . import device as _device

        return _device.decode_codes(self, block, side)


cuda_built = True
"""The library is always built with gfx950 device code (the HIP backend)."""
