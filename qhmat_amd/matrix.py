"""High-level H-matrix object with the reference's Python surface
(python/qhmat/matrix.py): ``HMatrix.from_dense``, ``from_partition``,
``matmul(x, device=...)``, ``@``, ``to_dense``, ``save``/``load``,
``storage_report``, ``InfeasibleError``, ``StorageReport`` and
``memory_bytes``.

``device="cuda"`` runs the MI355X kernels through a store uploaded once and
kept resident ("cuda" is the HIP device namespace under PyTorch-ROCm).  A
device tensor passed to ``matmul`` stays on the device.  ``precision``
selects FP64 (default, the parity anchor) or the FP32 split-precision apply.
"""

from __future__ import annotations

import os
from collections.abc import Sequence
from dataclasses import dataclass

import numpy as np

from ._native import NativeHMatrix, Rectangle, allocate, spatial_partition
from .blockwork import InfeasibleError, truncate


@dataclass(frozen=True)
class StorageReport:
    """Sizes for comparing a store with its dense FP64 source: the n x n
    FP64 array, the logical bit count of codes, scales and labels, and the
    allocated and resident byte counts.  ``source_array_bytes`` is known only
    for a store built in this process (None after ``load``)."""

    dense_fp64_bytes: int
    logical_bits: int
    allocated_bytes: int
    resident_bytes: int
    source_array_bytes: int | None = None

    @property
    def logical_bytes(self) -> int:
        return (self.logical_bits + 7) // 8

    @property
    def logical_compression(self) -> float:
        return self.dense_fp64_bytes / (self.logical_bits / 8)

    @property
    def resident_compression(self) -> float:
        return self.dense_fp64_bytes / self.resident_bytes


def _human(size: int) -> str:
    if size < 1024:
        return f"{size} B"
    value = float(size)
    for unit in ("KiB", "MiB", "GiB", "TiB"):
        value /= 1024
        if value < 1024 or unit == "TiB":
            return f"{value:.1f} {unit}"
    return f"{value:.1f} TiB"


def _print_report(report: StorageReport) -> None:
    lines = []
    if report.source_array_bytes is not None:
        lines.append(f"Input array: {_human(report.source_array_bytes)}")
    lines.append(f"Dense FP64 baseline: {_human(report.dense_fp64_bytes)}")
    lines.append(f"H-matrix logical storage: {report.logical_bits} bits "
                 f"({report.logical_compression:.2f}x compression vs FP64)")
    lines.append(f"H-matrix C++ resident estimate: {_human(report.resident_bytes)} "
                 f"({report.resident_compression:.2f}x compression vs FP64)")
    print("\n".join(lines))


def _check_common(tolerance, rank_tolerance, max_rank, truncation_fraction,
                  arithmetic_reserve, bit_widths) -> tuple:
    if not np.isfinite(tolerance) or tolerance <= 0:
        raise ValueError("tolerance must be finite and positive")
    if not np.isfinite(rank_tolerance) or not 0 < rank_tolerance <= 1:
        raise ValueError("rank_tolerance must lie in (0, 1]")
    if max_rank is not None and (not isinstance(max_rank, int) or max_rank < 1):
        raise ValueError("max_rank must be a positive integer or None")
    if not (np.isfinite(truncation_fraction) and np.isfinite(arithmetic_reserve)):
        raise ValueError("budget fractions must be finite")
    if not (0 <= arithmetic_reserve < 1 and 0 < truncation_fraction < 1 - arithmetic_reserve):
        raise ValueError("invalid truncation/arithmetic budget fractions")
    widths = tuple(bit_widths)
    valid = bool(widths) and len(set(widths)) == len(widths) and all(
        type(b) is int and 2 <= b <= 32 for b in widths)
    if not valid:
        raise ValueError("bit_widths must contain distinct integers in [2, 32]")
    return widths


def _square(matrix) -> np.ndarray:
    a = np.asarray(matrix, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != a.shape[1] or a.shape[0] == 0:
        raise ValueError("matrix must be nonempty and square")
    return a


def factor_block(block: np.ndarray, rect: Rectangle, *, tolerance: float, widths: tuple,
                 rank_tolerance: float, max_rank: int | None, truncation_fraction: float,
                 arithmetic_reserve: float, spectrum=None):
    """Rank, factors and widths of one off-diagonal block (matrix.py:213-245).

    Returns (u, sigma, v, bits, residual); rank 0 for an all-zero block.
    Raises InfeasibleError when the block cannot meet its budget.
    ``spectrum`` as in blockwork.truncate (the randomized range finder for
    large synthetic blocks; the full SVD when None).
    """
    box = (rect.row0, rect.row1, rect.col0, rect.col1)
    u, sigma, v, residual, quant_budget = truncate(
        block, box, tolerance=tolerance, rank_tolerance=rank_tolerance, max_rank=max_rank,
        truncation_fraction=truncation_fraction, arithmetic_reserve=arithmetic_reserve,
        spectrum=spectrum)
    if quant_budget is None:
        return u, sigma, v, [], residual
    return u, sigma, v, allocate_widths(sigma, widths, quant_budget, block.shape, box), residual


def allocate_widths(sigma, widths, quant_budget: float, shape, box) -> list:
    """Bit widths of the kept components within the quantization budget
    (native allocator); InfeasibleError when even the widest code misses it."""
    plan = allocate(sigma, widths, quant_budget, shape[0], shape[1])
    if not plan.feasible:
        raise InfeasibleError(f"block ({box[0]}:{box[1]}, {box[2]}:{box[3]}) "
                              "is infeasible even at the highest bit width")
    return plan.bits


class HMatrix:
    """H-matrix whose diagonal blocks stay dense FP64 and whose other blocks
    are truncated SVD factors stored as integer codes of per-component
    widths.

    Construction uses NumPy's LAPACK-backed SVD; packing, allocation, storage
    accounting and the host apply are native; ``device="cuda"`` applies on the
    MI355X through a resident device store.
    """

    def __init__(self, native: NativeHMatrix, source_array_bytes: int | None = None):
        self._native = native
        self._source_array_bytes = source_array_bytes

    # -- construction (matrix.py:107-276) --------------------------------------
    @classmethod
    def from_dense(cls, matrix, points, *, tolerance: float, leaf_size: int = 32,
                   admissibility: float = 1.0,
                   bit_widths: Sequence[int] = (4, 8, 12, 16, 24, 32),
                   rank_tolerance: float = 0.1, max_rank: int | None = None,
                   truncation_fraction: float = 0.25, arithmetic_reserve: float = 0.02,
                   support_lower=None, support_upper=None, verbose: bool = False) -> "HMatrix":
        """Cluster `points`, split `matrix` into dense diagonal and factored
        off-diagonal blocks, and pack each off-diagonal block's factors."""
        a = _square(matrix)
        p = np.asarray(points, dtype=np.float64)
        if p.ndim != 2 or p.shape[0] != a.shape[0] or p.shape[1] == 0:
            raise ValueError("points must have shape (matrix order, dimension)")
        if not (np.isfinite(a).all() and np.isfinite(p).all()):
            raise ValueError("matrix and points must be finite")
        if not np.isfinite(tolerance) or tolerance <= 0:
            raise ValueError("tolerance must be finite and positive")
        if not isinstance(leaf_size, int) or leaf_size < 1:
            raise ValueError("leaf_size must be a positive integer")
        widths = _check_common(tolerance, rank_tolerance, max_rank, truncation_fraction,
                               arithmetic_reserve, bit_widths)
        lo = p if support_lower is None else np.asarray(support_lower, dtype=np.float64)
        hi = p if support_upper is None else np.asarray(support_upper, dtype=np.float64)
        part = spatial_partition(p, lo, hi, leaf_size, admissibility)
        out = cls.from_partition(a, part.permutation, part.blocks, tolerance=tolerance,
                                 bit_widths=widths, rank_tolerance=rank_tolerance,
                                 max_rank=max_rank, truncation_fraction=truncation_fraction,
                                 arithmetic_reserve=arithmetic_reserve)
        out._source_array_bytes = int(np.asarray(matrix).nbytes)
        if verbose:
            _print_report(out.storage_report())
        return out

    @classmethod
    def from_partition(cls, matrix, permutation: Sequence[int], blocks: Sequence[Rectangle], *,
                       tolerance: float, bit_widths: Sequence[int] = (4, 8, 12, 16, 24, 32),
                       rank_tolerance: float = 0.1, max_rank: int | None = None,
                       truncation_fraction: float = 0.25, arithmetic_reserve: float = 0.02,
                       verbose: bool = False) -> "HMatrix":
        """Store the caller's block rectangles (rows and columns counted in the
        order `permutation` gives) of `matrix`."""
        a = _square(matrix)
        if not np.isfinite(a).all():
            raise ValueError("matrix must be finite")
        widths = _check_common(tolerance, rank_tolerance, max_rank, truncation_fraction,
                               arithmetic_reserve, bit_widths)
        entries = tuple(permutation)
        if any(isinstance(i, (bool, np.bool_)) or not isinstance(i, (int, np.integer))
               for i in entries):
            raise ValueError("permutation must contain integer indices")
        perm = [int(i) for i in entries]
        if len(perm) != a.shape[0] or set(perm) != set(range(a.shape[0])):
            raise ValueError("permutation must contain each index exactly once")
        native = NativeHMatrix(a.shape[0], perm)
        order = np.asarray(perm, dtype=np.int64)
        for rect in blocks:
            if not isinstance(rect, Rectangle):
                raise TypeError("blocks must contain Rectangle objects")
            block = np.ascontiguousarray(
                a[np.ix_(order[rect.row0:rect.row1], order[rect.col0:rect.col1])])
            if block.size == 0:
                raise ValueError("empty block in partition")
            if rect.row0 == rect.col0 and rect.row1 == rect.col1:
                native.add_dense(rect, block)
                continue
            if not (rect.row1 <= rect.col0 or rect.col1 <= rect.row0):
                raise ValueError("non-diagonal block intersects the diagonal")
            u, sigma, v, bits, residual = factor_block(
                block, rect, tolerance=tolerance, widths=widths, rank_tolerance=rank_tolerance,
                max_rank=max_rank, truncation_fraction=truncation_fraction,
                arithmetic_reserve=arithmetic_reserve)
            native.add_low_rank(rect, u, sigma, v, bits, len(widths), residual)
        native.finalize()
        out = cls(native, source_array_bytes=int(np.asarray(matrix).nbytes))
        if verbose:
            _print_report(out.storage_report())
        return out

    # -- properties ---------------------------------------------------------------
    @property
    def shape(self) -> tuple[int, int]:
        n = self._native.order
        return n, n

    @property
    def blocks(self):
        return self._native.blocks

    @property
    def logical_bits(self) -> int:
        """Bits of the codes plus, per component, its two FP64 scales and its
        width label."""
        return self._native.logical_bits

    @property
    def allocated_bytes(self) -> int:
        return self._native.allocated_bytes

    @property
    def resident_bytes(self) -> int:
        """Host bytes the store holds (estimate); the device copy is not counted."""
        return self._native.resident_bytes

    @property
    def native(self) -> NativeHMatrix:
        return self._native

    def storage_report(self) -> StorageReport:
        n = self._native.order
        return StorageReport(
            dense_fp64_bytes=n * n * 8,
            logical_bits=self.logical_bits,
            allocated_bytes=self.allocated_bytes,
            resident_bytes=self.resident_bytes,
            source_array_bytes=self._source_array_bytes,
        )

    @property
    def reconstruction_bound(self) -> float:
        """Upper bound on max |A - H| entrywise, H evaluated exactly (no
        rounding in the apply)."""
        return self._native.reconstruction_bound

    # -- apply (matrix.py:321-333) ---------------------------------------------------
    def matmul(self, x, *, device: str = "cpu", precision: str = "fp64"):
        """y = H x for x of shape ``(n,)`` or ``(n, p)``, in source order.

        ``device='cuda'`` runs on the GPU; a torch device tensor is applied in
        place on the device and a device tensor is returned, anything else is
        converted to FP64 numpy and a numpy array is returned.
        ``precision='fp32split'`` selects the FP32 split-precision kernels.
        """
        if device not in ("cpu", "cuda"):
            raise ValueError("device must be 'cpu' or 'cuda'")
        if device == "cuda" and _is_device_tensor(x):
            return self._native.apply_device(x, precision=precision)
        if device == "cpu" and precision != "fp64":
            raise ValueError("the host backend computes in fp64 only")
        return self._native.matmul(np.asarray(x, dtype=np.float64), device == "cuda",
                                   precision=precision)

    def __matmul__(self, x):
        """``h @ x``: the host FP64 apply."""
        return self.matmul(x)

    def to_dense(self, *, device: str = "cpu") -> np.ndarray:
        """The n x n FP64 matrix the store represents (an ``8*n*n``-byte array).

        ``device="cuda"`` forms it on the device from the resident store
        (bitwise the host result) and copies it back (matrix.py:309-311
        computes it on the host only)."""
        if device not in ("cpu", "cuda"):
            raise ValueError("device must be 'cpu' or 'cuda'")
        if device == "cuda":
            return self._native.reconstruct_device().cpu().numpy()
        return self._native.reconstruct()

    def save(self, path: str | os.PathLike[str]) -> None:
        self._native.save(os.fspath(path))

    @classmethod
    def load(cls, path: str | os.PathLike[str]) -> "HMatrix":
        return cls(NativeHMatrix.load(os.fspath(path)))


def _is_device_tensor(x) -> bool:
    t = type(x)
    if t.__module__.split(".")[0] != "torch":
        return False
    return bool(getattr(x, "is_cuda", False))


def memory_bytes(value) -> int:
    """Bytes of a NumPy array's data, or the resident host bytes of an H-matrix."""
    if isinstance(value, np.ndarray):
        return int(value.nbytes)
    if isinstance(value, (HMatrix, NativeHMatrix)):
        return value.resident_bytes
    raise TypeError("memory_bytes expects a NumPy array or qhmat H-matrix")
