"""Synthetic kernel-matrix workloads and a parallel blockwise constructor.

``build_kernel_hmatrix`` produces the same store as
``HMatrix.from_dense(kernel(points, points), points, ...)`` (same partition,
same block values, same SVD and allocation per block) but evaluates each
block from the points instead of forming the n x n dense matrix, and
factors blocks on a thread pool.  Blocks are added in partition order, so the
store is identical to the serial construction.

Workloads (seeds fixed: geometry seed 0, right-hand sides seed 1):
  line_inverse   n=4096 uniform on [0,1], 1/(r+0.4), leaf 32, tol 1e-6, X 4096x4
  square_exp     n=16384 uniform in [0,1]^2, exp(-r/0.2), leaf 64, tol 1e-6
  sphere_bem     n=32768 Fibonacci sphere + N(0,1e-3) jitter, 1/(4 pi r),
                 self term 1/(4 pi 0.5 h), support boxes +-h, leaf 64
"""

from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Callable, Sequence

import numpy as np

from ._native import NativeHMatrix, Rectangle, spatial_partition
from .matrix import HMatrix, factor_block

FULL_MENU = (4, 8, 12, 16, 24, 32)


# -- kernels: each maps (rows points, cols points) -> block values -----------------
def distances(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    diff = p[:, None, :] - q[None, :, :]
    return np.sqrt(np.sum(diff * diff, axis=2))


def inverse_shifted(p, q):
    """1 / (r + 0.4)."""
    return 1.0 / (distances(p, q) + 0.4)


def exponential(p, q):
    """exp(-r / 0.2)."""
    return np.exp(-distances(p, q) / 0.2)


def log_kernel(p, q):
    """log(r + 1e-2)."""
    return np.log(distances(p, q) + 1e-2)


KERNELS = {"inverse": inverse_shifted, "exp": exponential, "log": log_kernel}


def single_layer(h: float) -> Callable:
    """1/(4 pi r) off the diagonal, 1/(4 pi 0.5 h) on it (point-panel self term)."""

    def kernel(p, q):
        r = distances(p, q)
        with np.errstate(divide="ignore"):
            out = 1.0 / (4.0 * math.pi * r)
        out[r == 0.0] = 1.0 / (4.0 * math.pi * 0.5 * h)
        return out

    return kernel


# -- geometry -------------------------------------------------------------------------
def uniform_points(n: int, dim: int, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).uniform(0.0, 1.0, size=(n, dim))


def fibonacci_sphere(n: int, jitter: float = 1e-3, seed: int = 0) -> np.ndarray:
    i = np.arange(n, dtype=np.float64) + 0.5
    z = 1.0 - 2.0 * i / n
    rho = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    theta = math.pi * (3.0 - math.sqrt(5.0)) * i
    pts = np.column_stack((rho * np.cos(theta), rho * np.sin(theta), z))
    return pts + np.random.default_rng(seed).normal(0.0, jitter, size=pts.shape)


# -- construction ------------------------------------------------------------------
def _blas_single_thread():
    try:
        from threadpoolctl import threadpool_limits

        return threadpool_limits(1)
    except Exception:  # pragma: no cover - threadpoolctl is optional
        import contextlib

        return contextlib.nullcontext()


def build_kernel_hmatrix(points: np.ndarray, kernel: Callable, *, tolerance: float,
                         leaf_size: int = 32, admissibility: float = 1.0,
                         bit_widths: Sequence[int] = FULL_MENU, rank_tolerance: float = 0.1,
                         max_rank: int | None = None, truncation_fraction: float = 0.25,
                         arithmetic_reserve: float = 0.02, support_lower=None,
                         support_upper=None, workers: int | None = None,
                         progress: Callable[[int, int], None] | None = None,
                         stress_width: int | None = None) -> HMatrix:
    """Blockwise, parallel equivalent of from_dense(kernel(points, points), points, ...).

    ``progress(done, total)`` is called after every twentieth of the blocks.

    ``stress_width``: build a store the reference constructor refuses (a
    tolerance whose width budget no allocation meets, such as 1e-8 on the
    2-D and 3-D kernels): each block keeps the rank its truncation allowance
    needs (the SVD tail, as matrix.py:213-238 chooses it) and every component
    gets this width, added through add_low_rank (core.cpp:284-330), which
    checks no budget.  The store is labelled by its use: a stress case of the
    widest codes, not a reference-constructible store."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    lo = p if support_lower is None else np.asarray(support_lower, dtype=np.float64)
    hi = p if support_upper is None else np.asarray(support_upper, dtype=np.float64)
    part = spatial_partition(p, lo, hi, leaf_size, admissibility)
    perm = np.asarray(part.permutation, dtype=np.int64)
    clustered = p[perm]
    widths = tuple(bit_widths)
    n = p.shape[0]

    def work(rect: Rectangle):
        block = np.ascontiguousarray(kernel(clustered[rect.row0:rect.row1],
                                            clustered[rect.col0:rect.col1]))
        if rect.row0 == rect.col0 and rect.row1 == rect.col1:
            return ("dense", block)
        if stress_width is not None:
            return ("lowrank", _tail_rank_factors(block, tolerance, rank_tolerance,
                                                  truncation_fraction, stress_width))
        return ("lowrank", factor_block(block, rect, tolerance=tolerance, widths=widths,
                                        rank_tolerance=rank_tolerance, max_rank=max_rank,
                                        truncation_fraction=truncation_fraction,
                                        arithmetic_reserve=arithmetic_reserve))

    native = NativeHMatrix(n, part.permutation)
    workers = workers or min(32, os.cpu_count() or 1)
    with _blas_single_thread(), ThreadPoolExecutor(max_workers=workers) as pool:
        # Bounded look-ahead keeps memory flat while preserving insertion order.
        pending = []
        blocks = list(part.blocks)
        window = 4 * workers
        it = iter(blocks)
        for rect in it:
            pending.append((rect, pool.submit(work, rect)))
            if len(pending) >= window:
                break
        total, done, step = len(blocks), 0, max(1, len(blocks) // 20)
        while pending:
            rect, fut = pending.pop(0)
            done += 1
            if progress is not None and (done % step == 0 or done == total):
                progress(done, total)
            kind, payload = fut.result()
            if kind == "dense":
                native.add_dense(rect, payload)
            else:
                u, sigma, v, bits, residual = payload
                native.add_low_rank(rect, u, sigma, v, bits, len(widths), residual)
            nxt = next(it, None)
            if nxt is not None:
                pending.append((nxt, pool.submit(work, nxt)))
    native.finalize()
    return HMatrix(native, source_array_bytes=n * n * 8)


def _tail_rank_factors(block, tolerance, rank_tolerance, truncation_fraction, width):
    """(u, sigma, v, bits, residual) at the rank the truncation allowance
    needs, every component `width` bits (no allocation budget)."""
    budget = tolerance * float(np.max(np.abs(block)))
    m, n = block.shape
    if budget == 0.0:
        return np.zeros((m, 0)), np.empty(0), np.zeros((n, 0)), [], 0.0
    u, sigma, vt = np.linalg.svd(block, full_matrices=False)
    tail = np.sqrt(np.concatenate([np.cumsum(sigma[::-1] ** 2)[::-1], [0.0]]))
    allowance = min(rank_tolerance * tail[0], truncation_fraction * budget)
    fits = np.nonzero(tail[1:] <= allowance)[0]
    rank = int(fits[0]) + 1 if fits.size else sigma.size
    rank = min(rank, int(np.count_nonzero(sigma > 0)))
    u, sigma, v = u[:, :rank], sigma[:rank], vt[:rank, :].T
    residual = float(np.max(np.abs(block - (u * sigma) @ v.T)))
    return u, sigma, v, [int(width)] * rank, residual


@dataclass
class Workload:
    name: str
    points: np.ndarray
    kernel: Callable
    leaf_size: int
    tolerance: float
    rhs: int
    support_lower: np.ndarray | None = None
    support_upper: np.ndarray | None = None
    note: str = ""

    def build(self, workers: int | None = None, progress=None,
              stress_width: int | None = None) -> HMatrix:
        return build_kernel_hmatrix(self.points, self.kernel, tolerance=self.tolerance,
                                    leaf_size=self.leaf_size, support_lower=self.support_lower,
                                    support_upper=self.support_upper, workers=workers,
                                    progress=progress, stress_width=stress_width)

    def rows(self, index) -> np.ndarray:
        """Rows `index` (source order) of the dense kernel matrix, evaluated
        from the points: the exact A x of selected rows without the n x n
        matrix."""
        return self.kernel(self.points[np.asarray(index)], self.points)

    def dense(self) -> np.ndarray:
        return self.kernel(self.points, self.points)

    def rhs_matrix(self, k: int | None = None, seed: int = 1) -> np.ndarray:
        return np.random.default_rng(seed).normal(size=(self.points.shape[0], k or self.rhs))


def line_inverse(n: int = 4096) -> Workload:
    return Workload("line_inverse", uniform_points(n, 1), inverse_shifted, 32, 1e-6, 4)


def square_exp(n: int = 16384, tolerance: float = 1e-6) -> Workload:
    return Workload("square_exp", uniform_points(n, 2), exponential, 64, tolerance, 1,
                    note="tolerance 1e-8 is infeasible for the reference constructor; "
                         "the tightest feasible tolerance 1e-6 is used")


def sphere_bem(n: int = 32768, tolerance: float = 1e-6) -> Workload:
    pts = fibonacci_sphere(n)
    h = math.sqrt(4.0 * math.pi / n)
    return Workload("sphere_bem", pts, single_layer(h), 64, tolerance, 64,
                    support_lower=pts - h, support_upper=pts + h,
                    note="mean spacing h = sqrt(4 pi / n)")
