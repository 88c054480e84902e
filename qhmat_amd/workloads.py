"""Synthetic kernel-matrix workloads and a parallel blockwise constructor.

``build_kernel_hmatrix`` produces the same store as
``HMatrix.from_dense(kernel(points, points), points, ...)`` (same partition,
same block values, same SVD and allocation per block) but evaluates each
block from the points instead of forming the n x n dense matrix, and
factors blocks on a thread pool or, for large stores of the kernels below,
on worker processes (qhmat_amd.blockwork).  Blocks are added in partition
order, so the store is identical to the serial construction.

Workloads (seeds fixed: geometry seed 0, right-hand sides seed 1):
  line_inverse   n=4096 uniform on [0,1], 1/(r+0.4), leaf 32, tol 1e-6, X 4096x4
  square_exp     n=16384 uniform in [0,1]^2, exp(-r/0.2), leaf 64, tol 1e-6
  sphere_bem     n=32768 Fibonacci sphere + N(0,1e-3) jitter, 1/(4 pi r),
                 self term 1/(4 pi 0.5 h), support boxes +-h, leaf 64
"""

from __future__ import annotations

import contextlib
import math
import os
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import Callable, Sequence

import numpy as np

from . import blockwork as B
from ._native import NativeHMatrix, Rectangle, spatial_partition
from .blockwork import (RectangleSingleLayer, SingleLayer, distances, exponential,
                        inverse_distance, inverse_shifted, log_kernel, range_spectrum)
from .matrix import HMatrix, allocate_widths

FULL_MENU = (4, 8, 12, 16, 24, 32)


# -- kernels (blockwork: NumPy only, importable by the worker processes) ------------
KERNELS = {"inverse": inverse_shifted, "exp": exponential, "log": log_kernel}


def single_layer(h: float) -> Callable:
    """1/(4 pi r) off the diagonal, 1/(4 pi 0.5 h) on it (point-panel self term)."""
    return SingleLayer(h)


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
        return contextlib.nullcontext()


PROCESS_MIN_ORDER = 2048  # smaller stores build on threads (no process start-up)


def build_kernel_hmatrix(points: np.ndarray, kernel: Callable, *, tolerance: float,
                         leaf_size: int = 32, admissibility: float = 1.0,
                         bit_widths: Sequence[int] = FULL_MENU, rank_tolerance: float = 0.1,
                         max_rank: int | None = None, truncation_fraction: float = 0.25,
                         arithmetic_reserve: float = 0.02, support_lower=None,
                         support_upper=None, workers: int | None = None,
                         progress: Callable[[int, int], None] | None = None,
                         stress_width: int | None = None, svd: str = "full",
                         processes: bool | None = None, kernel_data: np.ndarray | None = None,
                         keep_factors: bool = False) -> HMatrix:
    """Blockwise, parallel equivalent of from_dense(kernel(points, points), points, ...).

    ``progress(done, total)`` is called after every twentieth of the blocks.

    ``stress_width``: build a store the reference constructor refuses (a
    tolerance whose width budget no allocation meets, such as 1e-8 on the
    2-D and 3-D kernels): each block keeps the rank its truncation allowance
    needs (the SVD tail, as matrix.py:213-238 chooses it) and every component
    gets this width, added through add_low_rank (core.cpp:284-330), which
    checks no budget.  The store is labelled by its use: a stress case of the
    widest codes, not a reference-constructible store.

    ``svd="range"`` factors blocks whose shorter side is at least 256 with
    range_spectrum (randomized, exact residual bookkeeping) instead of the full
    SVD the reference constructor takes: same tolerances and checks, ranks
    equal or one-off, and each block's measured residual still bounds its
    truncation; for the large synthetic stores whose full SVDs would take many
    CPU-minutes.

    ``kernel_data``: per-point rows handed to the kernel instead of the points
    (e.g. panel centres, half-widths and normals), in the points' order; the
    partition still uses the points and support boxes.

    ``keep_factors``: the result also carries ``factors``, block index ->
    (u sqrt(sigma), v sqrt(sigma)) in FP32, the balanced factors before
    quantization (the application harness's unquantized methods)."""
    if svd not in ("full", "range"):
        raise ValueError("svd must be 'full' or 'range'")
    p = np.ascontiguousarray(points, dtype=np.float64)
    lo = p if support_lower is None else np.asarray(support_lower, dtype=np.float64)
    hi = p if support_upper is None else np.asarray(support_upper, dtype=np.float64)
    part = spatial_partition(p, lo, hi, leaf_size, admissibility)
    perm = np.asarray(part.permutation, dtype=np.int64)
    data = p if kernel_data is None else np.ascontiguousarray(kernel_data, dtype=np.float64)
    if data.shape[0] != p.shape[0]:
        raise ValueError("kernel_data must have one row per point")
    clustered = data[perm]
    widths = tuple(bit_widths)
    n = p.shape[0]

    params = dict(tolerance=tolerance, rank_tolerance=rank_tolerance, max_rank=max_rank,
                  truncation_fraction=truncation_fraction,
                  arithmetic_reserve=arithmetic_reserve, stress=stress_width is not None,
                  svd=svd)
    native = NativeHMatrix(n, part.permutation)
    kept = {} if keep_factors else None
    workers = workers or min(16, os.cpu_count() or 1)
    blocks = list(part.blocks)
    if processes is None:
        # own kernels only: a worker imports the kernel by its module name
        processes = (n >= PROCESS_MIN_ORDER and workers > 1
                     and getattr(kernel, "__module__", "") == B.__name__)
    job = B.make_job(clustered, kernel, params)
    if processes:
        # separate interpreters that import qhmat_amd.blockwork only (no
        # HIP runtime, no GIL shared with this process)
        pool, limits = B.WorkerPool(workers, job), contextlib.nullcontext()
    else:
        pool, limits = ThreadPoolExecutor(max_workers=workers), _blas_single_thread()
    with limits, pool:
        # Bounded look-ahead keeps memory flat while preserving insertion order.
        pending = []
        window = 4 * workers
        it = iter(blocks)
        for rect in it:
            pending.append((rect, (pool.submit(_box(rect)) if processes else
                                    pool.submit(B.block_job, _box(rect), job))))
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
                u, sigma, v, residual, quant_budget = payload
                if stress_width is not None:
                    bits = [int(stress_width)] * sigma.size
                elif quant_budget is None:
                    bits = []
                else:
                    bits = allocate_widths(sigma, widths, quant_budget, (u.shape[0], v.shape[0]),
                                           _box(rect))
                native.add_low_rank(rect, u, sigma, v, bits, len(widths), residual)
                if kept is not None and sigma.size:
                    root = np.sqrt(sigma)
                    kept[done - 1] = ((u * root).astype(np.float32), (v * root).astype(np.float32))
            nxt = next(it, None)
            if nxt is not None:
                pending.append((nxt, (pool.submit(_box(nxt)) if processes else
                                   pool.submit(B.block_job, _box(nxt), job))))
    native.finalize()
    out = HMatrix(native, source_array_bytes=n * n * 8)
    if kept is not None:
        out.factors = kept
    return out


def _box(rect: Rectangle) -> tuple:
    return (rect.row0, rect.row1, rect.col0, rect.col1)


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


def batch_workload(count: int = 64, seed: int = 4, n_min: int = 1024, n_max: int = 16384,
                   progress: Callable[[str], None] | None = None):
    """Stores of uneven order, kernel, dimension and tolerance for the batched
    apply: n log-uniform in [n_min, n_max], kernels 1/(r+0.4), exp(-r/0.2),
    log(r+1e-2) on 1-D, 2-D or 3-D uniform points, tolerances drawn from
    {1e-4, 1e-6, 1e-8}, right-hand-side counts uniform in 1..16.  A store the
    reference constructor refuses (its width budget infeasible: every 1e-8
    store, and some larger 1e-6 ones) is built as a labelled stress store
    (stress_width=32).  Returns (natives, ks, shapes)."""
    from .blockwork import InfeasibleError

    rng = np.random.default_rng(seed)
    kernels = [inverse_shifted, exponential, log_kernel]
    natives, ks, shapes = [], [], []
    for i in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        dim = int(rng.integers(1, 4))
        kern = kernels[int(rng.integers(0, 3))]
        tol = (1e-4, 1e-6, 1e-8)[int(rng.integers(0, 3))]
        pts = uniform_points(n, dim, seed=100 + i)
        stress = tol <= 1e-8
        if not stress:
            try:
                h = build_kernel_hmatrix(pts, kern, tolerance=tol, leaf_size=64, svd="range")
            except InfeasibleError:
                stress = True
        if stress:
            h = build_kernel_hmatrix(pts, kern, tolerance=tol, leaf_size=64, stress_width=32,
                                     svd="range")
        natives.append(h.native)
        ks.append(int(rng.integers(1, 17)))
        shapes.append({"n": n, "dim": dim, "kernel": kern.__name__, "tol": tol,
                       "stress": stress})
        if progress is not None:
            progress(f"store {i + 1}/{count}: n={n} {kern.__name__} {dim}-D tol {tol}"
                     + (" (stress)" if stress else ""))
    return natives, ks, shapes


def serving_workload(count: int = 16, n_min: int = 2048, n_max: int = 32768,
                     progress: Callable[[str], None] | None = None):
    """Resident stores of the served load: orders spaced geometrically from
    n_min to n_max, kernels 1/(r+0.4), exp(-r/0.2), log(r+1e-2) in turn on
    1-D, 2-D and 3-D uniform points, tolerance 1e-6, leaf 64 (a store the
    reference constructor refuses at 1e-6 is built as a labelled stress
    store).  Returns (name -> native, shapes)."""
    from .blockwork import InfeasibleError

    kernels = [inverse_shifted, exponential, log_kernel]
    natives, shapes = {}, []
    for i in range(count):
        n = int(round(n_min * (n_max / n_min) ** (i / max(count - 1, 1))))
        kern, dim = kernels[i % 3], 1 + (i // 3) % 3
        pts = uniform_points(n, dim, seed=300 + i)
        stress = False
        try:
            h = build_kernel_hmatrix(pts, kern, tolerance=1e-6, leaf_size=64, svd="range")
        except InfeasibleError:  # labelled stress store, as in batch_workload
            stress = True
            h = build_kernel_hmatrix(pts, kern, tolerance=1e-6, leaf_size=64, stress_width=32,
                                     svd="range")
        name = f"{kern.__name__}_{dim}d_{n}"
        natives[name] = h.native
        shapes.append({"name": name, "n": n, "dim": dim, "kernel": kern.__name__,
                       "stress": stress})
        if progress is not None:
            progress(f"store {i + 1}/{count}: {name}" + (" (stress)" if stress else ""))
    return natives, shapes


# -- the reference application suite's two cases (experiments: geometric_matrices.py) --
def box_panels(subdivisions: int, extent=(1.0, 1.0, 1.0)):
    """Axis-aligned square panels covering a closed box: 6 subdivisions^2
    panels, subdivisions x subdivisions per face; returns the panel rows
    (centre, half-widths, normal axis) as kernel data."""
    ext = np.asarray(extent, dtype=np.float64)
    rows = []
    for axis in range(3):
        a, b = [d for d in range(3) if d != axis]
        ga = (np.arange(subdivisions) + 0.5) * ext[a] / subdivisions - ext[a] / 2
        gb = (np.arange(subdivisions) + 0.5) * ext[b] / subdivisions - ext[b] / 2
        ua, ub = np.meshgrid(ga, gb, indexing="ij")
        for side in (-1.0, 1.0):
            r = np.zeros((subdivisions * subdivisions, 7))
            r[:, axis] = side * ext[axis] / 2
            r[:, a], r[:, b] = ua.ravel(), ub.ravel()
            r[:, 3 + a] = ext[a] / (2 * subdivisions)
            r[:, 3 + b] = ext[b] / (2 * subdivisions)
            r[:, 6] = axis
            rows.append(r)
    return np.vstack(rows)


@dataclass
class SuiteCase:
    """A case of the reference's application suite, rebuilt with this
    library's constructor (leaf 32, admissibility 1, widths 4..32)."""
    name: str
    points: np.ndarray
    kernel: Callable
    support_lower: np.ndarray | None = None
    support_upper: np.ndarray | None = None
    kernel_data: np.ndarray | None = None

    def build(self, tolerance: float, keep_factors: bool = True) -> HMatrix:
        return build_kernel_hmatrix(self.points, self.kernel, tolerance=tolerance, leaf_size=32,
                                    support_lower=self.support_lower,
                                    support_upper=self.support_upper,
                                    kernel_data=self.kernel_data, keep_factors=keep_factors,
                                    processes=False)


def suite_inverse_distance_2d(n: int = 2048, seed: int = 7) -> SuiteCase:
    """kernel_inverse_distance_2d_uniform_n2048: 1/r on n points uniform in
    [-0.5, 0.5]^2, zero diagonal."""
    pts = np.random.default_rng(seed).uniform(-0.5, 0.5, size=(n, 2))
    return SuiteCase(f"inverse_distance_2d_uniform_n{n}", pts, inverse_distance)


def suite_cube_bem(subdivisions: int = 16) -> SuiteCase:
    """bem3d_inverse_cube_n1536: constant-panel collocation of the 1/(4 pi r)
    single layer on the unit cube's surface, 6 x 16^2 panels, each panel's
    integral in closed form (self terms included), support boxes = panels."""
    data = box_panels(subdivisions)
    centres, half = data[:, :3], data[:, 3:6]
    return SuiteCase(f"cube_bem_n{data.shape[0]}", centres, RectangleSingleLayer(),
                     support_lower=centres - half, support_upper=centres + half, kernel_data=data)
