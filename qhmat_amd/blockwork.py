"""Per-block numerical work of the constructors, NumPy only.

Kernel evaluation, the spectrum of a block (LAPACK SVD, or a randomized
range finder for large blocks), rank selection and the measured truncation
residual.  The bit allocation and packing that follow are native
(``_native.allocate``, ``NativeHMatrix.add_low_rank``) and run in the
calling process.  This module imports neither the native library nor the
rest of the package, so the blockwise constructor can run it in spawned
worker processes that never load the HIP runtime.

Reference: python/qhmat/matrix.py:213-245 (per-block SVD, truncation
allowance, rank choice and residual check of ``HMatrix.from_dense``).
"""

from __future__ import annotations

import math

import numpy as np


class InfeasibleError(ValueError):
    """Raised when an off-diagonal block's error budget cannot be met: no
    rank within the truncation allowance, or no width allocation within the
    quantization budget."""


# -- kernels: each maps (row points, column points) -> block values -----------------
def distances(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    """|p_i - q_j| with the squares summed over the coordinates in order
    (the same sums, bit for bit, as a sum over the last axis of p - q)."""
    s = None
    for d in range(p.shape[1]):
        t = p[:, d, None] - q[None, :, d]
        t = t * t
        s = t if s is None else s + t
    return np.sqrt(s)


def inverse_shifted(p, q):
    """1 / (r + 0.4)."""
    return 1.0 / (distances(p, q) + 0.4)


def exponential(p, q):
    """exp(-r / 0.2)."""
    return np.exp(-distances(p, q) / 0.2)


def log_kernel(p, q):
    """log(r + 1e-2)."""
    return np.log(distances(p, q) + 1e-2)


class SingleLayer:
    """1/(4 pi r) off the diagonal, 1/(4 pi 0.5 h) on it (point-panel self term)."""

    def __init__(self, h: float):
        self.h = float(h)

    def __call__(self, p, q):
        r = distances(p, q)
        with np.errstate(divide="ignore"):
            out = 1.0 / (4.0 * math.pi * r)
        out[r == 0.0] = 1.0 / (4.0 * math.pi * 0.5 * self.h)
        return out


def inverse_distance(p, q):
    """1 / r between distinct points, 0 where r = 0 (the omitted self term of
    the point-kernel suite cases)."""
    r = distances(p, q)
    with np.errstate(divide="ignore"):
        out = 1.0 / r
    out[r == 0.0] = 0.0
    return out


def _rect_primitive(u, v, z):
    """F(u, v) with d2F/du dv = 1 / sqrt(u^2 + v^2 + z^2) (z >= 0):
    u asinh(v / sqrt(u^2 + z^2)) + v asinh(u / sqrt(v^2 + z^2))
    - z atan(u v / (z sqrt(u^2 + v^2 + z^2))), each term taken as its limit 0
    where its argument is undefined (u = z = 0, v = z = 0, or z = 0)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        ru = np.sqrt(u * u + z * z)
        rv = np.sqrt(v * v + z * z)
        r = np.sqrt(u * u + v * v + z * z)
        t1 = np.where(ru > 0, u * np.arcsinh(v / np.where(ru > 0, ru, 1.0)), 0.0)
        t2 = np.where(rv > 0, v * np.arcsinh(u / np.where(rv > 0, rv, 1.0)), 0.0)
        t3 = np.where(z > 0, z * np.arctan(u * v / np.where(z > 0, z * r, 1.0)), 0.0)
    return t1 + t2 - t3


class RectangleSingleLayer:
    """Collocation single layer on axis-aligned rectangular panels: entry
    (i, j) = (1 / 4 pi) times the integral of 1 / |x_i - y| over panel j, x_i
    the centre of panel i, by the closed form of that integral (corners of
    _rect_primitive), the self terms included.  Rows of the kernel data:
    centre (3), half-widths (3, zero along the normal), normal axis (1)."""

    def __call__(self, p, q):
        out = np.empty((p.shape[0], q.shape[0]))
        normals = q[:, 6].astype(np.int64)
        for axis in range(3):
            cols = np.flatnonzero(normals == axis)
            if cols.size == 0:
                continue
            a, b = [d for d in range(3) if d != axis]
            src = q[cols]
            z = np.abs(src[None, :, axis] - p[:, None, axis])
            du = src[None, :, a] - p[:, None, a]
            dv = src[None, :, b] - p[:, None, b]
            hu, hv = src[None, :, 3 + a], src[None, :, 3 + b]
            total = np.zeros_like(z)
            for su in (-1.0, 1.0):
                for sv in (-1.0, 1.0):
                    total += su * sv * _rect_primitive(du + su * hu, dv + sv * hv, z)
            out[:, cols] = total / (4.0 * math.pi)
        return out


# -- spectra ----------------------------------------------------------------------
RANGE_MIN_SIDE = 256  # blocks with a shorter side take the full SVD
RANGE_STEP = 32


def full_spectrum(block: np.ndarray):
    """(u, sigma, vt, tail) of the LAPACK SVD; tail[r] is the Frobenius norm
    of the part discarded at rank r."""
    u, sigma, vt = np.linalg.svd(block, full_matrices=False)
    tail = np.sqrt(np.concatenate([np.cumsum(sigma[::-1] ** 2)[::-1], [0.0]]))
    return u, sigma, vt, tail


def _orthonormal(y: np.ndarray, q: np.ndarray | None) -> np.ndarray:
    # two rounds of block Gram-Schmidt against q, then a thin QR
    if q is not None:
        for _ in range(2):
            y = y - q @ (q.T @ y)
    return np.linalg.qr(y)[0]


def range_spectrum(block: np.ndarray, allowance: float, seed: int = 0):
    """(u, sigma, vt, tail) of a blocked randomized range finder (randQB,
    Martinsson and Voronin 2016, with one power iteration per step): grows an
    orthonormal basis q of 32 columns a step, keeping the exact residual
    e = block - q q^T block, until ||e||_F <= allowance / 2, then takes the SVD
    of q^T block.  tail[r]^2 = ||e||_F^2 + sum_{i>r} sigma_i^2 exactly for the
    factors returned, so rank selection sees the true discarded norm of the
    factorization it keeps.  Falls back to the full SVD for small blocks and
    for blocks whose basis would pass half their shorter side."""
    m, n = block.shape
    limit = min(m, n) // 2
    if min(m, n) < RANGE_MIN_SIDE:
        return full_spectrum(block)
    rng = np.random.default_rng(seed)
    e = np.array(block, dtype=np.float64, copy=True)
    q = None
    rows = []
    while True:
        y = e @ rng.standard_normal((n, RANGE_STEP))
        y = _orthonormal(y, q)
        y = _orthonormal(e @ (e.T @ y), q)  # one power iteration
        b = y.T @ e
        e -= y @ b
        q = y if q is None else np.hstack([q, y])
        rows.append(b)
        residual = float(np.linalg.norm(e))
        if residual <= 0.5 * allowance:
            break
        if q.shape[1] + RANGE_STEP > limit:
            return full_spectrum(block)
    ub, sigma, vt = np.linalg.svd(np.vstack(rows), full_matrices=False)
    tail = np.sqrt(residual ** 2 + np.concatenate([np.cumsum(sigma[::-1] ** 2)[::-1], [0.0]]))
    return q @ ub, sigma, vt, tail


def _label(rect) -> str:
    return f"block ({rect[0]}:{rect[1]}, {rect[2]}:{rect[3]})"


# -- truncation ---------------------------------------------------------------------
def truncate(block: np.ndarray, rect, *, tolerance: float, rank_tolerance: float,
             max_rank: int | None, truncation_fraction: float, arithmetic_reserve: float,
             spectrum=None):
    """Rank and factors of one off-diagonal block (matrix.py:213-238).

    ``rect`` = (row0, row1, col0, col1), for messages.  Returns (u, sigma, v,
    residual, quant_budget): the kept factors, the measured max-norm
    truncation residual and the budget left for quantization; rank 0 (and
    quant_budget None) for an all-zero block.  Raises InfeasibleError when
    the truncation allowance cannot be met.

    ``spectrum(block, allowance) -> (u, sigma, vt, tail)`` replaces the full
    SVD when given (range_spectrum)."""
    budget = tolerance * float(np.max(np.abs(block)))
    m, n = block.shape
    if budget == 0.0:
        return np.zeros((m, 0)), np.empty(0), np.zeros((n, 0)), 0.0, None
    if spectrum is None:
        u, sigma, vt, tail = full_spectrum(block)
        allowance = min(rank_tolerance * tail[0], truncation_fraction * budget)
    else:
        allowance = min(rank_tolerance * float(np.linalg.norm(block)),
                        truncation_fraction * budget)
        u, sigma, vt, tail = spectrum(block, allowance)
    cap = sigma.size if max_rank is None else min(max_rank, sigma.size)
    fits = np.nonzero(tail[1:cap + 1] <= allowance)[0]
    if fits.size == 0:
        raise InfeasibleError(f"{_label(rect)} cannot satisfy the truncation allowance")
    rank = min(int(fits[0]) + 1, int(np.count_nonzero(sigma > 0)))
    sigma = sigma[:rank]
    u = u[:, :rank]
    v = vt[:rank, :].T
    residual = float(np.max(np.abs(block - (u * sigma) @ v.T)))
    quant_budget = budget * (1 - arithmetic_reserve) - residual
    if residual > truncation_fraction * budget or quant_budget < 0:
        raise InfeasibleError("measured truncation error exceeds the block allowance")
    return u, sigma, v, residual, quant_budget


def tail_rank(block: np.ndarray, *, tolerance: float, rank_tolerance: float,
              truncation_fraction: float, spectrum=None):
    """(u, sigma, v, residual) at the rank the truncation allowance needs,
    with no allocation budget (stress stores of the widest codes)."""
    budget = tolerance * float(np.max(np.abs(block)))
    m, n = block.shape
    if budget == 0.0:
        return np.zeros((m, 0)), np.empty(0), np.zeros((n, 0)), 0.0
    allowance = min(rank_tolerance * float(np.linalg.norm(block)), truncation_fraction * budget)
    u, sigma, vt, tail = spectrum(block, allowance) if spectrum else full_spectrum(block)
    fits = np.nonzero(tail[1:] <= allowance)[0]
    rank = int(fits[0]) + 1 if fits.size else sigma.size
    rank = min(rank, int(np.count_nonzero(sigma > 0)))
    u, sigma, v = u[:, :rank], sigma[:rank], vt[:rank, :].T
    residual = float(np.max(np.abs(block - (u * sigma) @ v.T)))
    return u, sigma, v, residual


# -- blockwise construction workers ----------------------------------------------------
def _one_blas_thread() -> None:
    try:
        from threadpoolctl import threadpool_limits

        threadpool_limits(1)
    except Exception:  # pragma: no cover - threadpoolctl is optional
        pass


def make_job(points: np.ndarray, kernel, params: dict) -> dict:
    return {"points": points, "kernel": kernel, "params": dict(params)}


def block_job(rect, job: dict) -> tuple:
    """One block, rect = (row0, row1, col0, col1), of `job` (make_job):
    ("dense", values) for a diagonal block, ("lowrank", (u, sigma, v,
    residual, quant_budget)) otherwise (quant_budget None for a stress block
    or an all-zero one)."""
    pts, kernel, prm = job["points"], job["kernel"], job["params"]
    r0, r1, c0, c1 = rect
    block = np.ascontiguousarray(kernel(pts[r0:r1], pts[c0:c1]))
    if r0 == c0 and r1 == c1:
        return ("dense", block)
    spectrum = None
    if prm["svd"] == "range":
        seed = (r0 * 1_000_003 + c0) & 0xFFFFFFFF

        def spectrum(a, allowance):
            return range_spectrum(a, allowance, seed)
    if prm["stress"]:
        u, sigma, v, residual = tail_rank(block, tolerance=prm["tolerance"],
                                          rank_tolerance=prm["rank_tolerance"],
                                          truncation_fraction=prm["truncation_fraction"],
                                          spectrum=spectrum)
        return ("lowrank", (u, sigma, v, residual, None))
    return ("lowrank", truncate(block, rect, tolerance=prm["tolerance"],
                                rank_tolerance=prm["rank_tolerance"], max_rank=prm["max_rank"],
                                truncation_fraction=prm["truncation_fraction"],
                                arithmetic_reserve=prm["arithmetic_reserve"],
                                spectrum=spectrum))


# Worker processes: started as `python -c "from qhmat_amd import blockwork;
# blockwork.serve()"`, so they import NumPy and this module only (no HIP
# runtime, no copy of the caller's state, no re-run of the caller's main
# script).  Messages are length-prefixed pickles on the worker's stdin and
# stdout: first the job, then one block rectangle per task; each answer is
# ("ok", result) or ("error", exception).
def _send(stream, obj) -> None:
    import pickle
    import struct

    data = pickle.dumps(obj, protocol=pickle.HIGHEST_PROTOCOL)
    stream.write(struct.pack("<Q", len(data)))
    stream.write(data)
    stream.flush()


def _receive(stream):
    import pickle
    import struct

    head = stream.read(8)
    if len(head) < 8:
        raise EOFError("worker stream closed")
    (size,) = struct.unpack("<Q", head)
    return pickle.loads(stream.read(size))


def serve() -> None:
    """Worker main loop (see above); ends when its input closes."""
    import sys

    _one_blas_thread()
    inp, out = sys.stdin.buffer, sys.stdout.buffer
    job = _receive(inp)
    while True:
        try:
            rect = _receive(inp)
        except EOFError:
            return
        try:
            answer = ("ok", block_job(rect, job))
        except Exception as e:  # handed back to the caller, raised there
            answer = ("error", e)
        _send(out, answer)


class WorkerPool:
    """`workers` worker processes running block_job over one job.  Tasks go
    to the workers in turn and results come back in submission order
    (submit returns a handle whose result() blocks)."""

    def __init__(self, workers: int, job: dict):
        import os
        import subprocess
        import sys

        root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
        env = dict(os.environ)
        env["PYTHONPATH"] = root + os.pathsep + env.get("PYTHONPATH", "")
        env["OMP_NUM_THREADS"] = env["OPENBLAS_NUM_THREADS"] = "1"
        code = "from qhmat_amd import blockwork; blockwork.serve()"
        self._procs = [subprocess.Popen([sys.executable, "-c", code], stdin=subprocess.PIPE,
                                        stdout=subprocess.PIPE, env=env)
                       for _ in range(max(1, workers))]
        self._next = 0
        self._unread = [0] * len(self._procs)  # answers not yet received, per worker
        try:
            for p in self._procs:
                _send(p.stdin, job)
        except BaseException:
            for p in self._procs:
                p.kill()
                p.wait()
            raise

    def submit(self, rect):
        i = self._next
        self._next = (self._next + 1) % len(self._procs)
        _send(self._procs[i].stdin, rect)
        self._unread[i] += 1
        return _Pending(self, i)

    def close(self) -> None:
        """Ends the workers: at once for one with answers never read (the
        caller stopped early, e.g. on an error), else when its input closes."""
        for i, p in enumerate(self._procs):
            if self._unread[i]:
                p.kill()
            try:
                p.stdin.close()
            except OSError:
                pass
        for p in self._procs:
            try:
                p.wait(timeout=30)
            except Exception:
                p.kill()
                p.wait()
        for p in self._procs:
            if p.stdout is not None:
                p.stdout.close()
        self._procs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class _Pending:
    def __init__(self, pool: WorkerPool, index: int):
        self._pool, self._index = pool, index

    def result(self):
        status, value = _receive(self._pool._procs[self._index].stdout)
        self._pool._unread[self._index] -= 1
        if status == "error":
            raise value
        return value
