"""Shared builders for the tests: small stores shaped like the reference's own
test matrices (tests/test_core.py, tests/test_cuda.py) plus ragged and
user-partition cases."""

from __future__ import annotations

import os
import tempfile

import numpy as np

import qhmat_amd as q
from oracle import oracle as O


def two_by_two():
    """test_core.py:44-60: [[2, 1], [1, 3]] with 8-bit rank-1 off-diagonals."""
    m = q.NativeHMatrix(2, [0, 1])
    m.add_dense(q.Rectangle(0, 1, 0, 1), np.array([[2.0]]))
    m.add_dense(q.Rectangle(1, 2, 1, 2), np.array([[3.0]]))
    for row, col in [(0, 1), (1, 0)]:
        m.add_low_rank(q.Rectangle(row, row + 1, col, col + 1), np.ones((1, 1)),
                       np.array([1.0]), np.ones((1, 1)), [8], 3, 0.0)
    m.finalize()
    return m


def orthogonal_factors(bit: int):
    """test_core.py:170-192: 17x17, two rank-3 blocks at one width."""
    rng = np.random.default_rng(200 + bit)
    u, _ = np.linalg.qr(rng.normal(size=(9, 3)))
    v, _ = np.linalg.qr(rng.normal(size=(8, 3)))
    sigma = np.array([3.0, 1.0, 0.2])
    native = q.NativeHMatrix(17, list(range(17)))
    native.add_dense(q.Rectangle(0, 9, 0, 9), np.eye(9))
    native.add_dense(q.Rectangle(9, 17, 9, 17), np.eye(8))
    native.add_low_rank(q.Rectangle(0, 9, 9, 17), u, sigma, v, [bit] * 3, 1, 0.0)
    native.add_low_rank(q.Rectangle(9, 17, 0, 9), v, sigma, u, [bit] * 3, 1, 0.0)
    native.finalize()
    a = (u * sigma) @ v.T
    expected = np.eye(17)
    expected[:9, 9:] = a
    expected[9:, :9] = a.T
    return native, expected, rng


def line_kernel(n: int, leaf: int, tol: float, widths=(4, 8, 16, 32)):
    """test_core.py:70-83 / test_cuda.py:8-23 style: 1/(|dx|+0.4) on a line."""
    pts = np.linspace(0.0, 1.0, n)[:, None]
    src = 1.0 / (0.4 + np.abs(pts - pts.T))
    return q.HMatrix.from_dense(src, pts, tolerance=tol, leaf_size=leaf, bit_widths=widths), src


def ragged_square(n: int = 1500, leaf: int = 40, tol: float = 1e-6, seed: int = 3):
    """2-D exp(-r/0.2) with an order that is not leaf * 2^L: uneven slabs."""
    from qhmat_amd import workloads as W

    pts = np.random.default_rng(seed).uniform(size=(n, 2))
    return W.build_kernel_hmatrix(pts, W.exponential, tolerance=tol, leaf_size=leaf)


def user_partition(seed: int = 5):
    """from_partition with large dense diagonal blocks and uneven rectangles."""
    rng = np.random.default_rng(seed)
    n = 300
    x = np.sort(rng.uniform(size=n))
    src = 1.0 / (np.abs(x[:, None] - x[None, :]) + 0.3)
    cuts = [0, 70, 71, 200, 300]
    rects = []
    for i in range(4):
        for j in range(4):
            rects.append(q.Rectangle(cuts[i], cuts[i + 1], cuts[j], cuts[j + 1]))
    return q.HMatrix.from_partition(src, list(range(n)), rects, tolerance=1e-7), src


def oracle_of(native) -> O.OracleStore:
    """The oracle's view of a product store, through a QHMAT001 file."""
    d = tempfile.mkdtemp(prefix="qhmat-oracle-")
    path = os.path.join(d, "store.qhmat")
    native.save(path)
    try:
        return O.OracleStore.from_file(path, verify_checksum=os.path.getsize(path) < 4_000_000)
    finally:
        os.remove(path)
        os.rmdir(d)


def rel_err(a, b) -> float:
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


# Right-hand sides per pass from which the device passes form the code
# products on the matrix cores (the wide passes; 1, 2 or 4 per pass use the
# vector ALUs).  Within one of the two paths a column's bits do not depend on
# the pass width; across them results agree to the precision's tolerance
# (same_result).
MATRIX_MIN_KB = 16
FP32_REL_TOL = 1e-6


def same_result(got, want, precision: str) -> None:
    """The same apply through the other arithmetic path: FP64 within rtol =
    atol = 1e-12 (scaled), FP32-split within relative error 1e-6."""
    got, want = np.asarray(got), np.asarray(want)
    if precision == "fp64":
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * float(np.abs(want).max()))
    else:
        assert rel_err(got, want) <= FP32_REL_TOL


def wide_strips(tiles: int = 68, rank: int = 3, seed: int = 7):
    """Row strips of 64: a dense 64x64 diagonal block between two low-rank
    blocks spanning the rest of the strip.  The first and last strips hold
    blocks wider than 64 column tiles (their partial sums come in several
    runs); every strip also has a high-rank 64x64 block split over items."""
    rng = np.random.default_rng(seed)
    n = 64 * tiles
    native = q.NativeHMatrix(n, list(rng.permutation(n)))
    menu = [4, 8, 12, 16, 24, 32]

    def low_rank(r0, r1, c0, c1, r):
        u = np.linalg.qr(rng.normal(size=(r1 - r0, r)))[0]
        v = np.linalg.qr(rng.normal(size=(c1 - c0, r)))[0]
        sigma = np.sort(rng.uniform(0.5, 2.0, size=r))[::-1]
        bits = list(rng.choice(menu, size=r))
        native.add_low_rank(q.Rectangle(r0, r1, c0, c1), u, sigma, v, bits, len(menu), 0.0)

    for i in range(tiles):
        r0, r1 = 64 * i, 64 * i + 64
        native.add_dense(q.Rectangle(r0, r1, r0, r1), rng.normal(size=(64, 64)))
        if i > 0:
            if i == 1:
                low_rank(r0, r1, 0, 64, 40)  # 64x64 with many 32-bit-capable components
            else:
                low_rank(r0, r1, 0, r0, rank)
        if i < tiles - 1:
            low_rank(r0, r1, r1, n, rank)
    native.finalize()
    return native
