"""Writes the golden fixtures under tests/golden/ with the reference's own
compiled core (oracle/_ref, built from /root/reference/src/qhmat by
`make -C oracle ref`).

For each case the reference core does the partition (spatial_partition),
the width allocation (allocate), the balanced packing (add_low_rank), the
validation (finalize), the file (save) and the product y = H x (matmul).
The only steps not run by reference code are the per-block SVD and rank
selection of python/qhmat/matrix.py:213-245 (NumPy, restated in
qhmat_amd.matrix.factor_block with one BLAS thread).

Outputs per case: <name>.qhmat (QHMAT001, written by the reference) and
<name>.npz (x, y = reference matmul(x), reconstruction bound, and for small
orders the reference's dense reconstruction).

Run from the repository root:  python tests/golden/make_golden.py
"""

from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import oracle as O  # noqa: E402

HERE = Path(__file__).resolve().parent


def build_reference(points, source, *, tolerance, leaf, widths=(4, 8, 12, 16, 24, 32),
                    lower=None, upper=None, eta=1.0):
    """Reference partition + allocation + packing of `source` on `points`."""
    from threadpoolctl import threadpool_limits

    from qhmat_amd import matrix as M
    from qhmat_amd._native import Rectangle

    lo = points if lower is None else lower
    hi = points if upper is None else upper
    perm, rects = O.ref_partition(points, lo, hi, leaf, eta)
    ref = O.RefHMatrix(len(perm), perm)
    with threadpool_limits(1):
        for r in rects:
            r0, r1, c0, c1, adm = (int(v) for v in r)
            block = np.ascontiguousarray(source[np.ix_(perm[r0:r1], perm[c0:c1])])
            if (r0, r1) == (c0, c1):
                ref.add_dense((r0, r1, c0, c1), block)
                continue
            budget = tolerance * float(np.max(np.abs(block)))
            if budget == 0.0:
                ref.add_low_rank((r0, r1, c0, c1), np.zeros((r1 - r0, 0)), [],
                                 np.zeros((c1 - c0, 0)), [], len(widths), 0.0, adm)
                continue
            # SVD + rank selection (matrix.py:221-238) only: widths by the reference allocate
            u, sigma, v, _, residual = M.factor_block(
                block, Rectangle(r0, r1, c0, c1), tolerance=tolerance, widths=tuple(widths),
                rank_tolerance=0.1, max_rank=None, truncation_fraction=0.25,
                arithmetic_reserve=0.02)
            bits, _, _, feasible = O.ref_allocate(sigma, widths, budget * 0.98 - residual,
                                                  r1 - r0, c1 - c0)
            assert feasible
            ref.add_low_rank((r0, r1, c0, c1), u, sigma, v, bits, len(widths), residual, adm)
    ref.finalize()
    return ref


def cases():
    out = {}
    # tests/test_core.py:44-60
    r = O.RefHMatrix(2, [0, 1])
    r.add_dense((0, 1, 0, 1), [[2.0]])
    r.add_dense((1, 2, 1, 2), [[3.0]])
    for row, col in [(0, 1), (1, 0)]:
        r.add_low_rank((row, row + 1, col, col + 1), np.ones((1, 1)), [1.0], np.ones((1, 1)),
                       [8], 3, 0.0)
    r.finalize()
    out["two_by_two"] = r
    # tests/test_core.py:170-192 at every width of the reference test
    for bit in (2, 4, 8, 12, 16, 24, 32):
        rng = np.random.default_rng(200 + bit)
        u, _ = np.linalg.qr(rng.normal(size=(9, 3)))
        v, _ = np.linalg.qr(rng.normal(size=(8, 3)))
        sigma = np.array([3.0, 1.0, 0.2])
        r = O.RefHMatrix(17, list(range(17)))
        r.add_dense((0, 9, 0, 9), np.eye(9))
        r.add_dense((9, 17, 9, 17), np.eye(8))
        r.add_low_rank((0, 9, 9, 17), u, sigma, v, [bit] * 3, 1, 0.0)
        r.add_low_rank((9, 17, 0, 9), v, sigma, u, [bit] * 3, 1, 0.0)
        r.finalize()
        out[f"orthogonal_b{bit}"] = r
    # tests/test_cuda.py:8-23 matrix
    pts = np.linspace(0.0, 1.0, 10)[:, None]
    src = 1.0 / (0.4 + np.abs(pts - pts.T))
    out["line10"] = build_reference(pts, src, tolerance=0.25, leaf=2, widths=(4, 8, 16, 32))
    # multi-level 1-D store, 1/(r+0.4)
    pts = np.random.default_rng(0).uniform(size=(300, 1))
    src = 1.0 / (np.abs(pts - pts.T) + 0.4)
    out["line_inverse_300"] = build_reference(pts, src, tolerance=1e-6, leaf=16)
    # 2-D exp(-r/0.2), uneven leaves
    pts = np.random.default_rng(0).uniform(size=(320, 2))
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    out["square_exp_320"] = build_reference(pts, np.exp(-d / 0.2), tolerance=1e-6, leaf=20)
    # 3-D single layer on a jittered Fibonacci sphere with panel support boxes
    n = 256
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    rho = np.sqrt(1 - z * z)
    th = math.pi * (3 - math.sqrt(5)) * i
    pts = np.column_stack((rho * np.cos(th), rho * np.sin(th), z))
    pts = pts + np.random.default_rng(0).normal(0, 1e-3, size=pts.shape)
    h = math.sqrt(4 * math.pi / n)
    d = np.sqrt(((pts[:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    with np.errstate(divide="ignore"):
        src = 1.0 / (4 * math.pi * d)
    src[d == 0] = 1.0 / (4 * math.pi * 0.5 * h)
    out["sphere_256"] = build_reference(pts, src, tolerance=1e-6, leaf=16, lower=pts - h,
                                        upper=pts + h)
    return out


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref is not built: make -C oracle ref")
    for name, ref in cases().items():
        path = HERE / f"{name}.qhmat"
        if path.exists():
            path.unlink()
        ref.save(path)
        n = ref.order
        x = np.random.default_rng(1).normal(size=(n, 3))
        payload = {"x": x, "y": ref.matmul(x), "bound": ref.reconstruction_bound,
                   "logical_bits": ref.logical_bits}
        if n <= 64:
            payload["dense"] = ref.reconstruct()
        np.savez(HERE / f"{name}.npz", **payload)
        print(f"{name}: n={n} {path.stat().st_size} bytes")


if __name__ == "__main__":
    main()
