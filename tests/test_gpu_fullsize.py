"""Parity at the full sizes of the measured workloads.

* 2-D exp(-r/0.2) kernel, n=16384, leaf 64, tol 1e-6 (the headline store):
  k=1 and k=32 in both precisions against the CPU oracle.
* 3-D single-layer BEM, n=32768 jittered Fibonacci sphere, support boxes of
  half-width h: k=64 in both precisions against the CPU oracle, and
  |H x - A x| <= reconstruction_bound * ||x||_1 per column on a sample of
  rows of the source matrix A (evaluated from the points, not stored).
* A tol-1e-8 stress store (every component 32 bits, ranks from the SVD
  tail; the reference constructor refuses this tolerance): both precisions
  against the oracle, FP32-split inside reconstruction_bound * ||x||_1.

Bars: FP64 within rtol 1e-12 (atol 1e-12 of the largest |y|), as the
reference GPU test (tests/test_cuda.py:23); FP32 split within relative
Frobenius error 1e-6; decoded codes bit-exact on a sample of blocks.

The oracle is the reference's own compiled matmul (oracle/_ref) when it is
built, else the C restatement, one right-hand-side column per thread
(bitwise the serial matmul, tests/test_oracle.py).
"""

import os
import sys
import tempfile
import time

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    # straight to the terminal (past pytest's capture): a long build prints
    # progress instead of going quiet
    print(f"[fullsize] {msg}", file=sys.__stderr__, flush=True)


def _build(workload, **kw):
    t = time.time()
    last = [0.0]

    def progress(done, total):
        if time.time() - last[0] > 10 or done == total:
            last[0] = time.time()
            _say(f"{workload.name}: {done}/{total} blocks")

    h = workload.build(progress=progress, **kw)
    _say(f"{workload.name}: built in {time.time() - t:.1f}s")
    return h


def _oracle(h):
    d = tempfile.mkdtemp(prefix="qhmat-full-")
    path = os.path.join(d, "store.qhmat")
    h.save(path)
    try:
        if O.ref_available():
            return O.RefHMatrix.load(path)
        return O.OracleStore.from_file(path, verify_checksum=False)
    finally:
        os.remove(path)
        os.rmdir(d)


def _fp64_close(got, want):
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * float(np.abs(want).max()))


def _check_codes(native, step):
    for index, info in enumerate(native.blocks):
        if index % step or info.dense or info.rank == 0:
            continue
        r = info.rectangle
        for side, rows in ((0, r.row1 - r.row0), (1, r.col1 - r.col0)):
            words, _, widths = native.block_words(index, side)
            np.testing.assert_array_equal(native.decode_codes(index, side),
                                          O.decode_codes(words, rows, widths),
                                          err_msg=f"block {index} side {side}")


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    return torch


@pytest.fixture(scope="module")
def square(torch):
    from qhmat_amd import workloads as W

    w = W.square_exp(16384)
    h = _build(w)
    return w, h.native, _oracle(h)


@pytest.fixture(scope="module")
def sphere(torch):
    from qhmat_amd import workloads as W

    w = W.sphere_bem(32768)
    h = _build(w)
    return w, h.native, _oracle(h)


@pytest.mark.parametrize("k", [1, 32])
def test_square_exp_16384(square, k):
    w, native, orc = square
    x = w.rhs_matrix(k)
    want = orc.matmul_columns(x)
    _fp64_close(native.matmul(x, True), want)
    assert H.rel_err(native.matmul(x, True, precision="fp32split"), want) <= H.FP32_REL_TOL


def test_square_exp_16384_codes(square):
    _check_codes(square[1], 97)


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_sphere_bem_32768_k64(sphere, precision):
    w, native, orc = sphere
    x = w.rhs_matrix(64)
    want = orc.matmul_columns(x)
    got = native.matmul(x, True, precision=precision)
    if precision == "fp64":
        _fp64_close(got, want)
    else:
        assert H.rel_err(got, want) <= H.FP32_REL_TOL
    # against the source matrix on a sample of rows
    rows = np.random.default_rng(5).choice(native.order, size=256, replace=False)
    exact = w.rows(rows) @ x
    bound = native.reconstruction_bound * np.abs(x).sum(axis=0)
    assert np.all(np.abs(got[rows] - exact).max(axis=0) <= bound)


def test_sphere_bem_32768_codes(sphere):
    _check_codes(sphere[1], 401)


@pytest.fixture(scope="module")
def stress(torch):
    from qhmat_amd import workloads as W

    w = W.square_exp(4096, tolerance=1e-8)
    h = _build(w, stress_width=32)
    return w, h.native, _oracle(h)


@pytest.mark.parametrize("k", [1, 16])
def test_stress_store_all_32_bit(stress, k):
    w, native, orc = stress
    blocks = native.blocks
    assert all(blocks[i].dense or blocks[i].rank == 0 or set(native.block_words(i, 0)[2]) == {32}
               for i in range(0, len(blocks), 37))
    x = w.rhs_matrix(k)
    want = orc.matmul_columns(x)
    _fp64_close(native.matmul(x, True), want)
    got = native.matmul(x, True, precision="fp32split")
    assert H.rel_err(got, want) <= H.FP32_REL_TOL
    rows = np.arange(0, native.order, 5)
    exact = w.rows(rows) @ x
    bound = native.reconstruction_bound * np.abs(x).sum(axis=0)
    assert np.all(np.abs(got[rows] - exact).max(axis=0) <= bound)
