"""The tile-major device layout checked on the host: the library builds the
device plan and replays both passes' index arithmetic with bounds-checked
accesses (qh_plan_emulate); the result must match the oracle.  Catches
chunking, offset, staging and workspace-index mistakes without a GPU."""

from pathlib import Path

import numpy as np
import pytest

import qhmat_amd as qhmat
from tests import helpers as H

GOLDEN = Path(__file__).resolve().parent / "golden"


def _cases():
    yield "two_by_two", H.two_by_two()
    for bit in (2, 5, 12, 24, 27, 32):
        yield f"orthogonal_b{bit}", H.orthogonal_factors(bit)[0]
    yield "line12", H.line_kernel(12, 3, 0.25)[0].native
    yield "ragged_square", H.ragged_square(n=700, leaf=24).native
    yield "user_partition", H.user_partition()[0].native
    yield "line_wide_leaves", H.line_kernel(300, 100, 1e-6, widths=(4, 8, 12, 16, 24, 32))[0].native
    yield "wide_strips", H.wide_strips()
    for p in sorted(GOLDEN.glob("*.qhmat")):
        yield p.stem, qhmat.HMatrix.load(p).native


CASES = dict(_cases())


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("k", [1, 3])
@pytest.mark.parametrize("path", ["narrow", "wide_fp64", "wide_fp32split"])
def test_plan_emulation_matches_oracle(name, k, path):
    native = CASES[name]
    orc = H.oracle_of(native)
    x = np.random.default_rng(k).normal(size=(native.order, k))
    got = native.plan_emulate(x, wide=path != "narrow", precision=path.removeprefix("wide_"))
    np.testing.assert_allclose(got, orc.matmul(x), rtol=1e-12, atol=1e-12)


def test_rank_zero_and_dense_only_plans():
    source = np.diag([2.0, 3.0])
    rects = [qhmat.Rectangle(i, i + 1, j, j + 1) for i in range(2) for j in range(2)]
    h = qhmat.HMatrix.from_partition(source, [1, 0], rects, tolerance=0.1)
    for wide in (False, True):
        np.testing.assert_allclose(h.native.plan_emulate(np.ones(2), wide=wide).ravel(), [2.0, 3.0])
    one = qhmat.NativeHMatrix(1, [0])
    one.add_dense(qhmat.Rectangle(0, 1, 0, 1), np.array([[4.0]]))
    one.finalize()
    for wide in (False, True):
        np.testing.assert_allclose(one.plan_emulate(np.array([0.5]), wide=wide).ravel(), [2.0])
