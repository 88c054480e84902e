"""The launch-shape autotuner: every right-hand-side block of one arithmetic
path gives the same bits (and the two paths agree to the precision's
tolerance), the tuner keeps one of the candidates, and batches built
afterwards follow it without changing any member's result."""

import numpy as np
import pytest

from qhmat_amd import batch as B
from qhmat_amd import tune as T
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    return torch


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_every_block_gives_the_same_bits(torch, precision):
    native = H.ragged_square(n=900, leaf=30).native
    x = np.random.default_rng(3).normal(size=(native.order, 11))
    outs = {}
    for kb in (1, 2, 4, 16, 32, 64):
        native.set_rhs_block(11, kb)
        outs[kb] = native.matmul(x, True, precision=precision)
    native.set_rhs_block(11, 0)
    vector = [kb for kb in outs if kb < H.MATRIX_MIN_KB]
    matrix = [kb for kb in outs if kb >= H.MATRIX_MIN_KB]
    for group in (vector, matrix):
        for kb in group[1:]:
            np.testing.assert_array_equal(outs[kb], outs[group[0]])
    H.same_result(outs[matrix[0]], outs[vector[0]], precision)


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_several_launch_slices_give_the_same_bits(torch, precision):
    # k = 70: two 64-wide launch slices (the second 6 columns wide), three
    # 32-wide or five 16-wide ones; pass L splits each slice into row groups
    native = H.ragged_square(n=900, leaf=30).native
    x = np.random.default_rng(4).normal(size=(native.order, 70))
    outs = {}
    for kb in (16, 32, 64):
        native.set_rhs_block(70, kb)
        outs[kb] = native.matmul(x, True, precision=precision)
    native.set_rhs_block(70, 0)
    for kb in (32, 64):
        np.testing.assert_array_equal(outs[kb], outs[16])
    want = H.oracle_of(native).matmul(x)
    H.same_result(outs[16], want, precision)


def test_tuner_keeps_a_candidate_and_batches_follow(torch):
    a = H.ragged_square(n=700, leaf=24).native
    b = H.line_kernel(300, 16, 1e-6, widths=(4, 8, 12, 16, 24, 32))[0].native
    ks = [5, 3]
    results = T.tune_batch([a, b], ks, precision="fp32split", repeats=3)
    assert [r.k for r in results] == ks
    for r in results:
        assert r.best in T.candidates(r.k) and set(r.us_per_apply) == set(T.candidates(r.k))
    xs = [np.random.default_rng(i).normal(size=(n.order, k)) for i, (n, k) in enumerate(zip([a, b], ks))]
    batch = B.DeviceBatch([a, b], ks)
    buf = batch.pack([torch.from_numpy(x).cuda() for x in xs])
    out = batch.apply(buf, precision="fp32split")
    torch.cuda.synchronize()
    for view, native, x in zip(batch.views(out), [a, b], xs):
        np.testing.assert_array_equal(view.cpu().numpy(), native.matmul(x, True, precision="fp32split"))


def test_tuner_cache_skips_timing(torch, tmp_path):
    native = H.ragged_square(n=700, leaf=24).native
    cache = T.TuneCache(tmp_path / "tune.json")
    first = T.tune_rhs_block(native, 6, repeats=2, cache=cache)
    assert not first.cached and first.best in T.candidates(6)
    native.set_rhs_block(6, 0)
    again = T.tune_rhs_block(native, 6, repeats=2, cache=T.TuneCache(tmp_path / "tune.json"))
    assert again.cached and again.best == first.best and again.us_per_apply == {}
    native.set_rhs_block(6, 0)
