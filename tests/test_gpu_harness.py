"""The application harness (dense GEMM, batched GEMMs on unquantized,
per-application-decoded and predecoded factors, the fused apply): every
method computes the same product to its precision, and the device factor
decode equals the oracle's unpacker times the component scales."""

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    return torch


@pytest.fixture(scope="module")
def built(torch):
    from qhmat_amd import workloads as W

    pts = np.random.default_rng(3).uniform(size=(900, 2))
    return W.build_kernel_hmatrix(pts, W.exponential, tolerance=1e-6, leaf_size=30,
                                  keep_factors=True)


def test_methods_agree(built):
    from qhmat_amd.harness import Harness

    res = Harness(built.native, factors=built.factors).run(5, repeats=2, warmup=1)
    assert set(res) == {"dense_fp64", "dense_fp32", "h_fp32_batched", "unpacked_batched_fp32",
                        "predecoded_batched_fp32", "packed_fused_fp32split", "packed_fused_fp64"}
    assert res["packed_fused_fp64"]["rel_err_vs_fp64_apply"] == 0.0
    assert res["dense_fp64"]["rel_err_vs_fp64_apply"] <= 1e-12
    for m in ("dense_fp32", "unpacked_batched_fp32", "predecoded_batched_fp32",
              "packed_fused_fp32split"):
        assert res[m]["rel_err_vs_fp64_apply"] <= 1e-5, m
    # unquantized factors: the store's quantization error on top of FP32
    assert res["h_fp32_batched"]["rel_err_vs_fp64_apply"] <= 1e-4
    for v in res.values():
        assert v["ms"] > 0


def test_factor_decode_matches_oracle(built, torch):
    from qhmat_amd.harness import Harness

    native = built.native
    hn = Harness(native)
    buf = hn.predecoded_factors.cpu().numpy()
    checked = 0
    for rows, cols, dense, lay in hn.buckets:
        if dense is not None:
            continue
        off_u, off_vt, count, m, c, r = lay
        members = [i for i, b in enumerate(native.blocks)
                   if not b.dense and b.rank == r
                   and b.rectangle.row1 - b.rectangle.row0 == m
                   and b.rectangle.col1 - b.rectangle.col0 == c]
        for slot, index in enumerate(members[:3]):
            for side, rows_n, off in ((0, m, off_u + slot * m * r), (1, c, off_vt + slot * r * c)):
                words, scales, widths = native.block_words(index, side)
                codes = O.decode_codes(words, rows_n, widths)  # rank x rows
                want = (codes.astype(np.float64) * np.asarray(scales)[:, None]).astype(np.float32)
                got = buf[off:off + rows_n * r].reshape((rows_n, r) if side == 0 else (r, rows_n))
                np.testing.assert_array_equal(got.T if side == 0 else got, want)
                checked += 1
    assert checked > 0
    # decoding again per application gives the same bits
    hn.plan.decode(hn.live)
    torch.cuda.synchronize()
    assert torch.equal(hn.live, hn.predecoded_factors)


def test_factor_plan_rejects_bad_slots(built, torch):
    from qhmat_amd import _native as N

    native = built.native
    dense = next(i for i, b in enumerate(native.blocks) if b.dense)
    low = next(i for i, b in enumerate(native.blocks) if not b.dense and b.rank > 0)
    with pytest.raises(ValueError):
        N.FactorPlan(native, [dense], [0], [0], 1 << 20)
    with pytest.raises(ValueError):
        N.FactorPlan(native, [low], [2], [0], 1 << 20)
    with pytest.raises(ValueError):
        N.FactorPlan(native, [low], [0], [0], 1)  # output too small


def test_store_release_refused_while_a_factor_plan_exists(torch):
    import gc

    from qhmat_amd import _native as N

    native = H.ragged_square(n=400, leaf=20, seed=4).native
    low = next(i for i, b in enumerate(native.blocks) if not b.dense and b.rank > 0)
    plan = N.FactorPlan(native, [low], [0], [0], 1 << 16)
    with pytest.raises(ValueError):
        native.release_device(0)
    del plan
    gc.collect()
    native.release_device(0)
