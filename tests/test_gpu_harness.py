"""The application harness (dense GEMM, batched GEMMs on decoded factors, the
fused apply): every method computes the same product, to its precision."""

import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_methods_agree(tmp_path):
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    from qhmat_amd.harness import Harness

    native = H.ragged_square(n=900, leaf=30).native
    res = Harness(native).run(5, repeats=2, warmup=1)
    assert set(res) == {"dense_fp64", "dense_fp32", "predecoded_batched_fp32",
                        "packed_fused_fp32split", "packed_fused_fp64"}
    assert res["packed_fused_fp64"]["rel_err_vs_fp64_apply"] == 0.0
    assert res["dense_fp64"]["rel_err_vs_fp64_apply"] <= 1e-12
    for m in ("dense_fp32", "predecoded_batched_fp32", "packed_fused_fp32split"):
        assert res[m]["rel_err_vs_fp64_apply"] <= 1e-5, m
    for v in res.values():
        assert v["ms"] > 0
