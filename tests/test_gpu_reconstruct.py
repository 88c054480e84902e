"""Dense reconstruction on the device (qh_reconstruct, HMatrix.to_dense with
device="cuda"): bitwise the oracle's reconstruction (orc_reconstruct, the
reference core's loop order: sum over components of (code_l lscale)
(code_r rscale), tests/test_oracle.py) and the reference-written golden
dense matrices; the reference's own check that the apply equals the
reconstructed matrix times X (tests/test_core.py:192) on the device."""

from pathlib import Path

import numpy as np
import pytest

import qhmat_amd as qhmat
from tests import helpers as H

GOLDEN = Path(__file__).resolve().parent / "golden"

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    return torch


BUILDERS = {
    "two_by_two": lambda: H.two_by_two(),
    "ragged_square": lambda: H.ragged_square(n=900, leaf=30).native,
    "line_kernel": lambda: H.line_kernel(600, 16, 1e-6, widths=(4, 8, 12, 16, 24, 32))[0].native,
    "wide_strips": lambda: H.wide_strips(tiles=68, rank=3),
    "user_partition": lambda: H.user_partition()[0].native,
}


@pytest.mark.parametrize("name", list(BUILDERS))
def test_bitwise_oracle_reconstruction(torch, name):
    native = BUILDERS[name]()
    got = native.reconstruct_device().cpu().numpy()
    np.testing.assert_array_equal(got, H.oracle_of(native).reconstruct())
    np.testing.assert_array_equal(got, native.reconstruct())


GOLDEN_DENSE = sorted(p.stem for p in GOLDEN.glob("*.qhmat")
                      if "dense" in np.load(GOLDEN / f"{p.stem}.npz"))


@pytest.mark.parametrize("name", GOLDEN_DENSE)
def test_bitwise_golden_dense(torch, name):
    h = qhmat.HMatrix.load(GOLDEN / f"{name}.qhmat")
    dense = np.load(GOLDEN / f"{name}.npz")["dense"]
    np.testing.assert_array_equal(h.native.reconstruct_device().cpu().numpy(), dense)


def test_strided_output_and_apply_consistency(torch):
    native = H.ragged_square(n=700, leaf=24).native
    n = native.order
    wide = torch.full((n, n + 5), 7.0, dtype=torch.float64, device="cuda")
    out = wide[:, 2:2 + n]
    native.reconstruct_device(out)
    torch.cuda.synchronize()
    dense = out.cpu().numpy()
    np.testing.assert_array_equal(dense, native.reconstruct())
    assert torch.all(wide[:, :2] == 7.0) and torch.all(wide[:, 2 + n:] == 7.0)
    x = np.random.default_rng(1).normal(size=(n, 3))
    np.testing.assert_allclose(native.matmul(x, True), dense @ x, rtol=1e-12, atol=1e-12)


def test_to_dense_surface(torch):
    h, _ = H.line_kernel(200, 8, 1e-6, widths=(4, 8, 16, 32))
    np.testing.assert_array_equal(h.to_dense(device="cuda"), h.to_dense())
    with pytest.raises(ValueError):
        h.to_dense(device="tpu")
    with pytest.raises(ValueError):
        h.native.reconstruct_device(torch.empty((3, 3), dtype=torch.float64, device="cuda"))
