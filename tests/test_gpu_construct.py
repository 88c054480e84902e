"""Device construction (batched FP64 SVD per block shape) against the host
NumPy construction: same partition, same ranks and widths, applies equal to
rounding, and the reference's error bound holds."""

import numpy as np
import pytest

from qhmat_amd import construct as K
from qhmat_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kernel,dim,leaf", [("exp", 2, 40), ("inverse", 1, 32), ("log", 3, 30)])
def test_device_construction_matches_host(kernel, dim, leaf):
    pts = np.random.default_rng(dim).uniform(size=(1200, dim))
    dev = K.build_on_device(pts, kernel, tolerance=1e-6, leaf_size=leaf)
    host = W.build_kernel_hmatrix(pts, W.KERNELS[kernel], tolerance=1e-6, leaf_size=leaf)
    assert [b.rank for b in dev.blocks] == [b.rank for b in host.blocks]
    x = np.random.default_rng(1).normal(size=(1200, 2))
    a = W.KERNELS[kernel](pts, pts)
    yd = dev.matmul(x, device="cuda")
    np.testing.assert_allclose(yd, host @ x, rtol=1e-9, atol=1e-9 * np.abs(host @ x).max())
    bound = dev.reconstruction_bound * np.abs(x).sum(axis=0)
    assert np.all(np.abs(yd - a @ x).max(axis=0) <= bound + 1e-9)
