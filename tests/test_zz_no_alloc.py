"""qh_apply_noalloc: the apply with the library's static device scratch as
its workspace, called through the C ABI and checked against qh_apply (the
same kernels, so bitwise equal) and against the CPU oracle, on two calls in a
row, the second with a smaller input."""

import ctypes as C

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    return torch


def _call(lib, N, fn, store, x, y, prec, stream, *extra):
    k = x.shape[1]
    N.check(fn(store, C.c_void_p(x.data_ptr()), x.stride(0), k, C.c_void_p(y.data_ptr()),
               y.stride(0), prec, C.c_void_p(stream.cuda_stream), *extra))


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_noalloc_matches_apply(torch, precision):
    from qhmat_amd import _native as N
    from qhmat_amd import device as D

    lib = N.lib()
    prec = D.precision_code(precision)
    stream = torch.cuda.current_stream()
    big = H.ragged_square(n=1500).native
    small = H.line_kernel(200, 16, 1e-6, widths=(4, 8, 12, 16, 24, 32))[0].native
    # the larger input first, then a smaller store and right-hand-side count:
    # the second call must not see the first call's leftovers
    for native, k in ((big, 33), (small, 3)):
        store = native.device_store(0)
        orc = H.oracle_of(native)
        x_np = np.random.default_rng(k).normal(size=(native.order, k))
        x = torch.from_numpy(x_np).cuda()
        y_noalloc = torch.full_like(x, float("nan"))
        y_apply = torch.full_like(x, float("nan"))
        _call(lib, N, lib.qh_apply_noalloc, store, x, y_noalloc, prec, stream)
        ws = torch.empty(D.workspace_bytes(native, k), dtype=torch.uint8, device="cuda")
        _call(lib, N, lib.qh_apply, store, x, y_apply, prec, stream,
              C.c_void_p(ws.data_ptr()), ws.numel())
        torch.cuda.synchronize()
        a, b = y_noalloc.cpu().numpy(), y_apply.cpu().numpy()
        np.testing.assert_array_equal(a, b)
        want = orc.matmul(x_np)
        if precision == "fp64":
            np.testing.assert_allclose(a, want, rtol=1e-12, atol=1e-12)
        else:
            assert H.rel_err(a, want) <= H.FP32_REL_TOL


def test_noalloc_rejects_oversized_workspace(torch):
    from qhmat_amd import _native as N
    from qhmat_amd import device as D

    lib = N.lib()
    cap = C.c_uint64()
    N.check(lib.qh_noalloc_scratch_bytes(C.byref(cap)))
    native = H.ragged_square(n=1500).native
    per_rhs = D.workspace_bytes(native, 1)
    k = int(cap.value // per_rhs) + 1  # the workspace grows linearly in k
    assert D.workspace_bytes(native, k) > cap.value
    x = torch.zeros((native.order, k), dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    with pytest.raises(ValueError, match="static scratch"):
        _call(lib, N, lib.qh_apply_noalloc, native.device_store(0), x, y, 0,
              torch.cuda.current_stream())
