"""The serving loop's C ABI on the host: argument checks and the error a
missing device gives (no GPU work)."""

import ctypes as C

import pytest


def test_null_arguments_are_rejected():
    from qhmat_amd import _native as N
    from qhmat_amd import serve as S

    S._server_sigs()
    lib = N.lib()
    assert lib.qh_server_create(None, C.byref(C.c_void_p())) == N.QH_EINVAL
    assert b"null" in lib.qh_last_error()
    ticket = C.c_uint64()
    assert lib.qh_server_submit(None, 0, None, 1, None, C.byref(ticket)) == N.QH_EINVAL
    state = C.c_int32()
    assert lib.qh_server_wait(None, 1, 0.0, C.byref(state)) == N.QH_EINVAL
    assert lib.qh_server_destroy(None) == N.QH_OK


def test_no_device_raises_runtime_error():
    import torch

    from qhmat_amd import serve as S

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(RuntimeError, match="no CUDA device"):
        S.NativeServer()
