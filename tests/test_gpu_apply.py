"""Parity of the HIP apply against the CPU oracle (oracle/oracle.c).

Bars: decoded codes bit-exact; FP64 apply within rtol = atol = 1e-12 (the
reference GPU test's tolerance, tests/test_cuda.py:23); FP32 split-precision
apply within relative (Frobenius) error 1e-6 of the FP64 oracle; results
bitwise reproducible and independent of the right-hand-side count.
"""

import numpy as np
import pytest

import qhmat_amd as q
from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

FP32_REL_TOL = 1e-6


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    return torch


def _stores():
    out = {"two_by_two": H.two_by_two()}
    for bit in (2, 4, 8, 12, 16, 24, 32):
        out[f"orthogonal_b{bit}"] = H.orthogonal_factors(bit)[0]
    out["line12"] = H.line_kernel(12, 3, 0.25)[0].native
    out["line10"] = H.line_kernel(10, 2, 0.25)[0].native
    out["ragged_square"] = H.ragged_square().native
    out["user_partition"] = H.user_partition()[0].native
    out["wide_strips"] = H.wide_strips()
    return out


@pytest.fixture(scope="module")
def stores(torch):
    built = _stores()
    return {name: (native, H.oracle_of(native)) for name, native in built.items()}


@pytest.fixture(scope="module")
def line_inverse(torch):
    from qhmat_amd import workloads as W

    h = W.line_inverse().build()
    return h.native, H.oracle_of(h.native)


NAMES = ["two_by_two"] + [f"orthogonal_b{b}" for b in (2, 4, 8, 12, 16, 24, 32)] + [
    "line12", "line10", "ragged_square", "user_partition", "wide_strips"]


def _check_codes(native):
    for index, info in enumerate(native.blocks):  # one pass over the metadata
        if info.dense or info.rank == 0:
            continue
        r = info.rectangle
        for side, rows in ((0, r.row1 - r.row0), (1, r.col1 - r.col0)):
            words, _, widths = native.block_words(index, side)
            want = O.decode_codes(words, rows, widths)
            got = native.decode_codes(index, side)
            np.testing.assert_array_equal(got, want, err_msg=f"block {index} side {side}")


@pytest.mark.parametrize("name", NAMES)
def test_decoded_codes_bit_exact(stores, name):
    _check_codes(stores[name][0])


def test_decoded_codes_bit_exact_line_inverse(line_inverse):
    _check_codes(line_inverse[0])


@pytest.mark.parametrize("k", [1, 3, 4, 17, 33])
@pytest.mark.parametrize("name", NAMES)
def test_fp64_matches_oracle(stores, name, k):
    native, orc = stores[name]
    x = np.random.default_rng(k).normal(size=(native.order, k))
    np.testing.assert_allclose(native.matmul(x, True), orc.matmul(x), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [1, 4, 16])
def test_fp64_line_inverse(line_inverse, k):
    native, orc = line_inverse
    x = np.random.default_rng(1).normal(size=(native.order, k))
    np.testing.assert_allclose(native.matmul(x, True), orc.matmul(x), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("k", [1, 4, 32])
@pytest.mark.parametrize("name", NAMES + ["line_inverse"])
def test_fp32split_accuracy(stores, line_inverse, name, k):
    native, orc = line_inverse if name == "line_inverse" else stores[name]
    x = np.random.default_rng(7 + k).normal(size=(native.order, k))
    want = orc.matmul(x)
    got = native.matmul(x, True, precision="fp32split")
    assert H.rel_err(got, want) <= FP32_REL_TOL


def test_reference_cuda_case(stores):
    """tests/test_cuda.py:8-23 with the same matrix, seed and tolerance."""
    native = stores["line10"][0]
    rhs = np.random.default_rng(11).normal(size=(10, 3))
    np.testing.assert_allclose(native.matmul(rhs, True), native.matmul(rhs), rtol=1e-12,
                               atol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_bitwise_reproducible_and_rhs_count_independent(stores, precision):
    native = stores["ragged_square"][0]
    x = np.random.default_rng(2).normal(size=(native.order, 20))
    a = native.matmul(x[:, :9], True, precision=precision)  # one matrix-core pass
    b = native.matmul(x[:, :9], True, precision=precision)
    np.testing.assert_array_equal(a, b)
    # the same columns inside two 16-wide passes: the same bits
    np.testing.assert_array_equal(native.matmul(x, True, precision=precision)[:, :9], a)
    # narrow passes (vector path): a column alone and inside a 3-wide pass
    c = native.matmul(x[:, :3], True, precision=precision)
    for p in range(3):
        np.testing.assert_array_equal(native.matmul(x[:, p], True, precision=precision), c[:, p])
    # across the two paths: within the precision's tolerance
    for p in (0, 4, 8):
        H.same_result(native.matmul(x[:, p], True, precision=precision), a[:, p], precision)


def test_device_tensors_strided_and_workspace(stores, torch):
    native, orc = stores["ragged_square"]
    n = native.order
    wide = torch.randn(n, 8, dtype=torch.float64, device="cuda")
    x = wide[:, 2:5]  # row stride 8, k = 3
    from qhmat_amd import device as D

    ws = torch.empty(D.workspace_bytes(native, 3), dtype=torch.uint8, device="cuda")
    out = torch.full((n, 6), float("nan"), dtype=torch.float64, device="cuda")
    y = out[:, 1:4]
    native.apply_device(x, y, workspace=ws)
    torch.cuda.synchronize()
    want = orc.matmul(x.cpu().numpy())
    np.testing.assert_allclose(y.cpu().numpy(), want, rtol=1e-12, atol=1e-12)
    assert torch.isnan(out[:, 0]).all() and torch.isnan(out[:, 4:]).all()


def test_hmatrix_surface(torch):
    h, src = H.line_kernel(40, 4, 1e-6, widths=(4, 8, 12, 16, 24, 32))
    x = np.random.default_rng(3).normal(size=40)
    y = h.matmul(x, device="cuda")
    assert y.shape == (40,) and isinstance(y, np.ndarray)
    np.testing.assert_allclose(y, h @ x, rtol=1e-12, atol=1e-12)
    xt = torch.from_numpy(x).cuda()
    yt = h.matmul(xt, device="cuda", precision="fp32split")
    assert yt.is_cuda and yt.shape == (40,)
    assert H.rel_err(yt.cpu().numpy(), y) <= FP32_REL_TOL
    with pytest.raises(ValueError, match="finite"):
        h.matmul(np.full(40, np.nan), device="cuda")


def test_error_bound_against_source(line_inverse):
    """|H x - A x| <= reconstruction_bound * ||x||_1 per column, on the device."""
    from qhmat_amd import workloads as W

    native, _ = line_inverse
    w = W.line_inverse()
    x = np.random.default_rng(1).normal(size=(native.order, 4))
    exact = w.dense() @ x
    bound = native.reconstruction_bound * np.abs(x).sum(axis=0)
    for precision in ("fp64", "fp32split"):
        err = np.abs(native.matmul(x, True, precision=precision) - exact).max(axis=0)
        assert np.all(err <= bound + 1e-9)


def test_concurrent_host_threads(stores):
    """Several host threads apply one store at once (the reference releases
    the interpreter lock in matmul, bindings.cpp:164) while another retunes
    its passes: every result matches the oracle."""
    import threading

    native, orc = stores["ragged_square"]
    rng = np.random.default_rng(11)
    cases = [(k, prec, rng.normal(size=(native.order, k)))
             for k in (1, 3, 16, 33) for prec in ("fp64", "fp32split")]
    want = {i: orc.matmul(x) for i, (_, _, x) in enumerate(cases)}
    errors, stop = [], threading.Event()

    def worker(seed):
        order = np.random.default_rng(seed).permutation(len(cases))
        try:
            for _ in range(3):
                for i in order:
                    k, prec, x = cases[i]
                    got = native.matmul(x, True, precision=prec)
                    H.same_result(got, want[i], prec)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    def retune():
        kbs = [1, 2, 4, 16, 32, 0]
        i = 0
        while not stop.is_set():
            native.set_rhs_block(33, kbs[i % len(kbs)])
            native.set_rhs_block(16, kbs[(i + 1) % len(kbs)])
            i += 1
        native.set_rhs_block(33, 0)
        native.set_rhs_block(16, 0)

    threads = [threading.Thread(target=worker, args=(s,)) for s in range(6)]
    tuner = threading.Thread(target=retune)
    tuner.start()
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    stop.set()
    tuner.join(timeout=60)
    assert not errors, errors[0]


def test_side_stream_and_strided_input(stores, torch):
    """apply on a caller's side stream with a non-contiguous x: the input copy,
    the output and the workspace are ordered on that stream, and the result
    equals the default-stream one."""
    native, orc = stores["ragged_square"]
    x_full = torch.from_numpy(np.random.default_rng(7).normal(size=(native.order, 10))).cuda()
    x = x_full[:, ::2]  # non-contiguous, 5 columns
    want = native.apply_device(x.contiguous())
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    got = native.apply_device(x, stream=side)
    side.synchronize()
    assert torch.equal(got, want)
    H.same_result(got.cpu().numpy(), orc.matmul(x.cpu().numpy()), "fp64")
