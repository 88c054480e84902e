"""Serving loop: admission against the HBM budget, continuous batching, and
each request's result the same alone and in a mix (bitwise while the merged
pass stays on the request's arithmetic path, within the precision's tolerance
otherwise) and equal to the oracle's apply within the precision's tolerance."""

import numpy as np
import pytest

from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def natives():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    return {
        "square": H.ragged_square(n=900, leaf=30).native,
        "line": H.line_kernel(200, 12, 1e-6, widths=(4, 8, 12, 16, 24, 32))[0].native,
        "user": H.user_partition()[0].native,
    }


def test_admission_budget(natives):
    from qhmat_amd.serve import AdmissionError, Server

    server = Server(hbm_budget=1)
    with pytest.raises(AdmissionError):
        server.admit("square", natives["square"])
    big = Server(hbm_budget=1 << 30)
    size = big.admit("square", natives["square"])
    assert size > 0 and big.resident_bytes == size
    big.evict("square")
    assert big.resident_bytes == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_results_same_alone_and_in_mix(natives, precision):
    import torch

    from qhmat_amd.serve import Server

    rng = np.random.default_rng(3)
    plan = [(name, k) for _ in range(6) for name in natives for k in (1, 3, 4)]
    xs = [torch.from_numpy(rng.normal(size=(natives[n].order, k))).cuda() for n, k in plan]
    alone = []
    for (name, _), x in zip(plan, xs):
        alone.append(natives[name].apply_device(x, precision=precision))
    torch.cuda.synchronize()
    with Server(hbm_budget=1 << 32, precision=precision, slots=3) as server:
        for name, native in natives.items():
            server.admit(name, native)
        futures = [server.submit(name, x) for (name, _), x in zip(plan, xs)]
        results = [f.result(timeout=60) for f in futures]
        torch.cuda.synchronize()
        assert server.passes < len(plan)  # requests were merged into shared passes
    for got, want in zip(results, alone):
        H.same_result(got.cpu().numpy(), want.cpu().numpy(), precision)
    # every served result against the oracle's apply of its own x
    oracles = {name: H.oracle_of(native) for name, native in natives.items()}
    for (name, _), x, got in zip(plan, xs, results):
        H.same_result(got.cpu().numpy(), oracles[name].matmul(x.cpu().numpy()), precision)
    lat = server.latency_ms()
    assert lat["p50"] is not None and lat["p99"] >= lat["p50"]


def test_requests_wait_for_memory(natives):
    """A budget that fits the stores but only one request's operands at a
    time still serves everything (requests queue for memory)."""
    import torch

    from qhmat_amd.serve import Server

    native = natives["square"]
    x = torch.randn(native.order, 4, dtype=torch.float64, device="cuda")
    want = native.apply_device(x)
    torch.cuda.synchronize()
    server = Server(hbm_budget=1 << 62)
    size = server.admit("square", native)
    server.budget = size + 16 * native.order * 4 + 1  # room for one request in flight
    with server:
        futures = [server.submit("square", x) for _ in range(5)]
        for f in futures:
            assert torch.equal(f.result(timeout=60), want)
        assert server.passes == 5


def test_nonfinite_requests_raise(natives):
    import torch

    from qhmat_amd.serve import Server

    n = natives["line"].order
    bad = np.ones((n, 2))
    bad[3, 1] = np.nan
    with Server(hbm_budget=1 << 32) as server:
        server.admit("line", natives["line"])
        with pytest.raises(ValueError, match="nonfinite"):
            server.submit("line", bad)  # host array: checked before queueing
        fut = server.submit("line", torch.from_numpy(bad).cuda())  # device: checked on completion
        ok = server.submit("line", torch.ones(n, 1, dtype=torch.float64, device="cuda"))
        with pytest.raises(ValueError, match="nonfinite"):
            fut.result(timeout=60)
        assert ok.result(timeout=60).shape == (n, 1)


def test_oversized_request_rejected_and_lone_request_fits(natives):
    """A request wider than the server's max_rhs is refused at submit; one of
    exactly max_rhs columns is served alone."""
    import torch

    from qhmat_amd.serve import Server

    native = natives["line"]
    with Server(hbm_budget=1 << 32, max_rhs=8) as server:
        server.admit("line", native)
        with pytest.raises(ValueError, match="max_rhs"):
            server.submit("line", torch.ones(native.order, 9, dtype=torch.float64, device="cuda"))
        x = torch.randn(native.order, 8, dtype=torch.float64, device="cuda")
        got = server.submit("line", x).result(timeout=60)
        H.same_result(got.cpu().numpy(), H.oracle_of(native).matmul(x.cpu().numpy()), "fp64")


def test_evict_fails_queued_requests(natives):
    """Requests still queued for a store that is evicted fail with KeyError
    (the server is not started here, so nothing takes them first)."""
    import torch

    from qhmat_amd.serve import Server

    native = natives["user"]
    server = Server(hbm_budget=1 << 32)
    server.admit("user", native)
    futs = [server.submit("user", torch.ones(native.order, 2, dtype=torch.float64, device="cuda"))
            for _ in range(3)]
    server.evict("user")
    for f in futs:
        with pytest.raises(KeyError):
            f.result(timeout=10)
    assert server.resident_bytes == 0
