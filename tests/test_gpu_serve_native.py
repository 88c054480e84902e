"""The library's serving loop (qh_server_*, NativeServer): admission against
the HBM budget, merged passes, each result equal to the oracle's apply within
the precision's tolerance (and bitwise qh_apply's alone for a request served
in a pass of its own), evicted stores, and the closed-loop load."""

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
    from qhmat_amd.serve import NativeServer

    with NativeServer(hbm_budget=1) as server:
        with pytest.raises(ValueError, match="budget"):
            server.admit("square", natives["square"])
    with NativeServer(hbm_budget=1 << 30, max_rhs=16) as server:
        server.admit("square", natives["square"])
        st = server.stats()
        assert st["resident_bytes"] > 0 and st["buffer_bytes"] > 0
        # the store is pinned while the server holds it
        with pytest.raises(ValueError):
            natives["square"].release_device(0)
        server.evict("square")
        assert server.stats()["resident_bytes"] == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_single_requests_bitwise_alone(natives, precision):
    """Requests waited for one at a time run in passes of their own: bitwise
    qh_apply's result."""
    import torch

    from qhmat_amd.serve import NativeServer

    rng = np.random.default_rng(5)
    with NativeServer(hbm_budget=1 << 32, precision=precision, max_rhs=32) as server:
        for name, native in natives.items():
            server.admit(name, native)
        for name, native in natives.items():
            for k in (1, 3, 16, 21):
                x = torch.from_numpy(rng.normal(size=(native.order, k))).cuda()
                want = native.apply_device(x, precision=precision)
                torch.cuda.synchronize()
                got = server.wait(server.submit(name, x), timeout=60)
                torch.cuda.synchronize()
                assert torch.equal(got, want), (name, k)
        assert server.stats()["completed"] == 12


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_merged_requests_against_oracle(natives, precision):
    import torch

    from qhmat_amd.serve import NativeServer

    rng = np.random.default_rng(11)
    plan = [(name, k) for _ in range(8) for name in natives for k in (1, 2, 4)]
    xs = [torch.from_numpy(rng.normal(size=(natives[n].order, k))).cuda() for n, k in plan]
    torch.cuda.synchronize()
    with NativeServer(hbm_budget=1 << 32, precision=precision, slots=2, max_rhs=24) as server:
        for name, native in natives.items():
            server.admit(name, native)
        reqs = [server.submit(name, x, sync=False) for (name, _), x in zip(plan, xs)]
        results = [server.wait(r, timeout=60) for r in reqs]
        torch.cuda.synchronize()
        st = server.stats()
    assert st["completed"] == len(plan) and st["completed_rhs"] == sum(k for _, k in plan)
    assert st["passes"] < len(plan)  # requests were merged into shared passes
    assert 0 < st["p50_ms"] <= st["p99_ms"]
    oracles = {name: H.oracle_of(native) for name, native in natives.items()}
    for (name, _), x, got in zip(plan, xs, results):
        H.same_result(got.cpu().numpy(), oracles[name].matmul(x.cpu().numpy()), precision)


def test_requests_of_an_evicted_store_fail(natives):
    import torch

    from qhmat_amd.serve import NativeServer

    native = natives["line"]
    x = torch.randn(native.order, 2, dtype=torch.float64, device="cuda")
    with NativeServer(hbm_budget=1 << 30, max_rhs=8) as server:
        server.admit("line", native)
        server.wait(server.submit("line", x), timeout=60)
        server.evict("line")
        with pytest.raises(KeyError):
            server.submit("line", x)
        # the store is released and can be admitted again
        server.admit("line", native)
        y = server.wait(server.submit("line", x), timeout=60)
        torch.cuda.synchronize()
        H.same_result(y.cpu().numpy(), H.oracle_of(native).matmul(x.cpu().numpy()), "fp64")


def test_closed_loop_load(natives):
    from qhmat_amd.serve import NativeServer

    with NativeServer(hbm_budget=1 << 32, precision="fp32split", max_rhs=32) as server:
        for name, native in natives.items():
            server.admit(name, native)
        out = server.closed_loop(list(natives), ks=(1, 4), rate=2000.0, duration=0.5, session=10)
    assert out["errors"] == 0
    assert out["requests"] == out["sessions"] * 10 > 0
    assert out["passes"] <= out["requests"]
    assert 0 < out["latency_p50_ms"] <= out["latency_p99_ms"]
