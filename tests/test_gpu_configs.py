"""The batched and served workloads at mid size, against the oracle.

* Batch: uneven stores drawn as the batched benchmark draws them (orders
  log-uniform, three kernels in 1-D/2-D/3-D, tolerances 1e-4/1e-6 and
  tol-1e-8 stress members with every component at 32 bits, k in 1..16):
  every member of the batched apply and of its graph replays against the
  oracle's apply of that member, in both precisions.
* Serving: stores drawn as the served benchmark draws them; a mix of
  requests submitted together (continuous batching merges them) checked one
  by one against the oracle, and the closed-loop session generator run
  briefly (every session completes, the statistics are whole).

Bars as everywhere: FP64 within rtol = atol = 1e-12 (scaled), FP32 split
within relative Frobenius error 1e-6.
"""

import os
import sys
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[configs] {msg}", file=sys.__stderr__, flush=True)


def _oracle(native):
    d = tempfile.mkdtemp(prefix="qhmat-cfg-")
    path = os.path.join(d, "store.qhmat")
    native.save(path)
    try:
        if O.ref_available():
            return O.RefHMatrix.load(path)
        return O.OracleStore.from_file(path, verify_checksum=False)
    finally:
        os.remove(path)
        os.rmdir(d)


@pytest.fixture(scope="module")
def torch():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    return torch


@pytest.fixture(scope="module")
def batch_set(torch):
    from qhmat_amd import workloads as W

    natives, ks, shapes = W.batch_workload(12, seed=4, n_max=6000, progress=_say)
    assert any(s["stress"] for s in shapes) and not all(s["stress"] for s in shapes)
    rng = np.random.default_rng(21)
    xs = [rng.normal(size=(n.order, k)) for n, k in zip(natives, ks)]
    want = [_oracle(n).matmul_columns(x) for n, x in zip(natives, xs)]
    return natives, ks, xs, want


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_batch_members_match_oracle(batch_set, precision, torch):
    from qhmat_amd.batch import DeviceBatch

    natives, ks, xs, want = batch_set
    batch = DeviceBatch(natives, ks)
    x = batch.pack([torch.from_numpy(a).cuda() for a in xs])
    y = batch.apply(x, precision=precision)
    torch.cuda.synchronize()
    for i, (yi, w) in enumerate(zip(batch.views(y), want)):
        _say(f"batch member {i} ({precision})")
        H.same_result(yi.cpu().numpy(), w, precision)
    # graph replays reproduce the eager result
    out = batch.empty()
    graph = batch.graph(x, out, precision=precision)
    for _ in range(3):
        out.zero_()
        graph.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, y)


@pytest.fixture(scope="module")
def served(torch):
    from qhmat_amd import workloads as W

    natives, _ = W.serving_workload(6, n_min=2048, n_max=8192, progress=_say)
    return natives


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_served_results_match_oracle(served, precision, torch):
    from qhmat_amd.serve import Server

    rng = np.random.default_rng(8)
    plan = [(name, k) for _ in range(3) for name in served for k in (1, 4, 16)]
    xs = [rng.normal(size=(served[name].order, k)) for name, k in plan]
    with Server(hbm_budget=64 << 30, precision=precision, slots=3) as server:
        for name, native in served.items():
            server.admit(name, native)
        futures = [server.submit(name, torch.from_numpy(x).cuda()) for (name, _), x in zip(plan, xs)]
        got = [f.result(timeout=120).cpu().numpy() for f in futures]
        assert server.passes < len(plan)
    oracles = {name: _oracle(native) for name, native in served.items()}
    for (name, _), x, g in zip(plan, xs, got):
        H.same_result(g, oracles[name].matmul_columns(x), precision)


def test_closed_loop_sessions(served, torch):
    from qhmat_amd.serve import Server, closed_loop_load

    orders = {name: n.order for name, n in served.items()}
    with Server(hbm_budget=64 << 30, precision="fp32split", slots=3) as server:
        for name, native in served.items():
            server.admit(name, native)
        out = closed_loop_load(server, list(served), orders, rate=2000.0, duration=0.5,
                               session=20)
    assert out["sessions"] > 0
    assert out["requests"] == out["sessions"] * 20
    assert out["latency_p99_ms"] >= out["latency_p50_ms"] > 0
