"""Batched apply of uneven stores and its HIP-graph replay: every member's
result must be bitwise the result of applying that store alone and match
the oracle within the precision's tolerance, and every graph replay must
reproduce the eager result."""

import numpy as np
import pytest

import qhmat_amd as qhmat
from tests import helpers as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def members():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu tests need a HIP device")
    natives = [
        H.ragged_square(n=700, leaf=24).native,
        H.line_kernel(40, 4, 1e-6, widths=(4, 8, 12, 16, 24, 32))[0].native,
        H.user_partition()[0].native,
        H.two_by_two(),
        H.ragged_square(n=1500, leaf=40, seed=9).native,
        H.orthogonal_factors(24)[0],
    ]
    ks = [1, 3, 4, 7, 16, 20]
    return natives, ks


@pytest.mark.parametrize("precision", ["fp64", "fp32split"])
def test_batch_matches_single_applies(members, precision):
    import torch

    from qhmat_amd.batch import DeviceBatch

    natives, ks = members
    batch = DeviceBatch(natives, ks)
    rng = np.random.default_rng(4)
    xs = [torch.from_numpy(rng.normal(size=(n.order, k))).cuda() for n, k in zip(natives, ks)]
    x = batch.pack(xs)
    y = batch.apply(x, precision=precision)
    torch.cuda.synchronize()
    for native, xi, yi in zip(natives, xs, batch.views(y)):
        alone = native.apply_device(xi.contiguous(), precision=precision)
        torch.cuda.synchronize()
        assert torch.equal(yi, alone)
        # and the oracle's apply of the member's own x
        H.same_result(yi.cpu().numpy(), H.oracle_of(native).matmul(xi.cpu().numpy()), precision)


def test_batch_graph_replays(members):
    import torch

    from qhmat_amd.batch import DeviceBatch

    natives, ks = members
    batch = DeviceBatch(natives, ks)
    rng = np.random.default_rng(5)
    x = batch.pack([torch.from_numpy(rng.normal(size=(n.order, k))).cuda()
                    for n, k in zip(natives, ks)])
    want = batch.apply(x)
    torch.cuda.synchronize()
    y = batch.empty()
    graph = batch.graph(x, y)
    for replay in range(3):
        y.fill_(float("nan"))
        scratch = torch.randn(4096, device="cuda").sum()  # eager work between replays
        graph.launch()
        torch.cuda.synchronize()
        assert torch.isfinite(scratch)
        assert torch.equal(y, want), f"replay {replay}"


def test_batch_rejects_bad_input(members):
    from qhmat_amd.batch import DeviceBatch

    natives, _ = members
    with pytest.raises(ValueError):
        DeviceBatch(natives[:2], [1])
    with pytest.raises(ValueError, match="right-hand-side"):
        DeviceBatch(natives[:1], [0])


def test_store_release_refused_while_batched(members):
    """A store's device copy cannot be released while a batch (or its graph)
    holds its device pointers; it can once they are gone."""
    import gc

    import torch

    from qhmat_amd.batch import DeviceBatch

    natives, ks = members
    native = H.ragged_square(n=500, leaf=20, seed=21).native
    batch = DeviceBatch([native, natives[1]], [2, 3])
    graph = batch.graph(batch.empty(), batch.empty())
    with pytest.raises(ValueError):
        native.release_device(0)
    del graph, batch
    gc.collect()
    torch.cuda.synchronize()
    native.release_device(0)
    # and it comes back on the next use
    x = np.random.default_rng(2).normal(size=(native.order, 2))
    H.same_result(native.matmul(x, True), H.oracle_of(native).matmul(x), "fp64")
