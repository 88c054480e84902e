"""The blockwise constructor of the synthetic workloads (host only).

* worker processes and threads build byte-identical stores;
* the worker module never loads the native library (so spawned workers never
  touch the HIP runtime);
* the randomized range finder's tail is the true discarded Frobenius norm of
  the factors it returns, and its stores keep the full SVD's ranks;
* the distance evaluation is bit-identical to a sum over the last axis.
"""

import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from qhmat_amd import blockwork as B
from qhmat_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bytes(h) -> bytes:
    d = tempfile.mkdtemp(prefix="qhmat-wl-")
    path = os.path.join(d, "s.qhmat")
    h.save(path)
    try:
        with open(path, "rb") as f:
            return f.read()
    finally:
        os.remove(path)
        os.rmdir(d)


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_distances_bitwise(dim):
    rng = np.random.default_rng(dim)
    p, q = rng.uniform(size=(70, dim)), rng.uniform(size=(50, dim))
    diff = p[:, None, :] - q[None, :, :]
    assert np.array_equal(B.distances(p, q), np.sqrt(np.sum(diff * diff, axis=2)))


def test_processes_match_threads():
    pts = W.uniform_points(1536, 2, seed=7)
    a = W.build_kernel_hmatrix(pts, W.exponential, tolerance=1e-6, leaf_size=64,
                               processes=False, workers=4)
    b = W.build_kernel_hmatrix(pts, W.exponential, tolerance=1e-6, leaf_size=64,
                               processes=True, workers=3)
    assert _bytes(a) == _bytes(b)


def test_worker_module_loads_no_native_library():
    code = ("import sys, qhmat_amd.blockwork as b\n"
            "maps = open('/proc/self/maps').read()\n"
            "print(int('qhmat_amd._native' in sys.modules), int('libqhmat_hip' in maps),"
            " int('amdhip64' in maps))\n")
    out = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True,
                         text=True, timeout=120, check=True).stdout.split()
    assert out == ["0", "0", "0"]


def test_range_tail_is_discarded_norm():
    rng = np.random.default_rng(3)
    # a 400 x 300 matrix with geometric singular values
    u = np.linalg.qr(rng.normal(size=(400, 300)))[0]
    v = np.linalg.qr(rng.normal(size=(300, 300)))[0]
    a = (u * 0.7 ** np.arange(300)) @ v.T
    allowance = 1e-6 * np.linalg.norm(a)
    uu, s, vt, tail = B.range_spectrum(a, allowance, seed=5)
    assert s.size < 150  # the range finder stopped early, no full-SVD fallback
    assert tail[-1] <= 0.5 * allowance
    for r in (0, 5, 20, s.size // 2, s.size):
        exact = np.linalg.norm(a - (uu[:, :r] * s[:r]) @ vt[:r])
        assert tail[r] == pytest.approx(exact, rel=1e-6, abs=1e-12 * np.linalg.norm(a))
    # the leading singular values are the matrix's own
    np.testing.assert_allclose(s[:20], 0.7 ** np.arange(20), rtol=1e-10)


def test_range_spectrum_keeps_full_ranks():
    pts = W.uniform_points(2048, 1, seed=2)
    kw = dict(tolerance=1e-6, leaf_size=64, processes=False, workers=4)
    full = W.build_kernel_hmatrix(pts, W.log_kernel, svd="full", **kw)
    fast = W.build_kernel_hmatrix(pts, W.log_kernel, svd="range", **kw)
    assert [b.rank for b in full.native.blocks] == [b.rank for b in fast.native.blocks]
    # each block's quantized factors stay inside its tolerance budget
    assert fast.native.reconstruction_bound <= 1.01 * full.native.reconstruction_bound


def test_worker_error_raises_promptly():
    """An infeasible block raised in a worker reaches the caller, and the
    workers with answers still unread are ended at once."""
    import time

    from qhmat_amd import InfeasibleError

    pts = W.uniform_points(2500, 2, seed=1)
    t = time.time()
    with pytest.raises(InfeasibleError):
        W.build_kernel_hmatrix(pts, W.exponential, tolerance=1e-8, leaf_size=64,
                               processes=True, workers=3)
    assert time.time() - t < 25


def _panel_quadrature(x, centre, half, axis, n=300):
    """Midpoint rule for (1/4 pi) * integral of 1/|x - y| over a panel."""
    a, b = [d for d in range(3) if d != axis]
    us = centre[a] + half[a] * ((np.arange(n) + 0.5) / n * 2 - 1)
    vs = centre[b] + half[b] * ((np.arange(n) + 0.5) / n * 2 - 1)
    uu, vv = np.meshgrid(us, vs, indexing="ij")
    y = np.zeros(uu.shape + (3,))
    y[..., a], y[..., b], y[..., axis] = uu, vv, centre[axis]
    r = np.linalg.norm(y - x, axis=-1)
    return (1 / r).mean() * 4 * half[a] * half[b] / (4 * np.pi)


def test_panel_integral_closed_form():
    """The closed-form panel integral against quadrature away from the panel,
    and the textbook self term: integral of 1/r over a square of side s about
    its centre = 4 s ln(1 + sqrt 2)."""
    rng = np.random.default_rng(0)
    k = B.RectangleSingleLayer()
    for axis in range(3):
        centre = rng.uniform(-1, 1, 3)
        half = np.full(3, 0.1)
        half[axis] = 0.0
        panel = np.concatenate([centre, half, [axis]])[None]
        for x in (rng.uniform(-1, 1, 3), centre + np.array([0.3, -0.2, 0.25])):
            got = k(np.concatenate([x, np.zeros(4)])[None], panel)[0, 0]
            assert got == pytest.approx(_panel_quadrature(x, centre, half, axis), rel=1e-6)
        got = k(np.concatenate([centre, np.zeros(4)])[None], panel)[0, 0]
        assert got == pytest.approx(4 * 0.2 * np.log(1 + np.sqrt(2)) / (4 * np.pi), rel=1e-14)


def test_suite_cases():
    cube = W.suite_cube_bem(4)
    assert cube.kernel_data.shape == (96, 7) and cube.points.shape == (96, 3)
    # every panel lies on the unit cube's surface, half-width 1/8 in-plane
    assert np.allclose(np.abs(cube.points).max(axis=1), 0.5)
    h = cube.build(1e-2)
    x = np.random.default_rng(2).normal(size=(96, 2))
    a = cube.kernel(cube.kernel_data, cube.kernel_data)
    assert np.all(np.abs(h @ x - a @ x).max(axis=0) <= h.reconstruction_bound * np.abs(x).sum(axis=0))
    assert set(h.factors) == {i for i, b in enumerate(h.native.blocks) if not b.dense and b.rank}
    two = W.suite_inverse_distance_2d(300)
    d = two.kernel(two.points, two.points)
    assert np.all(np.diag(d) == 0) and np.all(d[~np.eye(300, dtype=bool)] > 0)


def test_tune_cache_round_trip(tmp_path):
    """The tuner's persisted cache: fingerprints tell stores apart and are
    stable, entries survive a save and a reload (timing itself needs a GPU,
    tests/test_gpu_tune.py)."""
    from qhmat_amd import tune as T

    a = W.build_kernel_hmatrix(W.uniform_points(600, 2, seed=1), W.exponential,
                               tolerance=1e-6, leaf_size=32).native
    b = W.build_kernel_hmatrix(W.uniform_points(600, 2, seed=2), W.exponential,
                               tolerance=1e-6, leaf_size=32).native
    fa = T.store_fingerprint(a)
    assert fa == T.store_fingerprint(a) and fa != T.store_fingerprint(b)
    cache = T.TuneCache(tmp_path / "tune.json")
    key = T.TuneCache.key(fa, "dev", "fp32split", 7)
    assert cache.lookup(key) is None
    cache.record(key, T.TuneResult(k=7, precision="fp32split", best=4, us_per_apply={4: 1.5, 16: 2.0}))
    cache.save()
    again = T.TuneCache(tmp_path / "tune.json")
    assert again.lookup(key) == 4
    assert again.lookup(T.TuneCache.key(fa, "dev", "fp64", 7)) is None
