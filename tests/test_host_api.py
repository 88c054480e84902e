"""The reference's own test strategy (tests/test_core.py) applied to the
drop-in surface, plus the product's construction pieces pinned against the
oracle and, where it is built, the reference's compiled core."""

import numpy as np
import pytest

import qhmat_amd as qhmat
from oracle import oracle as O
from oracle import qhmat001
from tests import helpers as H

ref_only = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def test_allocator_feasibility_and_storage():
    sigma = np.array([4.0, 1.0, 0.2])
    allocation = qhmat.allocate_bits(sigma, [16, 4, 8], 0.08, 20, 16)
    assert allocation.feasible
    assert allocation.bits == sorted(allocation.bits, reverse=True)
    assert allocation.bound <= 0.08
    assert allocation.payload_bits == 36 * sum(allocation.bits)
    assert not qhmat.allocate_bits(sigma, [4, 8], 1e-8, 20, 16).feasible


@pytest.mark.parametrize("bits", [2, 4, 8, 12, 16, 24, 32])
def test_coefficient_decreases(bits):
    assert qhmat.coefficient(bits) > 0
    if bits > 2:
        assert qhmat.coefficient(bits) < qhmat.coefficient(2)
    assert qhmat.coefficient(bits) == O.coefficient(bits)


def test_allocator_rejects_invalid_input():
    with pytest.raises(ValueError):
        qhmat.allocate_bits([1.0, 2.0], [4, 8], 0.1, 3, 3)
    with pytest.raises(ValueError):
        qhmat.allocate_bits([2.0, 1.0], [4, 4], 0.1, 3, 3)
    with pytest.raises(ValueError):
        qhmat.allocate_bits([2.0], [1, 8], 0.1, 3, 3)


def test_spatial_partition_covers_entries():
    points = np.arange(8, dtype=float)[:, None]
    partition = qhmat.spatial_partition(points, points, points, 2, 1.0)
    assert sorted(partition.permutation) == list(range(8))
    coverage = np.zeros((8, 8), dtype=int)
    for block in partition.blocks:
        coverage[block.row0:block.row1, block.col0:block.col1] += 1
    np.testing.assert_array_equal(coverage, np.ones((8, 8), dtype=int))
    assert any(block.admissible for block in partition.blocks)


def test_native_block_apply_and_metadata():
    matrix = H.two_by_two()
    expected = np.array([[2.0, 1.0], [1.0, 3.0]])
    np.testing.assert_allclose(matrix.reconstruct(), expected, rtol=0, atol=1e-12)
    np.testing.assert_allclose(matrix.matmul(np.array([1.0, 2.0])), expected @ [1.0, 2.0])
    np.testing.assert_allclose(matrix.matmul(np.eye(2)), expected)
    assert matrix.logical_bits == 2 * 64 + 2 * (2 * 8 + 128 + 2)
    assert matrix.reconstruction_bound >= 0


def test_native_rejects_partition_gap():
    matrix = qhmat.NativeHMatrix(2, [0, 1])
    matrix.add_dense(qhmat.Rectangle(0, 1, 0, 1), np.ones((1, 1)))
    with pytest.raises(ValueError, match="partition"):
        matrix.finalize()


def test_high_level_kernel_matrix():
    h, source = H.line_kernel(12, 3, 0.25)
    assert h.shape == source.shape
    assert any(not block.dense for block in h.blocks)
    assert np.max(np.abs(source - h.to_dense())) <= h.reconstruction_bound + 1e-12
    np.testing.assert_allclose(h @ np.eye(12), h.to_dense(), atol=1e-12)
    assert h.logical_bits < source.size * 64


def test_memory_report_and_verbose(capsys, tmp_path):
    source = np.diag([2.0, 3.0]).astype(np.float32)
    matrix = qhmat.HMatrix.from_dense(source, np.arange(2)[:, None], tolerance=0.1,
                                      leaf_size=1, verbose=1)
    output = capsys.readouterr().out
    assert output.count("Input array:") == 1
    assert "Dense FP64 baseline:" in output
    assert "H-matrix C++ resident estimate:" in output
    report = matrix.storage_report()
    assert report.source_array_bytes == source.nbytes == qhmat.memory_bytes(source)
    assert report.dense_fp64_bytes == source.size * 8
    assert report.logical_bits == matrix.logical_bits
    assert report.logical_bytes == (matrix.logical_bits + 7) // 8
    assert report.resident_bytes == qhmat.memory_bytes(matrix) == qhmat.memory_bytes(
        matrix._native)
    assert report.resident_bytes > matrix.allocated_bytes
    path = tmp_path / "memory.qhmat"
    matrix.save(path)
    loaded = qhmat.HMatrix.load(path).storage_report()
    assert loaded.source_array_bytes is None
    assert loaded.logical_bits == report.logical_bits
    with pytest.raises(TypeError, match="NumPy array"):
        qhmat.memory_bytes([1.0, 2.0])


def test_partition_verbose_and_infeasible_output(capsys):
    source = np.diag([2.0, 3.0])
    rectangles = [qhmat.Rectangle(i, i + 1, j, j + 1) for i in range(2) for j in range(2)]
    qhmat.HMatrix.from_partition(source, [0, 1], rectangles, tolerance=0.1, verbose=1)
    assert capsys.readouterr().out.count("Input array:") == 1
    qhmat.HMatrix.from_partition(source, [0, 1], rectangles, tolerance=0.1)
    assert capsys.readouterr().out == ""
    with pytest.raises(qhmat.InfeasibleError):
        qhmat.HMatrix.from_dense(np.ones((4, 4)), np.arange(4)[:, None], tolerance=1e-12,
                                 leaf_size=1, bit_widths=(4, 8), verbose=1)
    assert capsys.readouterr().out == ""


def test_nonfinite_rejected():
    source = np.eye(3)
    source[0, 1] = np.nan
    with pytest.raises(ValueError, match="finite"):
        qhmat.HMatrix.from_dense(source, np.arange(3)[:, None], tolerance=0.1)
    h, _ = H.line_kernel(12, 3, 0.25)
    with pytest.raises(ValueError, match="finite"):
        h @ np.full(12, np.inf)


def test_numpy_integer_permutation_and_zero_offdiagonal(tmp_path):
    source = np.diag([2.0, 3.0])
    rectangles = [qhmat.Rectangle(0, 1, 0, 1), qhmat.Rectangle(0, 1, 1, 2),
                  qhmat.Rectangle(1, 2, 0, 1), qhmat.Rectangle(1, 2, 1, 2)]
    matrix = qhmat.HMatrix.from_partition(source, np.array([1, 0], dtype=np.int64), rectangles,
                                          tolerance=0.1)
    np.testing.assert_array_equal(matrix.to_dense(), source)
    assert sum(not block.dense for block in matrix.blocks) == 2
    path = tmp_path / "zero-rank.qhmat"
    matrix.save(path)
    np.testing.assert_array_equal(qhmat.HMatrix.load(path).to_dense(), source)


@pytest.mark.parametrize("bit", [2, 4, 8, 12, 16, 24, 32])
def test_random_orthogonal_factor_quantization(bit):
    native, expected, rng = H.orthogonal_factors(bit)
    reconstructed = native.reconstruct()
    assert np.max(np.abs(reconstructed - expected)) <= native.reconstruction_bound + 1e-12
    x = rng.normal(size=(17, 4))
    np.testing.assert_allclose(native.matmul(x), reconstructed @ x, atol=1e-12)


def test_packed_file_round_trip_and_corruption(tmp_path):
    h, _ = H.line_kernel(8, 2, 0.25)
    path = tmp_path / "matrix.qhmat"
    h.save(path)
    loaded = qhmat.HMatrix.load(path)
    np.testing.assert_array_equal(loaded.to_dense(), h.to_dense())
    np.testing.assert_array_equal(loaded @ np.eye(8), h @ np.eye(8))
    assert loaded.logical_bits == h.logical_bits
    assert loaded.reconstruction_bound == h.reconstruction_bound
    with pytest.raises(ValueError, match="existing"):
        h.save(path)
    bad = tmp_path / "bad.qhmat"
    data = bytearray(path.read_bytes())
    data[-9] ^= 1
    bad.write_bytes(data)
    with pytest.raises(RuntimeError, match="checksum"):
        qhmat.HMatrix.load(bad)
    truncated = tmp_path / "truncated.qhmat"
    truncated.write_bytes(path.read_bytes()[:-5])
    with pytest.raises(RuntimeError, match="truncated|trailing"):
        qhmat.HMatrix.load(truncated)


def test_device_request_without_gpu_raises():
    """The reference raises RuntimeError for a GPU request it cannot serve
    (bindings.cpp:168, cuda.cu:115-117); there is no host fallback."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    h, _ = H.line_kernel(12, 3, 0.25)
    with pytest.raises(RuntimeError, match="no CUDA device"):
        h.matmul(np.ones(12), device="cuda")
    with pytest.raises(ValueError, match="device must be"):
        h.matmul(np.ones(12), device="gpu")


# -- product pieces against the oracle --------------------------------------------
@pytest.mark.parametrize("bit", list(range(2, 33)))
def test_product_pack_matches_oracle(bit):
    native, _, _ = H.orthogonal_factors(bit)
    rng = np.random.default_rng(200 + bit)
    u, _ = np.linalg.qr(rng.normal(size=(9, 3)))
    v, _ = np.linalg.qr(rng.normal(size=(8, 3)))
    root = np.sqrt([3.0, 1.0, 0.2])
    lw, ls = O.pack_factor(u * root, [bit] * 3)
    rw, rs = O.pack_factor(v * root, [bit] * 3)
    words, scales, widths = native.block_words(2, 0)
    np.testing.assert_array_equal(words, lw)
    np.testing.assert_array_equal(scales, ls)
    words, scales, _ = native.block_words(2, 1)
    np.testing.assert_array_equal(words, rw)
    np.testing.assert_array_equal(scales, rs)
    assert list(widths) == [bit] * 3


def test_product_host_apply_matches_oracle():
    for native in (H.ragged_square(n=700, leaf=24).native, H.user_partition()[0].native,
                   H.two_by_two()):
        orc = H.oracle_of(native)
        x = np.random.default_rng(0).normal(size=(native.order, 3))
        # same products and summation order as the reference (core.cpp:360-390),
        # whatever the thread count
        np.testing.assert_array_equal(native.matmul(x), orc.matmul(x))


def test_product_file_is_parsed_by_independent_reader(tmp_path):
    h = H.ragged_square(n=400, leaf=24)
    path = tmp_path / "s.qhmat"
    h.save(path)
    parsed = qhmat001.parse(path.read_bytes())
    assert parsed["order"] == 400
    assert qhmat001.write(parsed) == path.read_bytes()


# -- product pieces against the reference's compiled core ---------------------------
@ref_only
def test_allocate_matches_reference():
    rng = np.random.default_rng(11)
    menus = [[4, 8, 12, 16, 24, 32], [16, 4, 8], [2, 32], [32], [3, 7, 19]]
    for trial in range(300):
        r = int(rng.integers(0, 12))
        sigma = np.sort(rng.exponential(size=r))[::-1] * 10 ** rng.uniform(-6, 2)
        menu = menus[trial % len(menus)]
        budget = float(np.sum(sigma) * 10 ** rng.uniform(-10, -1)) if r else 0.1
        mine = qhmat.allocate_bits(sigma, menu, budget, 40, 30)
        bits, bound, payload, feasible = O.ref_allocate(sigma, menu, budget, 40, 30)
        assert (mine.bits, mine.bound, mine.payload_bits, mine.feasible) == (
            bits, bound, payload, feasible)


@ref_only
@pytest.mark.parametrize("dim,leaf,eta", [(1, 2, 1.0), (2, 8, 1.0), (3, 5, 0.7), (2, 1, 2.0)])
def test_partition_matches_reference(dim, leaf, eta):
    rng = np.random.default_rng(dim * 10 + leaf)
    pts = rng.uniform(size=(203, dim))
    pts[5] = pts[6]  # ties in the median sort
    lo, hi = pts - 0.01, pts + 0.02
    part = qhmat.spatial_partition(pts, lo, hi, leaf, eta)
    perm, rects = O.ref_partition(pts, lo, hi, leaf, eta)
    assert part.permutation == list(perm)
    mine = [(b.row0, b.row1, b.col0, b.col1, int(b.admissible)) for b in part.blocks]
    assert mine == [tuple(int(v) for v in r) for r in rects]


@ref_only
def test_files_byte_identical_with_reference(tmp_path):
    """The same blocks added to the product and to the reference core save to
    byte-identical QHMAT001 files, and each loads the other's file."""
    h = H.ragged_square(n=500, leaf=24)
    mine = tmp_path / "mine.qhmat"
    h.save(mine)
    r = O.RefHMatrix.load(mine)
    theirs = tmp_path / "theirs.qhmat"
    r.save(theirs)
    assert mine.read_bytes() == theirs.read_bytes()
    back = qhmat.HMatrix.load(theirs)
    x = np.random.default_rng(2).normal(size=(500, 2))
    np.testing.assert_allclose(back @ x, r.matmul(x), rtol=1e-12, atol=1e-12)
    assert back.reconstruction_bound == r.reconstruction_bound
    assert back.logical_bits == r.logical_bits
    assert back.allocated_bytes == r.allocated_bytes


@ref_only
def test_product_pack_matches_reference_core(tmp_path):
    """add_low_rank into both cores from the same SVD factors: same words."""
    rng = np.random.default_rng(8)
    n = 64
    x = np.sort(rng.uniform(size=n))
    a = 1.0 / (np.abs(x[:, None] - x[None, :]) + 0.2)
    r = O.RefHMatrix(n, list(range(n)))
    m = qhmat.NativeHMatrix(n, list(range(n)))
    cuts = [0, 20, 45, 64]
    for i in range(3):
        for j in range(3):
            rect = (cuts[i], cuts[i + 1], cuts[j], cuts[j + 1])
            blk = a[rect[0]:rect[1], rect[2]:rect[3]]
            if i == j:
                r.add_dense(rect, blk)
                m.add_dense(qhmat.Rectangle(*rect), blk)
                continue
            u, s, vt = np.linalg.svd(blk, full_matrices=False)
            k = 4
            widths = [32, 24, 12, 5][:k]
            r.add_low_rank(rect, u[:, :k], s[:k], vt[:k].T, widths, 6, 1e-3)
            m.add_low_rank(qhmat.Rectangle(*rect), u[:, :k], s[:k], vt[:k].T, widths, 6, 1e-3)
    r.finalize()
    m.finalize()
    p1, p2 = tmp_path / "ref.qhmat", tmp_path / "mine.qhmat"
    r.save(p1)
    m.save(p2)
    assert p1.read_bytes() == p2.read_bytes()
