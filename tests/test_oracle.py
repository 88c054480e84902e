"""Pins the CPU oracle (oracle/oracle.c) before anything is checked against it.

Each check is chosen so that a plausible mistake in the restatement -- a
wrong shift or sign extension, an off-by-one word, round-half-away instead of
half-even, a transposed factor, a dropped scale -- fails at least one:
  * hand-computed code words (not produced by any code under test);
  * the reference's own known answers (tests/test_core.py:44-60, 154-167,
    170-192) re-stated on stores the oracle packs itself;
  * a brute-force long-double sum over decoded entries;
  * word-for-word and bit-for-bit agreement with the reference's compiled
    core (oracle/_ref) where that is built.
"""

import numpy as np
import pytest

from oracle import oracle as O
from oracle import qhmat001

WIDTHS = list(range(2, 33))


# -- closed forms ---------------------------------------------------------------
@pytest.mark.parametrize("bits", WIDTHS)
def test_limit_and_coefficient_closed_form(bits):
    q = 2 ** (bits - 1) - 1
    assert O.limit(bits) == q
    assert O.coefficient(bits) == pytest.approx(1.0 / q + 1.0 / (4.0 * q * q), rel=1e-15)


# -- hand-computed code words ---------------------------------------------------------
KNOWN_WORDS = [
    # (column, bits, scale, words): every word worked out by hand, LSB first.
    ([7.0, -1.0], 4, 1.0, [0xF7]),
    ([2047.0, 1.0, -1.0, 0.0, 5.0, -2047.0, 100.0, -100.0], 12, 1.0,
     [0xFF0017FF, 0x1005000F, 0xF9C06480]),
    ([1.0, -1.0], 32, 1.0 / (2 ** 31 - 1), [0x7FFFFFFF, 0x80000001]),
    # round half to even: 0.5 -> 0, 1.5 -> 2, 2.5 -> 2, -0.5 -> 0
    ([7.0, 0.5, 1.5, 2.5, -0.5], 4, 1.0, [0x2207]),
    ([0.0, 0.0, 0.0], 8, 0.0, [0x0]),
    # b=8 crosses nothing: four codes per word, the fifth starts word 1
    ([127.0, -127.0, 64.0, -64.0, 1.0], 8, 1.0, [0xC040817F, 0x01]),
]


@pytest.mark.parametrize("column,bits,scale,words", KNOWN_WORDS)
def test_pack_known_words(column, bits, scale, words):
    got, s = O.pack_component(np.array(column), bits)
    assert s == pytest.approx(scale, rel=1e-15)
    assert [int(w) for w in got] == words


@pytest.mark.parametrize("column,bits,scale,words", KNOWN_WORDS)
def test_decode_known_words(column, bits, scale, words):
    codes = O.decode_codes(np.array(words, dtype=np.uint32), len(column), [bits])[0]
    expect = [0 if scale == 0 else int(np.rint(v / scale)) for v in column]
    assert list(codes) == expect


def test_twelve_bit_codes_fill_three_words():
    words, _ = O.pack_component(np.linspace(-1, 1, 8), 12)
    assert words.size == 3  # 8 * 12 = 96 bits, no spare word


@pytest.mark.parametrize("bits", WIDTHS)
def test_pack_decode_round_trip(bits):
    """|code * scale - x| <= scale / 2 (experiments/cuda/host_test.cpp:12-20)."""
    rng = np.random.default_rng(bits)
    for rows in (1, 7, 31, 32, 33, 100):
        x = rng.normal(size=rows)
        words, scale = O.pack_component(x, bits)
        assert words.size == (rows * bits + 31) // 32
        codes = O.decode_codes(words, rows, [bits])[0]
        q = 2 ** (bits - 1) - 1
        assert np.all(np.abs(codes) <= q)
        assert np.max(np.abs(codes * scale - x)) <= scale / 2 + 1e-14
        assert np.max(np.abs(codes)) == q  # the peak maps to +-q


# -- stores the oracle packs itself ------------------------------------------------------
def _lowrank_block(rect, u, sigma, v, widths, menu=1, trunc=0.0):
    root = np.sqrt(np.asarray(sigma, dtype=np.float64))
    lw, ls = O.pack_factor(np.asarray(u) * root, widths)
    rw, rs = O.pack_factor(np.asarray(v) * root, widths)
    qb = float(sum(s * O.coefficient(int(b)) for s, b in zip(sigma, widths)))
    return {"rect": rect, "admissible": False, "dense": False, "widths": np.asarray(widths),
            "menu_size": menu, "truncation_error": trunc, "quantization_bound": qb,
            "lscale": ls, "rscale": rs, "lwords": lw, "rwords": rw}


def _dense_block(rect, values):
    return {"rect": rect, "admissible": False, "dense": True,
            "values": np.asarray(values, dtype=np.float64)}


def _store(order, perm, blocks):
    return O.OracleStore({"order": order, "perm": np.asarray(perm), "blocks": blocks})


def test_two_by_two_known_answer():
    """tests/test_core.py:44-60: [[2,1],[1,3]] with 8-bit rank-1 blocks."""
    blocks = [_dense_block((0, 1, 0, 1), [[2.0]]), _dense_block((1, 2, 1, 2), [[3.0]])]
    for r, c in ((0, 1), (1, 0)):
        blocks.append(_lowrank_block((r, r + 1, c, c + 1), [[1.0]], [1.0], [[1.0]], [8]))
    s = _store(2, [0, 1], blocks)
    np.testing.assert_array_equal(s.matmul(np.array([1.0, 2.0])), [4.0, 7.0])
    np.testing.assert_array_equal(s.reconstruct(), [[2.0, 1.0], [1.0, 3.0]])
    np.testing.assert_array_equal(s.matmul(np.eye(2)), [[2.0, 1.0], [1.0, 3.0]])


def test_rank_zero_blocks_and_permutation():
    """tests/test_core.py:154-167: zero off-diagonal blocks, permutation [1, 0]."""
    blocks = [_dense_block((0, 1, 0, 1), [[3.0]]),
              _lowrank_block((0, 1, 1, 2), np.zeros((1, 0)), [], np.zeros((1, 0)), []),
              _lowrank_block((1, 2, 0, 1), np.zeros((1, 0)), [], np.zeros((1, 0)), []),
              _dense_block((1, 2, 1, 2), [[2.0]])]
    s = _store(2, [1, 0], blocks)
    # clustered 0 is source 1: A = diag(2, 3) in source order
    np.testing.assert_array_equal(s.reconstruct(), np.diag([2.0, 3.0]))
    np.testing.assert_array_equal(s.matmul(np.array([1.0, 1.0])), [2.0, 3.0])


def _orthogonal_store(bit):
    rng = np.random.default_rng(200 + bit)
    u, _ = np.linalg.qr(rng.normal(size=(9, 3)))
    v, _ = np.linalg.qr(rng.normal(size=(8, 3)))
    sigma = np.array([3.0, 1.0, 0.2])
    blocks = [_dense_block((0, 9, 0, 9), np.eye(9)), _dense_block((9, 17, 9, 17), np.eye(8)),
              _lowrank_block((0, 9, 9, 17), u, sigma, v, [bit] * 3),
              _lowrank_block((9, 17, 0, 9), v, sigma, u, [bit] * 3)]
    a = (u * sigma) @ v.T
    expected = np.eye(17)
    expected[:9, 9:] = a
    expected[9:, :9] = a.T
    bound = float(sum(s * O.coefficient(bit) for s in sigma))
    return _store(17, list(range(17)), blocks), expected, bound, rng


@pytest.mark.parametrize("bit", [2, 4, 8, 12, 16, 24, 32])
def test_random_orthogonal_factors(bit):
    """tests/test_core.py:170-192 on an oracle-packed store."""
    s, expected, bound, rng = _orthogonal_store(bit)
    rec = s.reconstruct()
    assert np.max(np.abs(rec - expected)) <= bound + 1e-12
    x = rng.normal(size=(17, 4))
    np.testing.assert_allclose(s.matmul(x), rec @ x, atol=1e-12)


def _decoded_dense(s: O.OracleStore, parsed) -> np.ndarray:
    """Entries formed from decoded codes in long double, by brute force."""
    n = parsed["order"]
    out = np.zeros((n, n), dtype=np.longdouble)
    perm = parsed["perm"]
    for blk in parsed["blocks"]:
        r0, r1, c0, c1 = blk["rect"]
        if blk["dense"]:
            tile = np.asarray(blk["values"], dtype=np.longdouble)
        else:
            w = blk["widths"]
            lc = O.decode_codes(blk["lwords"], r1 - r0, w).astype(np.longdouble)
            rc = O.decode_codes(blk["rwords"], c1 - c0, w).astype(np.longdouble)
            lf = lc * np.asarray(blk["lscale"], dtype=np.longdouble)[:, None]
            rf = rc * np.asarray(blk["rscale"], dtype=np.longdouble)[:, None]
            tile = lf.T @ rf
        out[np.ix_(perm[r0:r1], perm[c0:c1])] = tile
    return out


def test_brute_force_long_double():
    rng = np.random.default_rng(9)
    perm = rng.permutation(11)
    blocks = [_dense_block((0, 5, 0, 5), rng.normal(size=(5, 5))),
              _dense_block((5, 11, 5, 11), rng.normal(size=(6, 6)))]
    for rect, m, n in (((0, 5, 5, 11), 5, 6), ((5, 11, 0, 5), 6, 5)):
        u, _ = np.linalg.qr(rng.normal(size=(m, 3)))
        v, _ = np.linalg.qr(rng.normal(size=(n, 3)))
        blocks.append(_lowrank_block(rect, u, [2.0, 0.5, 0.1], v, [5, 13, 27]))
    parsed = {"order": 11, "perm": perm, "blocks": blocks}
    s = O.OracleStore(parsed)
    exact = _decoded_dense(s, parsed)
    x = rng.normal(size=(11, 3))
    want = (exact @ x.astype(np.longdouble)).astype(np.float64)
    np.testing.assert_allclose(s.matmul(x), want, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(s.reconstruct(), exact.astype(np.float64), rtol=1e-14,
                               atol=1e-15)


def test_linearity_and_nonfinite():
    s, _, _, rng = _orthogonal_store(12)
    x1, x2 = rng.normal(size=(17, 2)), rng.normal(size=(17, 2))
    np.testing.assert_allclose(s.matmul(2.5 * x1 - x2), 2.5 * s.matmul(x1) - s.matmul(x2),
                               atol=1e-12)
    bad = x1.copy()
    bad[3, 1] = np.inf
    with pytest.raises(ValueError, match="finite"):
        s.matmul(bad)


def test_file_writer_parser_round_trip():
    s, _, _, _ = _orthogonal_store(24)
    parsed = {"order": 17, "perm": np.arange(17), "blocks": []}
    blob = qhmat001.write({"order": 17, "perm": np.arange(17),
                           "blocks": [_dense_block((0, 17, 0, 17), np.eye(17))]})
    back = qhmat001.parse(blob)
    assert back["order"] == 17 and back["blocks"][0]["dense"]
    with pytest.raises(ValueError, match="checksum"):
        qhmat001.parse(blob[:-9] + bytes([blob[-9] ^ 1]) + blob[-8:])
    assert parsed["order"] == 17


# -- agreement with the reference's compiled core -------------------------------------------
ref_only = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


def _ref_store(bit, tmp_path):
    rng = np.random.default_rng(200 + bit)
    u, _ = np.linalg.qr(rng.normal(size=(9, 3)))
    v, _ = np.linalg.qr(rng.normal(size=(8, 3)))
    sigma = np.array([3.0, 1.0, 0.2])
    r = O.RefHMatrix(17, list(range(17)))
    r.add_dense((0, 9, 0, 9), np.eye(9))
    r.add_dense((9, 17, 9, 17), np.eye(8))
    r.add_low_rank((0, 9, 9, 17), u, sigma, v, [bit] * 3, 1, 0.0)
    r.add_low_rank((9, 17, 0, 9), v, sigma, u, [bit] * 3, 1, 0.0)
    r.finalize()
    path = tmp_path / f"ref{bit}.qhmat"
    r.save(path)
    return r, path, (u, sigma, v), rng


@ref_only
@pytest.mark.parametrize("bit", WIDTHS)
def test_pack_matches_reference_words(bit, tmp_path):
    r, path, (u, sigma, v), _ = _ref_store(bit, tmp_path)
    parsed = qhmat001.parse(path.read_bytes())
    root = np.sqrt(sigma)
    blk = parsed["blocks"][2]
    lw, ls = O.pack_factor(u * root, [bit] * 3)
    rw, rs = O.pack_factor(v * root, [bit] * 3)
    np.testing.assert_array_equal(blk["lwords"], lw)
    np.testing.assert_array_equal(blk["rwords"], rw)
    np.testing.assert_array_equal(blk["lscale"], ls)
    np.testing.assert_array_equal(blk["rscale"], rs)


@ref_only
@pytest.mark.parametrize("bit", [2, 4, 8, 12, 16, 24, 32])
def test_matmul_and_reconstruct_bitwise_equal_reference(bit, tmp_path):
    r, path, _, rng = _ref_store(bit, tmp_path)
    s = O.OracleStore.from_file(path)
    x = rng.normal(size=(17, 5))
    np.testing.assert_array_equal(s.matmul(x), r.matmul(x))
    np.testing.assert_array_equal(s.reconstruct(), r.reconstruct())


@ref_only
def test_kernel_store_bitwise_equal_reference(tmp_path):
    """A multi-level store (1/(|dx|+0.4), n=300) written by the reference."""
    import qhmat_amd as q

    pts = np.random.default_rng(4).uniform(size=(300, 1))
    src = 1.0 / (np.abs(pts - pts.T) + 0.4)
    h = q.HMatrix.from_dense(src, pts, tolerance=1e-6, leaf_size=16)
    mine = tmp_path / "mine.qhmat"
    h.save(mine)
    r = O.RefHMatrix.load(mine)  # the reference's loader validates the file
    s = O.OracleStore.from_file(mine)
    x = np.random.default_rng(1).normal(size=(300, 3))
    np.testing.assert_array_equal(s.matmul(x), r.matmul(x))


def test_column_parallel_matmul_is_bitwise_matmul():
    """matmul_columns (one right-hand side per thread) is the same function:
    the reference loop nest never mixes columns."""
    from tests import helpers as H

    native = H.ragged_square(n=600, leaf=24).native
    s = H.oracle_of(native)
    x = np.random.default_rng(8).normal(size=(native.order, 7))
    np.testing.assert_array_equal(s.matmul_columns(x, workers=4), s.matmul(x))


@ref_only
def test_reference_column_parallel_matmul_is_bitwise_matmul(tmp_path):
    from tests import helpers as H

    native = H.ragged_square(n=600, leaf=24).native
    path = tmp_path / "store.qhmat"
    native.save(path)
    r = O.RefHMatrix.load(path)
    x = np.random.default_rng(9).normal(size=(native.order, 6))
    np.testing.assert_array_equal(r.matmul_columns(x, workers=3), r.matmul(x))
    np.testing.assert_array_equal(r.matmul(x), O.OracleStore.from_file(path).matmul(x))
