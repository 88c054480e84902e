"""Golden fixtures written by the reference's compiled core
(tests/golden/make_golden.py): QHMAT001 stores and the reference matmul of a
fixed x.  CPU tests check the oracle and the product's host side against
them; the gpu tests check the device apply."""

from pathlib import Path

import numpy as np
import pytest

import qhmat_amd as qhmat
from oracle import oracle as O
from oracle import qhmat001
from tests import helpers as H

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.qhmat"))


def _load(name):
    data = np.load(GOLDEN / f"{name}.npz")
    return GOLDEN / f"{name}.qhmat", data


def test_fixtures_present():
    assert len(NAMES) >= 12


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_reference_bitwise(name):
    path, g = _load(name)
    s = O.OracleStore.from_file(path)
    np.testing.assert_array_equal(s.matmul(g["x"]), g["y"])
    if "dense" in g:
        np.testing.assert_array_equal(s.reconstruct(), g["dense"])


@pytest.mark.parametrize("name", NAMES)
def test_product_loads_reference_file(name, tmp_path):
    path, g = _load(name)
    h = qhmat.HMatrix.load(path)
    # the host backend follows the reference's loop order: bitwise its y
    np.testing.assert_array_equal(h @ g["x"], g["y"])
    assert h.reconstruction_bound == float(g["bound"])
    assert h.logical_bits == int(g["logical_bits"])
    if "dense" in g:
        np.testing.assert_array_equal(h.to_dense(), g["dense"])
    again = tmp_path / "again.qhmat"
    h.save(again)
    assert again.read_bytes() == path.read_bytes()


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_device_apply_matches_reference(name):
    path, g = _load(name)
    h = qhmat.HMatrix.load(path)
    np.testing.assert_allclose(h.matmul(g["x"], device="cuda"), g["y"], rtol=1e-12, atol=1e-12)
    y32 = h.matmul(g["x"], device="cuda", precision="fp32split")
    assert H.rel_err(y32, g["y"]) <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_device_codes_match_reference_words(name):
    path, _ = _load(name)
    parsed = qhmat001.parse(path.read_bytes())
    h = qhmat.HMatrix.load(path)
    for index, blk in enumerate(parsed["blocks"]):
        if blk["dense"] or len(blk["widths"]) == 0:
            continue
        r0, r1, c0, c1 = blk["rect"]
        np.testing.assert_array_equal(h.native.decode_codes(index, 0),
                                      O.decode_codes(blk["lwords"], r1 - r0, blk["widths"]))
        np.testing.assert_array_equal(h.native.decode_codes(index, 1),
                                      O.decode_codes(blk["rwords"], c1 - c0, blk["widths"]))
