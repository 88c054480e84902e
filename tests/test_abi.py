"""The C ABI: the library loads and exports every function include/*.h
declares (no compute calls here: this runs without a GPU)."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions() -> list[str]:
    names = []
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(qh_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("qh_store_create", "qh_apply", "qh_store_destroy", "qh_decode_codes",
                 "qh_last_error", "qh_host_create", "qh_host_load"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    from qhmat_amd import _native

    assert hasattr(_native.lib(), name)


def test_exports_are_plain_c_symbols():
    from qhmat_amd import _native

    out = subprocess.run(["nm", "-D", "--defined-only", str(_native.library_path())],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    for name in declared_functions():
        assert name in exported, f"{name} is not an unmangled export"


def test_version_and_error_string():
    from qhmat_amd import _native

    lib = _native.lib()
    assert lib.qh_version().decode().count(".") == 2
    status = lib.qh_coefficient(1, ctypes.byref(ctypes.c_double()))
    assert status == _native.QH_EINVAL
    assert "bit width" in lib.qh_last_error().decode()
