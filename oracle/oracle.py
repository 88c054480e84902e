"""ctypes access to the CPU checkers -- TEST INFRASTRUCTURE ONLY.

  * liboracle.so  -- the plain-C restatement of the reference FP64 path
                     (oracle.c; function list and reference lines in oracle.h)
  * _ref/libqhmat_ref.so -- the reference's own core.cpp/io.cpp, when built

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, always as the checker or the timed CPU baseline, never as the product.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from . import qhmat001

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libqhmat_ref.so"
REFERENCE_SRC = Path("/root/reference/src/qhmat")

_p = C.c_void_p
_sz = C.c_size_t


class _OrcStore(C.Structure):
    _fields_ = [
        ("order", _sz), ("perm", _p), ("nblocks", _sz), ("rect", _p), ("is_dense", _p),
        ("dense", _p), ("dense_off", _p), ("rank", _p), ("comp_off", _p), ("widths", _p),
        ("lscale", _p), ("rscale", _p), ("lwords", _p), ("rwords", _p), ("loff", _p),
        ("roff", _p),
    ]


def build(reference: bool | None = None) -> None:
    """make -C oracle (and the reference build when its sources exist)."""
    targets = ["all"]
    if reference or (reference is None and REFERENCE_SRC.exists()):
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", str(HERE), *targets], check=True)


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            build(reference=False)
        _lib = C.CDLL(str(ORACLE_SO))
        _lib.orc_limit.restype = C.c_uint64
        _lib.orc_limit.argtypes = [C.c_int]
        _lib.orc_coefficient.restype = C.c_double
        _lib.orc_coefficient.argtypes = [C.c_int]
        _lib.orc_words_for.restype = _sz
        _lib.orc_words_for.argtypes = [_sz, C.c_int]
        _lib.orc_pack_component.restype = None
        _lib.orc_pack_component.argtypes = [_p, _sz, C.c_int, _p, C.POINTER(C.c_double)]
        _lib.orc_decode_code.restype = C.c_int64
        _lib.orc_decode_code.argtypes = [_p, _sz, C.c_int]
        _lib.orc_decode.restype = C.c_double
        _lib.orc_decode.argtypes = [_p, C.c_double, _sz, C.c_int]
        _lib.orc_decode_factor.restype = None
        _lib.orc_decode_factor.argtypes = [_p, _sz, _p, _sz, _p]
        _lib.orc_matmul.restype = C.c_int
        _lib.orc_matmul.argtypes = [C.POINTER(_OrcStore), _p, _sz, _p]
        _lib.orc_reconstruct.restype = None
        _lib.orc_reconstruct.argtypes = [C.POINTER(_OrcStore), _p]
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_p) if a.size else None


def coefficient(bits: int) -> float:
    return float(lib().orc_coefficient(bits))


def limit(bits: int) -> int:
    return int(lib().orc_limit(bits))


def pack_component(column: np.ndarray, bits: int) -> tuple[np.ndarray, float]:
    col = np.ascontiguousarray(column, dtype=np.float64)
    words = np.zeros(qhmat001.words_for(col.size, bits), dtype=np.uint32)
    scale = C.c_double()
    lib().orc_pack_component(_ptr(col), col.size, bits, _ptr(words), C.byref(scale))
    return words, float(scale.value)


def pack_factor(factor: np.ndarray, widths) -> tuple[np.ndarray, np.ndarray]:
    """Packs a rows x r balanced factor; returns (words, scales)."""
    words, scales = [], []
    for j, b in enumerate(widths):
        w, s = pack_component(factor[:, j], int(b))
        words.append(w)
        scales.append(s)
    if not words:
        return np.zeros(0, np.uint32), np.zeros(0)
    return np.concatenate(words), np.asarray(scales)


def decode_codes(words: np.ndarray, rows: int, widths) -> np.ndarray:
    """(rank, rows) int64 codes of a factor, unpacked by the oracle."""
    # one spare word: the decoder may look one word past the last component
    w = np.ascontiguousarray(np.append(np.asarray(words, dtype=np.uint32), np.uint32(0)))
    wd = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
    out = np.zeros((len(widths), rows), dtype=np.int64)
    if out.size:
        lib().orc_decode_factor(_ptr(w), rows, _ptr(wd), wd.size, _ptr(out))
    return out


class OracleStore:
    """A parsed QHMAT001 store laid out for orc_matmul / orc_reconstruct."""

    def __init__(self, parsed: dict):
        self.order = int(parsed["order"])
        blocks = parsed["blocks"]
        nb = len(blocks)
        self.perm = np.ascontiguousarray(parsed["perm"], dtype=np.uint64)
        self.rect = np.zeros(4 * nb, dtype=np.uint64)
        self.is_dense = np.zeros(nb, dtype=np.uint8)
        self.dense_off = np.zeros(nb, dtype=np.uint64)
        self.rank = np.zeros(nb, dtype=np.uint64)
        self.comp_off = np.zeros(nb, dtype=np.uint64)
        dense, widths, lsc, rsc, lw, rw, loff, roff = [], [], [], [], [], [], [], []
        nd = ncomp = nl = nr = 0
        for b, blk in enumerate(blocks):
            r0, r1, c0, c1 = blk["rect"]
            self.rect[4 * b:4 * b + 4] = (r0, r1, c0, c1)
            if blk["dense"]:
                self.is_dense[b] = 1
                self.dense_off[b] = nd
                dense.append(np.ravel(blk["values"]))
                nd += blk["values"].size
                continue
            r = len(blk["widths"])
            self.rank[b] = r
            self.comp_off[b] = ncomp
            ncomp += r
            widths.append(np.asarray(blk["widths"], dtype=np.int32))
            lsc.append(blk["lscale"])
            rsc.append(blk["rscale"])
            loff.append(nl + qhmat001.component_offsets(r1 - r0, blk["widths"]))
            roff.append(nr + qhmat001.component_offsets(c1 - c0, blk["widths"]))
            lw.append(blk["lwords"])
            rw.append(blk["rwords"])
            nl += blk["lwords"].size
            nr += blk["rwords"].size

        def cat(parts, dtype):
            arrays = [np.asarray(p, dtype=dtype).ravel() for p in parts]
            out = np.concatenate(arrays) if arrays else np.zeros(0, dtype)
            return np.ascontiguousarray(out if out.size else np.zeros(1, dtype))

        self.dense = cat(dense, np.float64)
        self.widths = cat(widths, np.int32)
        self.lscale = cat(lsc, np.float64)
        self.rscale = cat(rsc, np.float64)
        # one spare word: the decoder may look one word past a component
        self.lwords = np.append(cat(lw, np.uint32), np.zeros(1, np.uint32))
        self.rwords = np.append(cat(rw, np.uint32), np.zeros(1, np.uint32))
        self.loff = cat(loff, np.uint64)
        self.roff = cat(roff, np.uint64)
        self._c = _OrcStore(
            self.order, _ptr(self.perm), nb, _ptr(self.rect), _ptr(self.is_dense),
            _ptr(self.dense), _ptr(self.dense_off), _ptr(self.rank), _ptr(self.comp_off),
            _ptr(self.widths), _ptr(self.lscale), _ptr(self.rscale), _ptr(self.lwords),
            _ptr(self.rwords), _ptr(self.loff), _ptr(self.roff))

    @classmethod
    def from_file(cls, path, verify_checksum: bool = True) -> "OracleStore":
        return cls(qhmat001.parse(Path(path).read_bytes(), verify_checksum))

    def matmul(self, x: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.float64)
        vec = a.ndim == 1
        a2 = a.reshape(self.order, -1)
        y = np.zeros_like(a2)
        if lib().orc_matmul(C.byref(self._c), _ptr(a2), a2.shape[1], _ptr(y)) != 0:
            raise ValueError("nonfinite right-hand side")
        return y.reshape(self.order) if vec else y

    def matmul_columns(self, x: np.ndarray, workers: int | None = None) -> np.ndarray:
        """matmul one right-hand-side column at a time on a thread pool.

        In the reference loop nest (core.cpp:360-390) column p's arithmetic
        reads and writes column p only, in an order that does not depend on
        the column count, so the result is bitwise matmul(x); the columns run
        in parallel because ctypes releases the interpreter lock."""
        import os
        from concurrent.futures import ThreadPoolExecutor

        a = np.asarray(x, dtype=np.float64).reshape(self.order, -1)
        cols = [np.ascontiguousarray(a[:, p]) for p in range(a.shape[1])]
        workers = workers or min(len(cols), os.cpu_count() or 1, 16)
        with ThreadPoolExecutor(max_workers=max(workers, 1)) as pool:
            out = list(pool.map(self.matmul, cols))
        return np.stack(out, axis=1).reshape(np.shape(x))

    def reconstruct(self) -> np.ndarray:
        out = np.zeros((self.order, self.order))
        lib().orc_reconstruct(C.byref(self._c), _ptr(out))
        return out


# -- the reference's compiled core (oracle/_ref) --------------------------------
_ref = None


def ref_available() -> bool:
    return REF_SO.exists()


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            raise FileNotFoundError(f"{REF_SO} is not built (make -C oracle ref)")
        r = C.CDLL(str(REF_SO))
        r.ref_last_error.restype = C.c_char_p
        r.ref_create.restype = _p
        r.ref_create.argtypes = [_sz, _p]
        r.ref_destroy.argtypes = [_p]
        r.ref_add_dense.argtypes = [_p, _sz, _sz, _sz, _sz, _p]
        r.ref_add_low_rank.argtypes = [_p, _sz, _sz, _sz, _sz, C.c_int, _p, _p, _sz, _p, _p,
                                       _sz, C.c_double]
        r.ref_finalize.argtypes = [_p]
        r.ref_matmul.argtypes = [_p, _p, _sz, _p]
        r.ref_reconstruct.argtypes = [_p, _p]
        r.ref_reconstruction_bound.restype = C.c_double
        r.ref_reconstruction_bound.argtypes = [_p]
        r.ref_logical_bits.restype = C.c_uint64
        r.ref_logical_bits.argtypes = [_p]
        r.ref_allocated_bytes.restype = C.c_uint64
        r.ref_allocated_bytes.argtypes = [_p]
        r.ref_order.restype = _sz
        r.ref_order.argtypes = [_p]
        r.ref_save.argtypes = [_p, C.c_char_p]
        r.ref_load.restype = _p
        r.ref_load.argtypes = [C.c_char_p]
        r.ref_coefficient.restype = C.c_double
        r.ref_coefficient.argtypes = [C.c_int]
        r.ref_allocate.argtypes = [_p, _sz, _p, _sz, C.c_double, _sz, _sz, _p,
                                   C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                   C.POINTER(C.c_int)]
        r.ref_partition.restype = _p
        r.ref_partition.argtypes = [_p, _p, _p, _sz, _sz, _sz, C.c_double]
        r.ref_partition_blocks.restype = _sz
        r.ref_partition_blocks.argtypes = [_p]
        r.ref_partition_copy.argtypes = [_p, _p, _p]
        r.ref_partition_free.argtypes = [_p]
        _ref = r
    return _ref


def _ref_check(status: int) -> None:
    if status != 0:
        raise RuntimeError(ref().ref_last_error().decode())


class RefHMatrix:
    """The reference HMatrix (core.hpp:52-102) through oracle/_ref."""

    def __init__(self, order: int = 0, perm=None, *, _handle=None):
        r = ref()
        if _handle is not None:
            self._h = _handle
            return
        p = np.ascontiguousarray(np.asarray(perm, dtype=np.uint64))
        self._h = r.ref_create(order, _ptr(p))
        if not self._h:
            raise RuntimeError(r.ref_last_error().decode())

    def __del__(self):
        # module globals may already be gone at interpreter shutdown
        if getattr(self, "_h", None) and callable(ref):
            ref().ref_destroy(self._h)
            self._h = None

    def add_dense(self, rect, values) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        _ref_check(ref().ref_add_dense(self._h, *rect[:4], _ptr(v)))

    def add_low_rank(self, rect, u, sigma, v, widths, menu_size, truncation_error,
                     admissible=False) -> None:
        s = np.ascontiguousarray(sigma, dtype=np.float64)
        uc = np.ascontiguousarray(np.asarray(u, dtype=np.float64).T)
        vc = np.ascontiguousarray(np.asarray(v, dtype=np.float64).T)
        w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
        _ref_check(ref().ref_add_low_rank(self._h, *rect[:4], int(admissible), _ptr(uc),
                                          _ptr(s), s.size, _ptr(vc), _ptr(w), int(menu_size),
                                          float(truncation_error)))

    def finalize(self) -> None:
        _ref_check(ref().ref_finalize(self._h))

    @property
    def order(self) -> int:
        return int(ref().ref_order(self._h))

    def matmul(self, x: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.float64)
        a2 = a.reshape(self.order, -1)
        y = np.zeros_like(a2)
        _ref_check(ref().ref_matmul(self._h, _ptr(a2), a2.shape[1], _ptr(y)))
        return y.reshape(a.shape)

    def matmul_columns(self, x: np.ndarray, workers: int | None = None) -> np.ndarray:
        """matmul one column at a time on a thread pool: bitwise matmul(x),
        since core.cpp:360-390 reads and writes only column p for column p
        (the same argument as OracleStore.matmul_columns)."""
        import os
        from concurrent.futures import ThreadPoolExecutor

        a = np.asarray(x, dtype=np.float64).reshape(self.order, -1)
        cols = [np.ascontiguousarray(a[:, p]) for p in range(a.shape[1])]
        workers = workers or min(len(cols), os.cpu_count() or 1, 16)
        with ThreadPoolExecutor(max_workers=max(workers, 1)) as pool:
            out = list(pool.map(self.matmul, cols))
        return np.stack(out, axis=1).reshape(np.shape(x))

    def reconstruct(self) -> np.ndarray:
        out = np.zeros((self.order, self.order))
        _ref_check(ref().ref_reconstruct(self._h, _ptr(out)))
        return out

    @property
    def reconstruction_bound(self) -> float:
        return float(ref().ref_reconstruction_bound(self._h))

    @property
    def logical_bits(self) -> int:
        return int(ref().ref_logical_bits(self._h))

    @property
    def allocated_bytes(self) -> int:
        return int(ref().ref_allocated_bytes(self._h))

    def save(self, path) -> None:
        _ref_check(ref().ref_save(self._h, str(path).encode()))

    @classmethod
    def load(cls, path) -> "RefHMatrix":
        h = ref().ref_load(str(path).encode())
        if not h:
            raise RuntimeError(ref().ref_last_error().decode())
        return cls(_handle=h)


def ref_allocate(sigma, menu, budget, rows, cols):
    s = np.ascontiguousarray(sigma, dtype=np.float64)
    m = np.ascontiguousarray(np.asarray(menu, dtype=np.int32))
    bits = np.zeros(max(s.size, 1), dtype=np.int32)
    bound, payload, feasible = C.c_double(), C.c_uint64(), C.c_int()
    _ref_check(ref().ref_allocate(_ptr(s), s.size, _ptr(m), m.size, float(budget), rows, cols,
                                  _ptr(bits), C.byref(bound), C.byref(payload),
                                  C.byref(feasible)))
    return [int(b) for b in bits[:s.size]], float(bound.value), int(payload.value), bool(feasible.value)


def ref_partition(points, lower, upper, leaf_size, eta):
    p = np.ascontiguousarray(points, dtype=np.float64)
    lo = np.ascontiguousarray(lower, dtype=np.float64)
    hi = np.ascontiguousarray(upper, dtype=np.float64)
    n, d = p.shape
    h = ref().ref_partition(_ptr(p), _ptr(lo), _ptr(hi), n, d, leaf_size, float(eta))
    if not h:
        raise RuntimeError(ref().ref_last_error().decode())
    try:
        nb = ref().ref_partition_blocks(h)
        perm = np.zeros(n, dtype=np.uint64)
        rects = np.zeros(5 * nb, dtype=np.uint64)
        ref().ref_partition_copy(h, _ptr(perm), _ptr(rects))
    finally:
        ref().ref_partition_free(h)
    return perm.astype(np.int64), rects.reshape(-1, 5).astype(np.int64)
