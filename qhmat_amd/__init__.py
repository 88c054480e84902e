"""MI355X-native apply of adaptive integer-quantized hierarchical matrices.

Drop-in for the reference package ``qhmat`` (python/qhmat/__init__.py): same
names, with ``device="cuda"`` served by hand-written gfx950 HIP kernels over a
device-resident packed store.
"""

from .blockwork import InfeasibleError

__version__ = "0.1.0"

# The names below load the native library on first use, not at package import:
# the blockwise constructor's worker processes import qhmat_amd.blockwork and
# must never load the HIP runtime.
_NATIVE = {"Allocation", "BlockInfo", "NativeHMatrix", "Partition", "Rectangle", "coefficient",
           "cuda_built", "spatial_partition"}
_MATRIX = {"HMatrix", "StorageReport", "memory_bytes"}


def __getattr__(name: str):
    if name in _NATIVE or name == "allocate_bits":
        from . import _native

        value = _native.allocate if name == "allocate_bits" else getattr(_native, name)
    elif name in _MATRIX:
        from . import matrix

        value = getattr(matrix, name)
    else:
        raise AttributeError(f"module 'qhmat_amd' has no attribute {name!r}")
    globals()[name] = value
    return value

__all__ = [
    "Allocation",
    "BlockInfo",
    "HMatrix",
    "InfeasibleError",
    "NativeHMatrix",
    "Partition",
    "Rectangle",
    "StorageReport",
    "allocate_bits",
    "coefficient",
    "cuda_built",
    "memory_bytes",
    "spatial_partition",
]
