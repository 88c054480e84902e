"""MI355X-native apply of adaptive integer-quantized hierarchical matrices.

Drop-in for the reference package ``qhmat`` (python/qhmat/__init__.py): same
names, with ``device="cuda"`` served by hand-written gfx950 HIP kernels over a
device-resident packed store.
"""

from ._native import (Allocation, BlockInfo, NativeHMatrix, Partition, Rectangle,
                      coefficient, cuda_built, spatial_partition)
from ._native import allocate as allocate_bits
from .matrix import HMatrix, InfeasibleError, StorageReport, memory_bytes

__version__ = "0.1.0"

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
