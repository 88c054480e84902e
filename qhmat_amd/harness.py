"""Application harness after the reference's method comparison
(experiments/cuda/application.cu:189-222, experiments/APPLICATION_BENCHMARK.md:89-98)
on one MI355X: y = A X for one resident store, per method:

  dense_fp64, dense_fp32        torch.mm (rocBLAS / hipBLASLt GEMM) on the
                                reconstructed matrix in source order
                                (the reference's ``dense_fp32``)
  predecoded_batched_fp32       per exact (m, n, rank) shape bucket, two
                                batched GEMMs (torch.bmm) on the decoded FP32
                                factors, diagonal blocks batched the same way,
                                one permutation gather and one scatter (the
                                reference's ``adaptive_predecoded_batched``:
                                decoding done once, outside the timing)
  packed_fused_fp32split,       this library's apply on the packed codes,
  packed_fused_fp64             decoding every time (``adaptive_packed_fused``)

The reference's unquantized-factor methods (``h_fp32_*``) have no
counterpart: a store holds only the packed factors.  Each result reports its
relative error against the FP64 fused apply.
"""

from __future__ import annotations

import collections
import ctypes as C

import numpy as np

from . import _native
from . import device as _device


class Harness:
    def __init__(self, native, device: int = 0):
        import torch

        self.torch = torch
        self.native = native
        n = native.order
        self.n = n
        dev = f"cuda:{device}"
        perm = torch.from_numpy(native.permutation.astype(np.int64)).to(dev)  # clustered -> source
        self.perm = perm
        self.dense64 = native.reconstruct_device()
        self.dense32 = self.dense64.float()
        clustered = self.dense64[perm][:, perm]
        buckets = collections.defaultdict(lambda: {"u": [], "vt": [], "rows": [], "cols": []})
        for index, info in enumerate(native.blocks):
            r = info.rectangle
            m, c = r.row1 - r.row0, r.col1 - r.col0
            b = buckets[(m, c, 0 if info.dense else info.rank)]
            b["rows"].append(r.row0)
            b["cols"].append(r.col0)
            if info.dense:
                b["u"].append(clustered[r.row0:r.row1, r.col0:r.col1].float())
            elif info.rank > 0:
                # decoded factor sides (rank x rows): codes from the device,
                # times each component's scale
                fac = [torch.from_numpy(native.decode_codes(index, side).astype(np.float64))
                       * torch.from_numpy(native.block_words(index, side)[1])[:, None]
                       for side in (0, 1)]
                b["u"].append(fac[0].T.contiguous().float().to(dev))   # m x r
                b["vt"].append(fac[1].contiguous().float().to(dev))    # r x n
        del clustered
        self.buckets = []
        for (m, c, rank), b in buckets.items():
            if not b["u"]:
                continue  # rank-0 blocks contribute nothing
            rows = torch.tensor(b["rows"], device=dev)[:, None] + torch.arange(m, device=dev)
            cols = torch.tensor(b["cols"], device=dev)[:, None] + torch.arange(c, device=dev)
            u = torch.stack(b["u"])
            vt = torch.stack(b["vt"]) if b["vt"] else None
            self.buckets.append((u, vt, rows.reshape(-1), cols))

    def batched(self, x32):
        """predecoded_batched_fp32: x32 (n, k) FP32 in source order."""
        torch = self.torch
        xc = x32[self.perm]
        yc = torch.zeros_like(xc)
        k = x32.shape[1]
        for u, vt, rows, cols in self.buckets:
            xb = xc[cols]                                   # (count, n_b, k)
            yb = torch.bmm(u, torch.bmm(vt, xb)) if vt is not None else torch.bmm(u, xb)
            yc.index_add_(0, rows, yb.reshape(-1, k))
        y = torch.empty_like(yc)
        y[self.perm] = yc
        return y

    def run(self, k: int, repeats: int = 20, warmup: int = 3, seed: int = 1) -> dict:
        torch = self.torch
        n = self.n
        gen = torch.Generator(device="cpu").manual_seed(seed)
        x64 = torch.randn(n, k, dtype=torch.float64, generator=gen).cuda()
        x32 = x64.float()
        ws = torch.empty(max(_device.workspace_bytes(self.native, k), 8), dtype=torch.uint8,
                         device="cuda")
        y64 = torch.empty_like(x64)
        # the fused apply through the C ABI with its arguments built once, as a
        # host code calling it in a loop would
        lib = _native.lib()
        store = self.native.device_store(torch.cuda.current_device())
        args = (C.c_void_p(x64.data_ptr()), x64.stride(0), k, C.c_void_p(y64.data_ptr()),
                y64.stride(0))
        tail = (C.c_void_p(torch.cuda.current_stream().cuda_stream), C.c_void_p(ws.data_ptr()),
                ws.numel())

        def fused(precision):
            _native.check(lib.qh_apply(store, *args, precision, *tail))
            return y64

        methods = {
            "dense_fp64": lambda: torch.mm(self.dense64, x64),
            "dense_fp32": lambda: torch.mm(self.dense32, x32),
            "predecoded_batched_fp32": lambda: self.batched(x32),
            "packed_fused_fp32split": lambda: fused(_native.PRECISION_FP32SPLIT),
            "packed_fused_fp64": lambda: fused(_native.PRECISION_FP64),
        }
        ref = _device.apply(self.native, x64, precision="fp64").clone()
        out = {}
        for name, fn in methods.items():
            for _ in range(warmup):
                fn()
            torch.cuda.synchronize()
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(repeats):
                y = fn()
            end.record()
            torch.cuda.synchronize()
            ms = start.elapsed_time(end) / repeats
            err = float(torch.linalg.norm(y.double() - ref) / torch.linalg.norm(ref))
            out[name] = {"ms": round(ms, 5), "matvecs_per_s": round(k / (ms * 1e-3), 1),
                         "rel_err_vs_fp64_apply": float(f"{err:.3g}")}
        return out
