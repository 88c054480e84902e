"""Application harness after the reference's method comparison
(experiments/cuda/application.cu:189-222, experiments/APPLICATION_BENCHMARK.md:89-98)
on one MI355X: y = A X for one resident store, per method:

  dense_fp64, dense_fp32        torch.mm (rocBLAS / hipBLASLt GEMM) on the
                                reconstructed matrix in source order
                                (the reference's ``dense_fp32``)
  h_fp32_batched                per exact (m, n, rank) shape bucket, two
                                batched GEMMs (torch.bmm) on the unquantized
                                FP32 balanced factors kept at construction,
                                diagonal blocks batched the same way, one
                                permutation gather and one scatter
  unpacked_batched_fp32         the same GEMMs, every factor decoded from the
                                resident codes first, in one launch per
                                application (qh_factor_plan_decode; the
                                reference's ``adaptive_unpacked_batched``)
  predecoded_batched_fp32       the same GEMMs on factors decoded once, outside
                                the timing (``adaptive_predecoded_batched``)
  packed_fused_fp32split,       this library's apply on the packed codes,
  packed_fused_fp64             decoding every time (``adaptive_packed_fused``)

``h_fp32_fused`` (the reference's tiled kernels on unquantized factors) has
no counterpart here: this library's fused kernels read packed codes only.
Each result reports its relative error against the FP64 fused apply.
"""

from __future__ import annotations

import collections
import ctypes as C

import numpy as np

from . import _native
from . import device as _device


class Harness:
    """The methods of one store, set up once (outside the timings): the dense
    matrix in FP64 and FP32, the exact-shape buckets, a factor-decode plan
    and the decoded factors; with `factors` (block index -> balanced FP64
    factors (u sqrt(sigma), v sqrt(sigma)) kept at construction,
    workloads.build_kernel_hmatrix(keep_factors=True)), the unquantized FP32
    factors as well."""

    def __init__(self, native, device: int = 0, factors: dict | None = None):
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
        groups = collections.defaultdict(lambda: {"blocks": [], "rows": [], "cols": [], "dense": []})
        for index, info in enumerate(native.blocks):
            r = info.rectangle
            m, c = r.row1 - r.row0, r.col1 - r.col0
            if not info.dense and info.rank == 0:
                continue  # contributes nothing
            g = groups[(m, c, 0 if info.dense else info.rank)]
            g["blocks"].append(index)
            g["rows"].append(r.row0)
            g["cols"].append(r.col0)
            if info.dense:
                g["dense"].append(clustered[r.row0:r.row1, r.col0:r.col1].float())
        del clustered
        # factor layout: per low-rank bucket, count x m x r left factors then
        # count x r x c transposed right factors, in one FP32 buffer
        blocks, sides, offsets = [], [], []
        self.buckets = []  # (rows index, cols index, dense stack | None, (off_u, off_vt, shape))
        total = 0
        for (m, c, rank), g in groups.items():
            count = len(g["blocks"])
            rows = (torch.tensor(g["rows"], device=dev)[:, None] + torch.arange(m, device=dev)).reshape(-1)
            cols = torch.tensor(g["cols"], device=dev)[:, None] + torch.arange(c, device=dev)
            if rank == 0:
                self.buckets.append((rows, cols, torch.stack(g["dense"]), None))
                continue
            off_u, off_vt = total, total + count * m * rank
            for i, b in enumerate(g["blocks"]):
                blocks += [b, b]
                sides += [0, 1]
                offsets += [off_u + i * m * rank, off_vt + i * rank * c]
            total = off_vt + count * rank * c
            self.buckets.append((rows, cols, None, (off_u, off_vt, count, m, c, rank)))
        self.factor_floats = total
        self.plan = _native.FactorPlan(native, blocks, sides, offsets, max(total, 1), device)
        self.live = torch.empty(max(total, 1), dtype=torch.float32, device=dev)
        self.plan.decode(self.live)
        self.predecoded_factors = self.live.clone()
        self.fp32_factors = None
        if factors is not None:
            host = np.zeros(max(total, 1), dtype=np.float32)
            for b, side, off in zip(blocks, sides, offsets):
                u, v = factors[b]
                f = np.asarray(u if side == 0 else v, dtype=np.float32)
                a = f if side == 0 else f.T
                host[off:off + a.size] = a.reshape(-1)
            self.fp32_factors = torch.from_numpy(host).to(dev)
        torch.cuda.synchronize()

    def _apply(self, x32, buf):
        """y = A x over the buckets, low-rank factors from `buf`: x32 (n, k)
        FP32 in source order; one gather, two batched GEMMs per low-rank
        bucket (one per dense bucket), one scatter."""
        torch = self.torch
        xc = x32[self.perm]
        yc = torch.zeros_like(xc)
        k = x32.shape[1]
        for rows, cols, dense, lay in self.buckets:
            xb = xc[cols]                                   # (count, n_b, k)
            if dense is not None:
                yb = torch.bmm(dense, xb)
            else:
                off_u, off_vt, count, m, c, r = lay
                u = buf[off_u:off_u + count * m * r].view(count, m, r)
                vt = buf[off_vt:off_vt + count * r * c].view(count, r, c)
                yb = torch.bmm(u, torch.bmm(vt, xb))
            yc.index_add_(0, rows, yb.reshape(-1, k))
        y = torch.empty_like(yc)
        y[self.perm] = yc
        return y

    def batched(self, x32):
        """predecoded_batched_fp32: the factors decoded once, at set-up."""
        return self._apply(x32, self.predecoded_factors)

    def unpacked(self, x32):
        """unpacked_batched_fp32: every factor decoded from the resident
        codes in one launch per application, then the batched GEMMs."""
        self.plan.decode(self.live)
        return self._apply(x32, self.live)

    def fp32_batched(self, x32):
        """h_fp32_batched: the unquantized FP32 factors, no decoding."""
        return self._apply(x32, self.fp32_factors)

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
            "unpacked_batched_fp32": lambda: self.unpacked(x32),
            "packed_fused_fp32split": lambda: fused(_native.PRECISION_FP32SPLIT),
            "packed_fused_fp64": lambda: fused(_native.PRECISION_FP64),
        }
        if self.fp32_factors is not None:
            methods["h_fp32_batched"] = lambda: self.fp32_batched(x32)
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
