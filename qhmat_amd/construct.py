"""Device-side construction of a store from a kernel function.

The reference builds a store block by block with NumPy's SVD on the host
(python/qhmat/matrix.py:205-246).  Here the off-diagonal blocks of one shape
are evaluated, factorised (batched FP64 SVD), rank-selected and checked on
the device in one go; the width allocation and the packing stay in the
library's C++ (the same code the host path uses).  Rank selection, budgets
and feasibility checks follow matrix.py:213-245; the SVD backend differs,
so factors agree with the host construction to rounding, not bit for bit
(the reference pins no SVD output either), and the truncation tail near the
chosen rank is measured from the factors rather than summed from trailing
singular values, which a batched device SVD resolves less accurately.
"""

from __future__ import annotations

from collections import defaultdict
from typing import Callable, Sequence

import numpy as np

from ._native import NativeHMatrix, allocate, spatial_partition
from .matrix import HMatrix, InfeasibleError


def _torch_kernel(kernel_name: str, h: float | None = None) -> Callable:
    import math

    import torch

    def dist(p, q):
        d = p[:, :, None, :] - q[:, None, :, :]
        return torch.sqrt(torch.sum(d * d, dim=-1))

    if kernel_name == "inverse":
        return lambda p, q: 1.0 / (dist(p, q) + 0.4)
    if kernel_name == "exp":
        return lambda p, q: torch.exp(-dist(p, q) / 0.2)
    if kernel_name == "log":
        return lambda p, q: torch.log(dist(p, q) + 1e-2)
    if kernel_name == "single_layer":
        if h is None:
            raise ValueError("single_layer needs the panel size h")

        def sl(p, q):
            r = dist(p, q)
            out = 1.0 / (4.0 * math.pi * r)
            return torch.where(r == 0, torch.full_like(r, 1.0 / (4.0 * math.pi * 0.5 * h)), out)

        return sl
    raise ValueError(f"unknown kernel {kernel_name!r}")


def build_on_device(points: np.ndarray, kernel_name: str, *, tolerance: float,
                    leaf_size: int = 32, admissibility: float = 1.0,
                    bit_widths: Sequence[int] = (4, 8, 12, 16, 24, 32),
                    rank_tolerance: float = 0.1, max_rank: int | None = None,
                    truncation_fraction: float = 0.25, arithmetic_reserve: float = 0.02,
                    support_lower=None, support_upper=None, h: float | None = None,
                    device: str = "cuda", max_batch_bytes: int = 1 << 30) -> HMatrix:
    """Store of kernel(points, points) built with batched device SVDs."""
    import torch

    p = np.ascontiguousarray(points, dtype=np.float64)
    lo = p if support_lower is None else np.asarray(support_lower, dtype=np.float64)
    hi = p if support_upper is None else np.asarray(support_upper, dtype=np.float64)
    part = spatial_partition(p, lo, hi, leaf_size, admissibility)
    perm = np.asarray(part.permutation, dtype=np.int64)
    n = p.shape[0]
    widths = tuple(bit_widths)
    kern = _torch_kernel(kernel_name, h)
    pts = torch.from_numpy(p[perm]).to(device)

    dense_blocks, groups = [], defaultdict(list)
    for i, r in enumerate(part.blocks):
        if r.row0 == r.col0 and r.row1 == r.col1:
            dense_blocks.append(i)
        else:
            groups[(r.row1 - r.row0, r.col1 - r.col0)].append(i)

    results: dict[int, tuple] = {}

    def gather(idx_list, m, c):
        rows = torch.stack([pts[part.blocks[i].row0:part.blocks[i].row0 + m] for i in idx_list])
        cols = torch.stack([pts[part.blocks[i].col0:part.blocks[i].col0 + c] for i in idx_list])
        return kern(rows, cols)

    for i in dense_blocks:
        r = part.blocks[i]
        results[i] = ("dense", gather([i], r.row1 - r.row0, r.col1 - r.col0)[0].cpu().numpy())

    for (m, c), members in groups.items():
        per = max(1, max_batch_bytes // (8 * m * c * 4))
        for s0 in range(0, len(members), per):
            chunk = members[s0:s0 + per]
            a = gather(chunk, m, c)  # (B, m, c)
            budget = tolerance * a.abs().amax(dim=(1, 2))
            u, sigma, vh = torch.linalg.svd(a, full_matrices=False)
            # A batched device SVD need not return the singular values in
            # order; rank selection and the allocation expect them descending.
            sigma, order = torch.sort(sigma, dim=1, descending=True)
            u = torch.gather(u, 2, order[:, None, :].expand(-1, u.shape[1], -1))
            vh = torch.gather(vh, 1, order[:, :, None].expand(-1, -1, vh.shape[2]))
            kmin = sigma.shape[1]
            # tail[:, r] = Frobenius norm of the part discarded at rank r
            sq = torch.flip(torch.cumsum(torch.flip(sigma * sigma, [1]), 1), [1])
            tail = torch.sqrt(torch.cat([sq, torch.zeros_like(sq[:, :1])], dim=1))
            allowance = torch.minimum(rank_tolerance * tail[:, 0], truncation_fraction * budget)
            cap = kmin if max_rank is None else min(max_rank, kmin)
            ok = tail[:, 1:cap + 1] <= allowance[:, None]
            has = ok.any(dim=1)
            first = torch.where(has, ok.float().argmax(dim=1) + 1, torch.zeros_like(has, dtype=torch.long))
            rank = torch.minimum(first, (sigma > 0).sum(dim=1))
            cols = torch.arange(kmin, device=a.device)[None, :]

            def truncated(r):
                keep = (cols < r[:, None]).to(sigma.dtype)
                return a - (u * (sigma * keep)[:, None, :]) @ vh

            # A batched device SVD resolves the trailing singular values less
            # accurately than LAPACK, which can move the first rank meeting the
            # allowance by one.  tail[r] is ||A - A_r||_F, so near the cut it
            # is evaluated directly from the factors (accurate to rounding of
            # ||A||) and the first rank in the window that meets the
            # allowance is taken.
            window = range(-2, 3)
            chosen = rank.clone()
            for d in reversed(window):
                r = torch.clamp(rank + d, 1, cap)
                fits = torch.linalg.matrix_norm(truncated(r)) <= allowance
                chosen = torch.where(has & fits & (r <= (sigma > 0).sum(dim=1)), r, chosen)
            rank = chosen
            # residual max-norm at the chosen rank
            residual = truncated(rank).abs().amax(dim=(1, 2))
            budget_h, has_h = budget.cpu().numpy(), has.cpu().numpy()
            rank_h, res_h = rank.cpu().numpy(), residual.cpu().numpy()
            rmax = int(rank_h.max()) if rank_h.size else 0
            u_h = u[:, :, :rmax].cpu().numpy()
            s_h = sigma[:, :rmax].cpu().numpy()
            v_h = vh[:, :rmax, :].transpose(1, 2).cpu().numpy()
            for t, i in enumerate(chunk):
                rect = part.blocks[i]
                if budget_h[t] == 0.0:
                    results[i] = ("lowrank", (np.zeros((m, 0)), np.empty(0), np.zeros((c, 0)),
                                              [], 0.0))
                    continue
                if not has_h[t]:
                    raise InfeasibleError(f"block ({rect.row0}:{rect.row1}, {rect.col0}:{rect.col1}) "
                                          "cannot satisfy the truncation allowance")
                rk = int(rank_h[t])
                q_budget = budget_h[t] * (1 - arithmetic_reserve) - res_h[t]
                if res_h[t] > truncation_fraction * budget_h[t] or q_budget < 0:
                    raise InfeasibleError("measured truncation error exceeds the block allowance")
                sig = s_h[t, :rk]
                plan = allocate(sig, widths, q_budget, m, c)
                if not plan.feasible:
                    raise InfeasibleError(f"block ({rect.row0}:{rect.row1}, {rect.col0}:{rect.col1}) "
                                          "is infeasible even at the highest bit width")
                results[i] = ("lowrank", (u_h[t, :, :rk], sig, v_h[t, :, :rk], plan.bits,
                                          float(res_h[t])))
            del a, u, sigma, vh

    native = NativeHMatrix(n, part.permutation)
    for i, rect in enumerate(part.blocks):
        kind, payload = results.pop(i)
        if kind == "dense":
            native.add_dense(rect, payload)
        else:
            uu, sig, vv, bits, residual = payload
            native.add_low_rank(rect, uu, sig, vv, bits, len(widths), residual)
    native.finalize()
    return HMatrix(native, source_array_bytes=n * n * 8)
