// MI355X (gfx950) apply of a qhmat H-matrix, y = H x.
//
// The packed store is uploaded once as a list of work items (layout in
// device_store.hpp).  An apply is three launches on the caller's stream, all
// deterministic (fixed summation order, one writer per output element, no
// floating-point atomics):
//
//  permute x in clustered order into the workspace.
//  pass 1  four workgroups of four waves per row tile, each wave running one
//          of the tile's 16 item runs: the tile's diagonal (kDense) and
//          fused (kFused: z and U z inside the item) contributions, summed
//          in a fixed tree per workgroup into one of four row-sum sets of
//          the workspace, and a share of the kRight items: partial sums
//          z[c-tile][j] = sum_{c in tile} codeR(c, j) x[c] into the workspace.
//  pass 2  one workgroup of sixteen waves per row tile over the tile's kLeft
//          items (U z with z summed over the block's column tiles); it adds
//          its waves and the four pass-1 sets in a fixed tree and writes
//          y[perm[row]] once.
//
// Every wave walks a contiguous list of items with the data of the next two
// items (<= 4 KiB each, four 16-byte buffer loads per lane), their component
// widths, scales and inputs in flight while the current item is decoded from
// LDS.  The loads are buffer-resource loads: a lane past an item's end reads
// zero instead of branching, and a wave past its run's end fetches an empty
// item, so the loop has no load under a branch and no load waits on another.
//
// Decoding: a code is v_alignbit_b32 of two staged words and a sign-extending
// shift pair, branch-free for every width.  FP64 mode computes in FP64
// throughout.  FP32 split mode converts codes of at most 24 bits and x to
// FP32 (exact for the codes); where wider codes occur (per group of
// components) it splits every code into a high half and a sign-extended low
// 16-bit half, both exact in FP32, accumulated apart and joined in FP64.  An
// item's contribution to a row is summed in FP32; everything across items,
// the scales and the diagonal stay in FP64.
//
// Reference: src/qhmat/cuda.cu:64-164 (one thread per output entry scanning
// every block, store uploaded per call) and core.cpp:360-390 (FP64 loop nest).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <set>
#include <tuple>
#include <mutex>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "device_common.hpp"
#include "device_store.hpp"
#include "wide_passes.hpp"

namespace qhamd {
namespace {


// LDS of one wave: a head word that decoding may read below the first
// stream, the current item's data plus slack that lanes decoding past the end
// of a stream may read (64 codes of 32 bits and the word after), then the
// item's inputs (x rows of its columns, or the workspace partial sums of
// kLeft), KB values per lane, then the item's scale products and widths
// (lanes = components).
template <int KB> constexpr int wave_lds_meta() { return kLdsHead + kStageBytes + kTile * KB * 8; }
// Narrow passes in FP32 mode also keep the x rows in FP32 (lanes = columns).
template <int KB> constexpr int wave_lds_xf() { return wave_lds_meta<KB>() + 64 * 12; }
template <int KB> constexpr int wave_lds_bytes() {
    return wave_lds_xf<KB>() + 64 * KB * 4;
}

// Uniform fields of item `slot` of the record group held in r0/r1 (lanes =
// items of the group).
struct Item {
    std::uint32_t quad, quads, kind, count, comp0, col, cols, zrow, zstride, nseg;
};

__device__ __forceinline__ Item item_at(const quad_t& r0, const quad_t& r1, std::uint32_t slot) {
    Item f;
    f.quad = __builtin_amdgcn_readlane(r0.x, slot);
    const std::uint32_t packed = __builtin_amdgcn_readlane(r0.y, slot);
    f.quads = packed & 0xFFFFu;
    f.kind = (packed >> 16) & 0xFFu;
    f.count = min(packed >> 24, std::uint32_t(kItemComponents));
    f.comp0 = __builtin_amdgcn_readlane(r0.z, slot);
    f.col = __builtin_amdgcn_readlane(r0.w, slot);
    f.cols = min(std::uint32_t(__builtin_amdgcn_readlane(r1.x, slot)), std::uint32_t(kTile));
    f.zrow = __builtin_amdgcn_readlane(r1.y, slot);
    f.zstride = __builtin_amdgcn_readlane(r1.z, slot);
    f.nseg = max(std::uint32_t(__builtin_amdgcn_readlane(r1.w, slot)), 1u);
    f.quads = min(f.quads, std::uint32_t(kItemQuads));
    return f;
}

// The fields of a record held in scalar registers (a wave's head record).
__device__ __forceinline__ Item item_from(const ItemDev& d) {
    Item f;
    f.quad = d.quad;
    f.quads = min(std::uint32_t(d.quads), std::uint32_t(kItemQuads));
    f.kind = d.kind;
    f.count = min(std::uint32_t(d.count), std::uint32_t(kItemComponents));
    f.comp0 = d.comp0;
    f.col = d.col;
    f.cols = min(d.cols, std::uint32_t(kTile));
    f.zrow = d.zrow;
    f.zstride = d.zstride;
    f.nseg = max(d.nseg, 1u);
    return f;
}

// Columns of x an item reads (lanes = columns).
__device__ __forceinline__ std::uint32_t x_columns(const Item& f) {
    return f.kind == kDense ? min(f.count, std::uint32_t(kDenseColumns))
                            : (f.kind == kLeft ? 0u : f.cols);
}

// Workspace partial sums a kLeft item reads (lanes = (component, column tile)).
__device__ __forceinline__ std::uint32_t z_pairs(const Item& f) {
    return f.kind == kLeft ? min(f.count * f.nseg, 64u) : 0u;
}

// Registers of the item being fetched.
template <int KB> struct Fetch {
    quad_t q[kItemQuads / 64];
    std::uint32_t bits;  // lanes = components
    double scale;        // lanes = components (kFused, kLeft)
    double in[KB];       // x rows (lanes = columns) or partial sums (lanes = pairs)
};


// Row sets per tile in the workspace (pass 1's four workgroups), and x in
// clustered order after them: xc[i][p] = x[perm[i]][p], written once per
// apply (permute_tile) so that an item's x rows are one contiguous run.
constexpr int kRowSets = kPass1Groups;
__device__ __forceinline__ double* clustered_x(const MemberView& M) {
    return M.zp + (M.z_rows + std::uint64_t(kRowSets) * M.order) * std::uint64_t(M.k);
}

template <int KB>
__device__ __forceinline__ void fetch(const MemberView& M, const Item& f, int lane, int p0, int kk,
                                      Fetch<KB>& F) {
    const __amdgpu_buffer_rsrc_t data =
        buffer(M.words + 4ull * f.quad, f.quads * 16u);
#pragma unroll
    for (int t = 0; t < kItemQuads / 64; ++t)
        F.q[t] = __builtin_amdgcn_raw_buffer_load_b128(data, std::uint32_t(lane + 64 * t) * 16u, 0, 0);
    const std::uint32_t ncomp = f.kind == kDense ? 0u : f.count;
    F.bits = __builtin_amdgcn_raw_buffer_load_b32(buffer(M.widths + f.comp0, ncomp * 4u),
                                                  std::uint32_t(lane) * 4u, 0, 0);
    const std::uint32_t nscale = (f.kind == kFused || f.kind == kLeft) ? f.count : 0u;
    F.scale = load_f64(buffer(M.cscale + f.comp0, nscale * 8u), std::uint32_t(lane) * 8u);
    // clustered x and the partial sums: byte offsets in the workspace (the
    // host keeps it below kOutOfRange bytes)
    const std::uint32_t k = std::uint32_t(M.k);
    std::uint32_t row;
    bool mine;
    __amdgpu_buffer_rsrc_t r;
    if (f.kind == kLeft) {
        const std::uint32_t s = lane % f.nseg, j = lane / f.nseg;
        mine = std::uint32_t(lane) < z_pairs(f);
        row = (f.zrow + s * f.zstride + j) * k;
        r = buffer(M.zp, kOutOfRange);
    } else {
        mine = lane < int(x_columns(f));
        row = (f.col + std::uint32_t(lane)) * k;
        r = buffer(clustered_x(M), kOutOfRange);
    }
#pragma unroll
    for (int p = 0; p < KB; ++p)
        F.in[p] = load_f64(r, (mine && p < kk) ? (row + std::uint32_t(p0 + p)) * 8u : kOutOfRange);
}

// Per-component fields packed in one lane value (lanes = the item's
// components): the end bit E of the component's code in every lane's stream
// of a side block (the inclusive prefix sum of the widths) and its width b,
// as 64 E + b.  Lanes past the item's components hold a dummy (E = 32,
// b = 1: the top bit of a lane's first word) whose input is zero, so loops
// over whole groups of components need no tests.  `total` receives the bits
// per lane of the side block.
__device__ __forceinline__ std::uint32_t side_meta(bool comp, std::uint32_t bits, int lane,
                                                   std::uint32_t& total) {
    const std::uint32_t E = wave_scan(comp ? bits : 0u, lane);
    total = __builtin_amdgcn_readlane(E, 63);
    return comp ? E * 64u + bits : 32u * 64u + 1u;
}
// Words of a side block of `total` bits per lane.
__device__ __forceinline__ std::uint32_t side_words(std::uint32_t total) {
    return 64u * ((total + 31u) >> 5);
}

// ---------------------------------------------------------------------------
// Narrow passes (1, 2 or 4 right-hand sides per pass): vector ALUs.  Lanes are the item's rows
// (left factors, diagonal columns) or columns (right factors, summed over the
// lanes by a butterfly of lane swaps).

// Lane j's value, broadcast to the wave (v_readlane_b32; j is wave-uniform).
__device__ __forceinline__ float lane_value(float v, std::uint32_t j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double lane_value(double v, std::uint32_t j) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(unsigned(b)), j));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(unsigned(b >> 32)), j));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Cross-lane moves without LDS: DPP within rows of 16 lanes, and the gfx950
// lane-half swaps (v_permlane32_swap / v_permlane16_swap).
template <int CTRL> __device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp_move(double v) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned lo = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b)), CTRL, 0xF, 0xF, false));
    const unsigned hi = unsigned(__builtin_amdgcn_update_dpp(0, int(unsigned(b >> 32)), CTRL, 0xF, 0xF, false));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}
// (a, b) -> (a, b) with lanes 32-63 of a exchanged for lanes 0-31 of b.
__device__ __forceinline__ void swap_halves32(float& a, float& b) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap_halves32(double& a, double& b) {
    const unsigned long long x = static_cast<unsigned long long>(__double_as_longlong(a));
    const unsigned long long y = static_cast<unsigned long long>(__double_as_longlong(b));
    const auto lo = __builtin_amdgcn_permlane32_swap(unsigned(x), unsigned(y), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(unsigned(x >> 32), unsigned(y >> 32), false, false);
    a = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi[0]) << 32) | lo[0]));
    b = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi[1]) << 32) | lo[1]));
}
// (a, b) -> (a, b) with the odd 16-lane rows of a exchanged for the even rows of b.
__device__ __forceinline__ void swap_rows16(float& a, float& b) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b), false, false);
    a = __uint_as_float(r[0]);
    b = __uint_as_float(r[1]);
}
__device__ __forceinline__ void swap_rows16(double& a, double& b) {
    const unsigned long long x = static_cast<unsigned long long>(__double_as_longlong(a));
    const unsigned long long y = static_cast<unsigned long long>(__double_as_longlong(b));
    const auto lo = __builtin_amdgcn_permlane16_swap(unsigned(x), unsigned(y), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(unsigned(x >> 32), unsigned(y >> 32), false, false);
    a = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi[0]) << 32) | lo[0]));
    b = __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi[1]) << 32) | lo[1]));
}

constexpr int kDppRowRor4 = 0x124;  // lane i <- lane (i - 4) mod 16 within its row of 16
constexpr int kDppRowRor8 = 0x128;  // lane i <- lane (i - 8) mod 16 within its row of 16

// Sum of a lane's value over the lanes L, L + N, L + 2N, ... (N = 16, 8 or
// 4 lanes apart, 64 / N of them): rotations within each row of 16 lanes
// (DPP) for N < 16, then the 16-lane rows (v_permlane16_swap) and the two
// halves (v_permlane32_swap).  Every lane of a set ends with the same total,
// formed by the same tree.
template <class T> __device__ __forceinline__ T fold_rows16(T v) {
    T a = v, b = v;
    swap_rows16(a, b);  // a: rows 0,0,2,2; b: rows 1,1,3,3
    return a + b;
}
template <class T> __device__ __forceinline__ T fold_halves32(T v) {
    T a = v, b = v;
    swap_halves32(a, b);  // a: the low half twice; b: the high half twice
    return a + b;
}
template <int N, class T> __device__ __forceinline__ T group_sum(T v) {
    static_assert(N == 16 || N == 8 || N == 4, "groups 16, 8 or 4 lanes apart");
    if constexpr (N == 4) v += dpp_move<kDppRowRor4>(v);
    if constexpr (N <= 8) v += dpp_move<kDppRowRor8>(v);
    return fold_halves32(fold_rows16(v));
}

// Row contributions of four left-factor components (lanes are rows); zs[u]
// is component u's scaled input, zero for padding components.  In FP32 mode
// the four take the split path when a real one of them is wider than 24
// bits (groups of four start at multiples of four, whatever the
// right-hand-side count).
template <int KB, bool FP32>
__device__ __forceinline__ void valu_left_four(const std::uint32_t* stage, const std::uint32_t (&meta)[4],
                                          bool wide, int lane, const real_t<FP32> (&zs)[4][KB],
                                          double (&acc)[KB], float (&blk)[KB]) {
    int code[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) code[u] = decode_lm(stage, meta[u], lane);
    if (!FP32) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double cd = double(code[u]);
#pragma unroll
            for (int p = 0; p < KB; ++p) acc[p] = fma(cd, double(zs[u][p]), acc[p]);
        }
    } else if (!wide) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float cf = float(code[u]);
#pragma unroll
            for (int p = 0; p < KB; ++p) blk[p] = fmaf(cf, float(zs[u][p]), blk[p]);
        }
    } else {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float h, l;
            split_code(code[u], h, l);
#pragma unroll
            for (int p = 0; p < KB; ++p) {
                blk[p] = fmaf(h, float(zs[u][p]) * 65536.0f, blk[p]);
                blk[p] = fmaf(l, float(zs[u][p]), blk[p]);
            }
        }
    }
}

// Left-factor rows of components [j0, jend) of an item, four at a time
// (j0 a multiple of 4 and jend <= 64, so every lane index stays below 64):
// meta held one per lane (lanes = components), zs_of(j, u, p) the scaled input
// of component j + u for right-hand side p (lane-uniform), zero for
// components past jend, whose dummy meta decodes in-bounds 1-bit codes;
// bit c of wide_mask set for a component c wider than 24 bits (FP32 mode).
template <int KB, bool FP32, class ZsOf>
__device__ __forceinline__ void valu_left_rows(const std::uint32_t* stage, std::uint32_t lmeta_l,
                                          std::uint64_t wide_mask, std::uint32_t j0,
                                          std::uint32_t jend, int lane, ZsOf zs_of,
                                          double (&acc)[KB], float (&blk)[KB]) {
    using Real = real_t<FP32>;
    for (std::uint32_t j = j0; j < jend; j += 4) {
        std::uint32_t meta[4];
        Real zs[4][KB];
        const bool wide = ((wide_mask >> j) & 0xFull) != 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            meta[u] = __builtin_amdgcn_readlane(lmeta_l, j + u);
#pragma unroll
            for (int p = 0; p < KB; ++p) zs[u][p] = zs_of(j, u, p);
        }
        valu_left_four<KB, FP32>(stage, meta, wide, lane, zs, acc, blk);
    }
}

// Right-factor sums of a block of N components [j0, j0 + N) of an item
// (N = 16, 8 or 4; fewer real ones at the end: the rest decode the dummy
// field and are never used).  Lane L works on component j0 + L mod N and on
// the N columns starting at N (L / N): it sums code x x over them in order
// (x rows from LDS, zero past the item's columns), and group_sum adds the
// 64 / N column groups.  tot[p] in every lane of component j is its sum for
// right-hand side p.  In FP32 mode, when a real component of the block is
// wider than 24 bits, every code of the block is split (split_code) and the
// halves are summed apart and joined in FP64.  The blocks depend only on
// the item's component count, so a component's rounding is the same for
// every right-hand-side count.
template <int N, int KB, bool FP32>
__device__ __forceinline__ void valu_right_block(const std::uint32_t* stage,
                                                 const real_t<FP32>* xs, std::uint32_t rmeta_l,
                                                 std::uint64_t wide_mask, std::uint32_t j0,
                                                 int lane, double (&tot)[KB]) {
    using Real = real_t<FP32>;
    const int c0 = (lane / N) * N;
    const std::uint32_t meta =
        std::uint32_t(__shfl(int(rmeta_l), int(min(j0 + std::uint32_t(lane % N), 63u)), 64));
    const bool wide = FP32 && ((wide_mask >> j0) & ((1ull << N) - 1ull)) != 0;
    if (!wide) {
        Real sum[KB];
#pragma unroll
        for (int p = 0; p < KB; ++p) sum[p] = Real(0);
#pragma unroll 1
        for (int t = 0; t < N / 4; ++t) {
            int code[4];
            decode_right4(stage, meta, c0 + 4 * t, code);
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int p = 0; p < KB; ++p)
                    sum[p] = fma(Real(code[u]), xs[(c0 + 4 * t + u) * KB + p], sum[p]);
        }
#pragma unroll
        for (int p = 0; p < KB; ++p) tot[p] = double(group_sum<N>(sum[p]));
        return;
    }
    float sh[KB], sl[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) sh[p] = sl[p] = 0.0f;
#pragma unroll 1
    for (int t = 0; t < N / 4; ++t) {
        int code[4];
        decode_right4(stage, meta, c0 + 4 * t, code);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            float h, l;
            split_code(code[u], h, l);
#pragma unroll
            for (int p = 0; p < KB; ++p) {
                const float xv = float(xs[(c0 + 4 * t + u) * KB + p]);  // FP32 mode: already FP32
                sh[p] = fmaf(h, xv, sh[p]);
                sl[p] = fmaf(l, xv, sl[p]);
            }
        }
    }
#pragma unroll
    for (int p = 0; p < KB; ++p)
        tot[p] = double(group_sum<N>(sh[p])) * 65536.0 + double(group_sum<N>(sl[p]));
}

// One block of N components of a kRight or kFused item: right-factor sums
// from the right side block at `stage`, then (kRight) their workspace rows
// or (kFused) the scaled inputs and the left side block at `lstage` (the
// same components, so the same packed fields).
template <int N, int KB, bool FP32>
__device__ __forceinline__ void valu_right_then_left(const MemberView& M, const Item& f,
                                                const std::uint32_t* stage, std::uint32_t rmeta_l,
                                                const std::uint32_t* lstage, std::uint64_t wide_mask,
                                                double scale_l, std::uint32_t j0, int lane, int p0,
                                                int kk, const real_t<FP32>* xs,
                                                double (&acc)[KB], float (&blk)[KB]) {
    using Real = real_t<FP32>;
    double tot[KB];
    valu_right_block<N, KB, FP32>(stage, xs, rmeta_l, wide_mask, j0, lane, tot);
    const std::uint32_t jl = j0 + std::uint32_t(lane % N);
    if (f.kind == kRight) {
        if (lane < N && jl < f.count) {
            double* out = M.zp + (std::uint64_t(f.zrow) + jl) * std::uint64_t(M.k) + p0;
#pragma unroll
            for (int p = 0; p < KB; ++p)
                if (p < kk) out[p] = tot[p];
        }
        return;
    }
    const double sc = __shfl(scale_l, int(min(jl, 63u)), 64);
    Real zs_lane[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) zs_lane[p] = jl < f.count ? Real(sc * tot[p]) : Real(0);
    valu_left_rows<KB, FP32>(lstage, rmeta_l, wide_mask, j0, min(j0 + N, f.count), lane,
                        [&](std::uint32_t j, int u, int p) {
                            return lane_value(zs_lane[p], j + u - j0);
                        },
                        acc, blk);
}


// Component widths and scale products of the staged item (lanes =
// components), read back from LDS after the staging.
struct Staged {
    std::uint32_t bits_l;
    double scale_l;
};

// Moves a fetched item into the wave's LDS: its data, its x rows (lanes =
// columns; zero past the item's columns; in FP32 mode a kFused or kRight
// item's rows go straight to the FP32 table) or kLeft partial sums (lanes =
// pairs; zero past them), and its widths and scale products.  Every
// register of the fetch buffer is dead afterwards, so the next fetch into it
// needs no copy of a load still in flight.
template <int KB, bool FP32>
__device__ __forceinline__ void stage_item(const Item& f, const Fetch<KB>& F, int lane,
                                           std::uint8_t* wave_lds) {
    quad_t* stage_q = reinterpret_cast<quad_t*>(wave_lds + kLdsHead);
    double* xs = reinterpret_cast<double*>(wave_lds + kLdsHead + kStageBytes);
#pragma unroll
    for (int t = 0; t < kItemQuads / 64; ++t)
        if (std::uint32_t(lane + 64 * t) < f.quads) stage_q[lane + 64 * t] = F.q[t];
    if (FP32 && (f.kind == kFused || f.kind == kRight)) {
        float* xf = reinterpret_cast<float*>(wave_lds + wave_lds_xf<KB>());
#pragma unroll
        for (int p = 0; p < KB; ++p) xf[lane * KB + p] = float(F.in[p]);
    } else {
#pragma unroll
        for (int p = 0; p < KB; ++p) xs[lane * KB + p] = F.in[p];
    }
    reinterpret_cast<double*>(wave_lds + wave_lds_meta<KB>())[lane] = F.scale;
    reinterpret_cast<std::uint32_t*>(wave_lds + wave_lds_meta<KB>() + 512)[lane] = F.bits;
}
template <int KB>
__device__ __forceinline__ Staged staged_meta(const std::uint8_t* wave_lds, int lane) {
    return Staged{reinterpret_cast<const std::uint32_t*>(wave_lds + wave_lds_meta<KB>() + 512)[lane],
                  reinterpret_cast<const double*>(wave_lds + wave_lds_meta<KB>())[lane]};
}

// Narrow passes: one staged item on the vector ALUs.  Row contributions are
// added to acc (lanes = the R rows of the tile), in FP32 per item for the
// codes in FP32 mode.
template <int KB, bool FP32, int PASS>
__device__ __forceinline__ void valu_item(const MemberView& M, const Item& f, const Staged& st,
                                          int lane, int p0, int kk, std::uint32_t R,
                                          std::uint8_t* wave_lds, double (&acc)[KB]) {
    using Real = real_t<FP32>;
    const std::uint32_t* stage = reinterpret_cast<const std::uint32_t*>(wave_lds + kLdsHead);
    const double* xs = reinterpret_cast<const double*>(wave_lds + kLdsHead + kStageBytes);
    const bool comp = std::uint32_t(lane) < f.count;
    const std::uint32_t bits_l = st.bits_l;
    const double scale_l = st.scale_l;
    // components wider than 24 bits (split in FP32 mode), one bit each
    const std::uint64_t wide_mask = FP32 ? __ballot(comp && bits_l > 24u) : 0ull;
    float blk[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) blk[p] = 0.0f;
    if (PASS == 1 && f.kind == kDense) {
        // lanes = rows; column c of the item at stage[2 (c R + row)]
        const double* dv = reinterpret_cast<const double*>(stage);
        const std::uint32_t cnt = min(f.count, std::uint32_t(kDenseColumns));
#pragma unroll
        for (int c = 0; c < kDenseColumns; ++c) {
            if (std::uint32_t(c) < cnt) {
                const double v = dv[c * R + lane];
#pragma unroll
                for (int p = 0; p < KB; ++p) acc[p] = fma(v, xs[c * KB + p], acc[p]);
            }
        }
        return;
    }
    if (PASS == 2) {
        // kLeft. z of each component: its column tiles' partial sums in order
        double z[KB];
#pragma unroll
        for (int p = 0; p < KB; ++p) z[p] = 0.0;
        const std::uint32_t nseg = f.nseg;
        if (comp && (lane + 1) * nseg <= 64u) {
            for (std::uint32_t sg = 0; sg < nseg; ++sg) {
#pragma unroll
                for (int p = 0; p < KB; ++p) z[p] += xs[(lane * nseg + sg) * KB + p];
            }
        }
        Real zs_l[KB];
#pragma unroll
        for (int p = 0; p < KB; ++p) zs_l[p] = Real(scale_l * z[p]);
        std::uint32_t total;
        const std::uint32_t lmeta_l = side_meta(comp, bits_l, lane, total);
        valu_left_rows<KB, FP32>(stage, lmeta_l, wide_mask, 0, f.count, lane,
                                 [&](std::uint32_t j, int u, int p) { return lane_value(zs_l[p], j + u); },
                                 acc, blk);
    } else {
        // kFused / kRight.  One packed field per component for both side
        // blocks (the left block of a kFused item follows the right one).
        std::uint32_t total;
        const std::uint32_t rmeta_l = side_meta(comp, bits_l, lane, total);
        const std::uint32_t* lstage = stage + side_words(total);
        // x rows in the arithmetic precision (FP32 mode: converted at staging)
        const Real* xr;
        if constexpr (FP32)
            xr = reinterpret_cast<const float*>(wave_lds + wave_lds_xf<KB>());
        else
            xr = xs;
        // In FP32 mode the components up to the last one wider than 24 bits
        // (the widest come first) go in blocks of 4, each deciding the split
        // for itself; then blocks of 16 while more than 8 remain, then one of
        // 8 or 4.
        std::uint32_t j0 = 0;
        const std::uint32_t wide_end = wide_mask ? 64u - std::uint32_t(__builtin_clzll(wide_mask)) : 0u;
        for (; j0 < wide_end; j0 += 4)
            valu_right_then_left<4, KB, FP32>(M, f, stage, rmeta_l, lstage, wide_mask, scale_l, j0,
                                              lane, p0, kk, xr, acc, blk);
        for (; j0 + 8 < f.count; j0 += 16)
            valu_right_then_left<16, KB, FP32>(M, f, stage, rmeta_l, lstage, wide_mask, scale_l, j0,
                                               lane, p0, kk, xr, acc, blk);
        if (j0 + 4 < f.count) {
            valu_right_then_left<8, KB, FP32>(M, f, stage, rmeta_l, lstage, wide_mask, scale_l, j0,
                                              lane, p0, kk, xr, acc, blk);
            j0 += 8;
        }
        if (j0 < f.count)
            valu_right_then_left<4, KB, FP32>(M, f, stage, rmeta_l, lstage, wide_mask, scale_l, j0,
                                              lane, p0, kk, xr, acc, blk);
    }
    if (FP32) {
#pragma unroll
        for (int p = 0; p < KB; ++p) acc[p] += double(blk[p]);
    }
}

// Decodes and applies one staged item on the vector ALUs.  Pass 1 holds
// kDense, kFused and kRight items, pass 2 kLeft items only.
template <int KB, bool FP32, int PASS>
__device__ __forceinline__ void compute_item(const MemberView& M, const Item& f,
                                             const Staged& st, int lane, int p0, int kk,
                                             std::uint32_t R, std::uint8_t* wave_lds,
                                             double (&acc)[KB]) {
    valu_item<KB, FP32, PASS>(M, f, st, lane, p0, kk, R, wave_lds, acc);
}

// One wave's items: [b1, e1) followed by [b2, e2), as its head record gives
// them.  Row contributions (kDense, kFused, kLeft) are added to acc (lanes =
// the R rows of the tile); kRight items write their partial sums to the
// workspace.  Two items are in flight while one is decoded: the fetch
// buffers alternate.  The first two items' records come with the head, so
// their fetches wait for one load, not for the run's records.  Each step
// sets the wave's issue priority by its progress through the run.
template <int KB, bool FP32, int PASS>
__device__ __forceinline__ void run_items(const MemberView& M, const WaveHead* head, int p0,
                                          int kk, std::uint32_t R,
                                          std::uint8_t* __restrict__ wave_lds, double (&acc)[KB]) {
    const int lane = threadIdx.x & 63;
    // the head as wave-uniform dwords (scalar registers whatever the
    // compiler can prove about the pointer)
    WaveHead h;
    {
        const std::uint32_t* src = reinterpret_cast<const std::uint32_t*>(head);
        std::uint32_t* dst = reinterpret_cast<std::uint32_t*>(&h);
#pragma unroll
        for (int i = 0; i < int(sizeof(WaveHead) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    }
    const std::uint32_t b1 = h.b1, e1 = h.e1, b2 = h.b2, e2 = h.e2;
    const std::uint32_t n1 = e1 - b1, total = n1 + (e2 - b2);
    for (std::uint32_t g0 = 0; g0 < total; g0 += 64) {
        const std::uint32_t n = min(64u, total - g0);
        // item records, lanes = items of this group
        const std::uint32_t i = g0 + lane;
        const std::uint32_t idx = i < n1 ? b1 + i : b2 + (i - n1);
        const __amdgpu_buffer_rsrc_t recs = buffer(M.items, kOutOfRange);
        const std::uint32_t roff = std::uint32_t(lane) < n ? idx * 32u : kOutOfRange;
        const quad_t r0 = __builtin_amdgcn_raw_buffer_load_b128(recs, roff, 0, 0);
        const quad_t r1 = __builtin_amdgcn_raw_buffer_load_b128(recs, roff + 16u, 0, 0);

        // Every step issues the same loads: past the end of the group it
        // fetches an empty item (all loads out of range, reading zero), so
        // the loads in flight are counted exactly and never waited for early.
        const Item none = {};
        auto slot = [&](std::uint32_t i) { return i < n ? item_at(r0, r1, i) : none; };
        Item fa = g0 == 0 ? item_from(h.first[0]) : slot(0);
        Item fb = g0 == 0 ? item_from(h.first[1]) : slot(1);
        Fetch<KB> FA, FB;
        fetch<KB>(M, fa, lane, p0, kk, FA);
        fetch<KB>(M, fb, lane, p0, kk, FB);
        // the records of items 2 and 3 are read only now, so that the first
        // two fetches do not wait for the group's record loads
        Item fc = slot(2), fd = slot(3);

        // Steady state: buffer A holds item s, buffer B item s+1.  Step s
        // waits for item s (fetched two steps earlier) and fetches item s+2
        // into the freed buffer, so two items' data stay in flight during a
        // decode.
        for (std::uint32_t s = 0; s < n; s += 2) {
            {
                progress_priority(g0 + s, total);
                stage_item<KB, FP32>(fa, FA, lane, wave_lds);
                fetch<KB>(M, fc, lane, p0, kk, FA);
                wave_lds_sync();
                compute_item<KB, FP32, PASS>(M, fa, staged_meta<KB>(wave_lds, lane), lane, p0, kk,
                                             R, wave_lds, acc);
                wave_lds_sync();  // the next item overwrites the stage
                fa = fc;
                fc = slot(s + 4);
            }
            // No exit here: an odd count runs one empty item, whose loads all
            // read zero, so that both steps always run and the waits for
            // loads in flight stay exact across the loop.
            {
                progress_priority(g0 + s + 1, total);
                stage_item<KB, FP32>(fb, FB, lane, wave_lds);
                fetch<KB>(M, fd, lane, p0, kk, FB);
                wave_lds_sync();
                compute_item<KB, FP32, PASS>(M, fb, staged_meta<KB>(wave_lds, lane), lane, p0, kk,
                                             R, wave_lds, acc);
                wave_lds_sync();
                fb = fd;
                fd = slot(s + 5);
            }
        }
    }
}

// Completes the tree over the W wave totals in wave 0 (two right-hand sides
// per round through LDS), adds it to `base` and hands the result to `store`.
template <int KB, int W, class Store>
__device__ __forceinline__ void combine_rows(const double (&total)[KB], const double (&base)[KB],
                                             std::uint8_t* lds, int wave, int lane, Store store) {
    double (*partial)[2][64] = reinterpret_cast<double (*)[2][64]>(lds);
#pragma unroll
    for (int pc = 0; pc < KB; pc += 2) {
        __syncthreads();
        if (wave > 0) {
            partial[wave - 1][0][lane] = total[pc];
            partial[wave - 1][1][lane] = pc + 1 < KB ? total[pc + 1] : 0.0;
        }
        __syncthreads();
        if (wave == 0) {
            double v0[W], v1[W];
            v0[0] = total[pc];
            v1[0] = pc + 1 < KB ? total[pc + 1] : 0.0;
#pragma unroll
            for (int w = 1; w < W; ++w) {
                v0[w] = partial[w - 1][0][lane];
                v1[w] = partial[w - 1][1][lane];
            }
#pragma unroll
            for (int width = 1; width < W; width *= 2) {
#pragma unroll
                for (int i = 0; i + width < W; i += 2 * width) {
                    v0[i] += v0[i + width];
                    v1[i] += v1[i + width];
                }
            }
            store(pc, base[pc] + v0[0], pc + 1 < KB ? base[pc + 1] + v1[0] : 0.0);
        }
    }
}

// The row sums of a tile: pass 1's 16 item runs (each summed in item order
// from zero) go to four workgroups of four waves; workgroup q sums its waves
// as 0 + ((r0 + r1) + (r2 + r3)) into set q of the workspace's rows, and pass
// 2 (tile_pass2) forms
//   y = ((g0 + g1) + (g2 + g3)) + ((h0 + h1) + (h2 + h3))
// over these sets g and its own four groups of waves h: one fixed tree over
// the 32 runs, the same for every right-hand-side count and launch shape.
constexpr int kGroupWaves = kWaves / kPass1Groups;

template <int KB, bool FP32>
__device__ __forceinline__ void pass1_rows(const MemberView& M, std::uint32_t group, int p0,
                                           std::uint8_t* lds) {
    constexpr int W = kGroupWaves;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int kk = min(KB, M.k - p0);
    const std::uint32_t t = group / kPass1Groups, q = group % kPass1Groups;
    const TileDev* __restrict__ tile = M.tiles + t;
    const std::uint32_t R = min(tile->rows, std::uint32_t(kTile));
    const int r = int(q) * W + wave;
    double total[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) total[p] = 0.0;
    run_items<KB, FP32, 1>(M, M.heads + t * kWaves + std::uint32_t(r), p0, kk, R,
                           lds + wave * wave_lds_bytes<KB>(), total);
    double zero[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) zero[p] = 0.0;
    double* rows = M.zp + (M.z_rows + q * M.order + tile->row0 + lane) * std::uint64_t(M.k) + p0;
    combine_rows<KB, W>(total, zero, lds, wave, lane, [&](int pc, double a0, double a1) {
        if (lane < int(R)) {
            if (pc < kk) rows[pc] = a0;
            if (pc + 1 < kk) rows[pc + 1] = a1;
        }
    });
}

// Pass-1 waves per SIMD the register allocation must allow, by right-hand
// sides per pass.
constexpr int pass1_occupancy(int kb) { return kb <= 1 ? 4 : 2; }
// The batch kernels carry the member table besides.
constexpr int batch1_occupancy(int kb) { return kb <= 1 ? 3 : 2; }

template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kGroupWaves, pass1_occupancy(KB)) void apply_pass1(
    MemberView M) {
    __shared__ __align__(16) std::uint8_t lds[kGroupWaves * wave_lds_bytes<KB>()];
    pass1_rows<KB, FP32>(M, blockIdx.x, blockIdx.y * KB, lds);
}

// Pass 2 of tile t as one workgroup of kWaves waves (wave w runs the tile's
// pass-2 item run w), which then writes y for its rows and right-hand sides
// [p0, p0 + KB) itself, as the fixed tree over the tile's 32 runs
//   y = ((g0 + g1) + (g2 + g3)) + ((h0 + h1) + (h2 + h3)),
// with g the four pass-1 sets from the workspace and h_q = 0 + ((r_4q +
// r_4q+1) + (r_4q+2 + r_4q+3)) formed from the wave totals r in LDS.  The
// pass-2 sets never leave the workgroup, and no launch follows.
template <int KB, bool FP32>
__device__ __forceinline__ void tile_pass2(const MemberView& M, std::uint32_t t, int p0,
                                           std::uint8_t* lds) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int kk = min(KB, M.k - p0);
    const TileDev* __restrict__ tile = M.tiles + t;
    const std::uint32_t R = min(tile->rows, std::uint32_t(kTile));
    double total[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) total[p] = 0.0;
    run_items<KB, FP32, 2>(M, M.heads + (M.tiles_count + t) * kWaves + std::uint32_t(wave), p0, kk,
                           R, lds + wave * wave_lds_bytes<KB>(), total);
    __syncthreads();  // every wave is done with its stage
    double* part = reinterpret_cast<double*>(lds);  // [wave][p][row]
#pragma unroll
    for (int p = 0; p < KB; ++p) part[(wave * KB + p) * 64 + lane] = total[p];
    __syncthreads();
    const std::uint32_t e = threadIdx.x;
    if (e < R * std::uint32_t(kk)) {
        const std::uint32_t i = e / std::uint32_t(kk), p = e - i * std::uint32_t(kk);
        const std::uint64_t k = std::uint64_t(M.k);
        const double* sets = M.zp + (M.z_rows + tile->row0 + i) * k + std::uint64_t(p0) + p;
        double g[kPass1Groups], h[kPass1Groups];
#pragma unroll
        for (int q = 0; q < kPass1Groups; ++q) {
            g[q] = sets[std::uint64_t(q) * M.order * k];
            const double* v = part + (4 * q * KB + int(p)) * 64 + int(i);
            h[q] = 0.0 + ((v[0] + v[KB * 64]) + (v[2 * KB * 64] + v[3 * KB * 64]));
        }
        M.y[std::int64_t(M.perm[tile->row0 + i]) * M.ldy + p0 + std::int64_t(p)] =
            ((g[0] + g[1]) + (g[2] + g[3])) + ((h[0] + h[1]) + (h[2] + h[3]));
    }
}
static_assert(kWaves == 4 * kPass1Groups, "pass 2 tree: four waves per pass-1 set");

template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kWaves, 4) void apply_pass2(MemberView M) {
    __shared__ __align__(16) std::uint8_t lds[kWaves * wave_lds_bytes<KB>()];
    tile_pass2<KB, FP32>(M, blockIdx.x, blockIdx.y * KB, lds);
}


// x rows of tile t's indices in clustered order (the tiles cover every index
// once), one thread per (row, right-hand side).
__device__ __forceinline__ void permute_tile(const MemberView& M, std::uint32_t t) {
    const TileDev* __restrict__ tile = M.tiles + t;
    const std::uint32_t R = min(tile->rows, std::uint32_t(kTile));
    const std::uint32_t k = std::uint32_t(M.k);
    double* xc = clustered_x(M) + std::uint64_t(tile->row0) * k;
    for (std::uint32_t e = threadIdx.x; e < R * k; e += blockDim.x) {
        const std::uint32_t i = e / k, p = e - i * k;
        xc[e] = M.x[std::int64_t(M.perm[tile->row0 + i]) * M.ldx + p];
    }
}
__global__ __launch_bounds__(256) void apply_permute(MemberView M) { permute_tile(M, blockIdx.x); }

// The member owning global work index `index`: first[m] <= index < first[m+1].
__device__ __forceinline__ std::uint32_t locate_member(const std::uint32_t* __restrict__ first,
                                                       std::uint32_t count, std::uint32_t index) {
    std::uint32_t lo = 0, hi = count;  // first[lo] <= index < first[hi]
    while (hi - lo > 1) {
        const std::uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= index) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Batched: the work of every member of one launch, x / y / workspace given
// as offsets from the call's base pointers.
__device__ __forceinline__ MemberView bind_member(const MemberView* __restrict__ members,
                                                  std::uint32_t m, const double* x, double* y,
                                                  double* zp) {
    MemberView M = members[m];
    M.x = x + reinterpret_cast<std::uintptr_t>(M.x);
    M.y = y + reinterpret_cast<std::uintptr_t>(M.y);
    M.zp = zp + reinterpret_cast<std::uintptr_t>(M.zp);
    return M;
}

template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kGroupWaves, batch1_occupancy(KB)) void batch_pass1(
    const MemberView* __restrict__ members, const std::uint32_t* __restrict__ first,
    std::uint32_t count, const double* x, double* y, double* zp) {
    __shared__ __align__(16) std::uint8_t lds[kGroupWaves * wave_lds_bytes<KB>()];
    const std::uint32_t g = blockIdx.x / kPass1Groups;
    const std::uint32_t m = locate_member(first, count, g);
    const MemberView M = bind_member(members, m, x, y, zp);
    const int p0 = blockIdx.y * KB;
    if (p0 >= M.k) return;  // the whole workgroup leaves together
    pass1_rows<KB, FP32>(M, blockIdx.x - first[m] * kPass1Groups, p0, lds);
}

__global__ __launch_bounds__(256) void batch_permute(const MemberView* __restrict__ members,
                                                    const std::uint32_t* __restrict__ first,
                                                    std::uint32_t count, const double* x,
                                                    double* y, double* zp) {
    const std::uint32_t m = locate_member(first, count, blockIdx.x);
    const MemberView M = bind_member(members, m, x, y, zp);
    permute_tile(M, blockIdx.x - first[m]);
}

// Batched pass 2: one workgroup of kWaves waves per (member tile, slice),
// writing y as tile_pass2 does.
template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kWaves, 4) void batch_pass2(
    const MemberView* __restrict__ members, const std::uint32_t* __restrict__ first,
    std::uint32_t count, const double* x, double* y, double* zp) {
    __shared__ __align__(16) std::uint8_t lds[kWaves * wave_lds_bytes<KB>()];
    const std::uint32_t m = locate_member(first, count, blockIdx.x);
    const MemberView M = bind_member(members, m, x, y, zp);
    const int p0 = blockIdx.y * KB;
    if (p0 >= M.k) return;  // the whole workgroup leaves together
    tile_pass2<KB, FP32>(M, blockIdx.x - first[m], p0, lds);
}

// Test hook: codes of one factor side of one block, from its chunks, written
// component-major in the block's own row order.
// Reconstruction, one workgroup per (block, row tile, column tile): every
// entry formed as the reference forms it (core.cpp:392-410), in FP64 without
// contraction: the diagonal block's value, or zero plus, component by
// component, (code_l lscale) (code_r rscale).  The factors of 32 components
// at a time are decoded into LDS; thread 4i + q holds row i and columns
// q, q + 4, ... of the tile and writes them through the permutation.
constexpr int kRecComps = 32;
__global__ __launch_bounds__(256) void reconstruct_tiles(
    const RecTask* __restrict__ tasks, const ChunkRef* __restrict__ chunks,
    const std::uint32_t* __restrict__ words, const std::int32_t* __restrict__ widths,
    const double* __restrict__ lscale, const double* __restrict__ rscale,
    const std::int32_t* __restrict__ perm, double* __restrict__ out, std::int64_t ld) {
    __shared__ double L[kTile][kRecComps + 1], Rt[kTile][kRecComps + 1];
    __shared__ std::uint64_t off[2][kRecComps];
    __shared__ std::uint32_t obit[2][kRecComps];
    const RecTask t = tasks[blockIdx.x];
    const int tid = int(threadIdx.x);
    const int i = tid >> 2, kq = tid & 3;
    const int rows = int(min(t.rows, std::uint32_t(kTile))), cols = int(min(t.cols, std::uint32_t(kTile)));
    double v[kTile / 4];
#pragma unroll
    for (int e = 0; e < kTile / 4; ++e) v[e] = 0.0;
    if (t.kind == 1) {
        // the block's row tile: items of 8 columns, each R x 8 column-major
        const double* dv = reinterpret_cast<const double*>(words + 4ull * t.quad);
#pragma unroll
        for (int e = 0; e < kTile / 4; ++e) {
            const int k = kq + 4 * e;
            if (i < rows && k < cols) {
                const std::uint32_t cb = t.col0 + std::uint32_t(k) - t.bcol0;
                v[e] = dv[std::uint64_t(cb / kDenseColumns) * kDenseColumns * std::uint64_t(rows) +
                          (cb % kDenseColumns) * std::uint64_t(rows) + std::uint64_t(i)];
            }
        }
    } else if (t.kind == 2) {
        for (std::uint32_t w0 = 0; w0 < t.rank; w0 += kRecComps) {
            const std::uint32_t wn = min(std::uint32_t(kRecComps), t.rank - w0);
            __syncthreads();
            if (tid < 2 * kRecComps) {
                // side block holding component w0 + jj in this tile's chunks,
                // and the component's bit offset in its lanes
                const int side = tid / kRecComps, jj = tid % kRecComps;
                const std::uint32_t j = w0 + std::uint32_t(jj);
                const ChunkRef* c = chunks + (side ? t.rchunk : t.lchunk);
                const std::uint32_t n = side ? t.nr : t.nl;
                std::uint64_t w = 0;
                std::uint32_t o = 0;
                for (std::uint32_t q = 0; q < n; ++q) {
                    if (j >= c[q].j0 && j < c[q].j0 + c[q].jn) {
                        w = c[q].word;
                        o = 0;
                        for (std::uint32_t jp = c[q].j0; jp < j; ++jp)
                            o += std::uint32_t(widths[t.comp_base + jp]);
                    }
                }
                off[side][jj] = w;
                obit[side][jj] = o;
            }
            __syncthreads();
            for (int p = tid; p < kTile * kRecComps; p += 256) {
                const int row = p / kRecComps, jj = p % kRecComps;
                if (std::uint32_t(jj) < wn) {
                    const std::uint32_t c = t.comp_base + w0 + std::uint32_t(jj);
                    const std::uint32_t bits = std::uint32_t(widths[c]);
                    if (row < rows)
                        L[row][jj] = __dmul_rn(double(decode_side_code(words + off[0][jj], std::uint32_t(row),
                                                                       obit[0][jj], bits, false)),
                                               lscale[c]);
                    if (row < cols)
                        Rt[row][jj] = __dmul_rn(double(decode_side_code(words + off[1][jj], std::uint32_t(row),
                                                                        obit[1][jj], bits, true)),
                                                rscale[c]);
                }
            }
            __syncthreads();
            if (i < rows) {
                for (std::uint32_t jj = 0; jj < wn; ++jj) {
                    const double l = L[i][jj];
#pragma unroll
                    for (int e = 0; e < kTile / 4; ++e) {
                        const int k = kq + 4 * e;
                        if (k < cols) v[e] = __dadd_rn(v[e], __dmul_rn(l, Rt[k][jj]));
                    }
                }
            }
        }
    }
    if (i < rows) {
        const std::int64_t orow = std::int64_t(perm[t.row0 + std::uint32_t(i)]) * ld;
#pragma unroll
        for (int e = 0; e < kTile / 4; ++e) {
            const int k = kq + 4 * e;
            if (k < cols) out[orow + perm[t.col0 + std::uint32_t(k)]] = v[e];
        }
    }
}

__global__ void decode_side(const std::uint32_t* __restrict__ words,
                            const ChunkRef* __restrict__ chunks,
                            const std::int32_t* __restrict__ widths, std::uint32_t rows,
                            int swz, std::int32_t* __restrict__ out) {
    const ChunkRef ch = chunks[blockIdx.x];
    std::uint32_t o = 0;
    for (std::uint32_t j = ch.j0; j < ch.j0 + ch.jn; ++j) {
        const std::uint32_t bits = widths[j];
        for (std::uint32_t i = threadIdx.x; i < ch.count; i += blockDim.x)
            out[std::uint64_t(j) * rows + ch.first + i] =
                decode_side_code(words + ch.word, i, o, bits, swz != 0);
        o += bits;
    }
}

// One chunk of a factor side per workgroup, decoded to FP32 code x scale:
// side 0 at out[offset + row rank + j], side 1 at out[offset + j rows + row].
__global__ __launch_bounds__(256) void decode_factors(const FactorTask* __restrict__ tasks,
                                                      const std::uint32_t* __restrict__ words,
                                                      const std::int32_t* __restrict__ widths,
                                                      const double* __restrict__ lscale,
                                                      const double* __restrict__ rscale,
                                                      float* __restrict__ out) {
    const FactorTask t = tasks[blockIdx.x];
    const double* scale = t.side ? rscale : lscale;
    std::uint32_t o = 0;
    for (std::uint32_t j = t.j0; j < t.j0 + t.jn; ++j) {
        const std::uint32_t bits = std::uint32_t(widths[t.comp_base + j]);
        const double sc = scale[t.comp_base + j];
        for (std::uint32_t i = threadIdx.x; i < t.count; i += blockDim.x) {
            const float v = float(double(decode_side_code(words + t.word, i, o, bits, t.side != 0)) * sc);
            const std::uint64_t row = t.first + i;
            out[t.offset + (t.side ? std::uint64_t(j) * t.rows + row : row * t.rank + j)] = v;
        }
        o += bits;
    }
}

// ---------------------------------------------------------------------------
// Host side

MemberView member_view(const DeviceStore& s, const double* x, std::int64_t ldx, int k, double* y,
                       std::int64_t ldy, double* zp) {
    MemberView M{};
    M.perm = s.perm;
    M.words = s.words;
    M.widths = s.widths;
    M.cscale = s.cscale;
    M.items = s.items;
    M.tiles = s.tiles;
    M.heads = s.heads;
    M.x = x;
    M.y = y;
    M.zp = zp;
    M.ldx = ldx;
    M.ldy = ldy;
    M.k = k;
    M.tiles_count = s.n_tiles;
    M.order = std::uint32_t(s.order);
    M.z_rows = s.z_rows;
    return M;
}

// Event slots of the next profiled apply, or null when not profiling.
hipEvent_t* profile_slot(const DeviceStore& s) {
    auto& p = s.profile;
    std::lock_guard<std::mutex> guard(p.lock);
    if (!p.armed || 3 * (p.used + 1) > p.events.size()) return nullptr;
    return reinterpret_cast<hipEvent_t*>(p.events.data() + 3 * p.used++);
}

// Right-hand sides per pass for an apply of k: the tuned choice when one
// was set for this k (qh_store_set_rhs_block), else 1, 2 or 4 on the vector
// ALUs for k <= 4 and 16, 32 or 64 on the matrix cores above.
int default_rhs_block(std::int64_t k) {
    return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : k <= 16 ? 16 : k <= 32 ? 32 : 64;
}
int rhs_block(const DeviceStore& s, std::int64_t k) {
    std::shared_lock<std::shared_mutex> guard(s.rhs_lock);
    const auto it = s.rhs_block.find(k);
    return it == s.rhs_block.end() ? default_rhs_block(k) : it->second;
}
bool is_wide(int kb) { return kb >= 16; }

WideView wide_view(const DeviceStore& s, const double* x, std::int64_t ldx, int k, double* y,
                   std::int64_t ldy, void* zw) {
    WideView V{};
    V.perm = s.perm;
    V.words = s.words;
    V.widths = s.widths;
    V.cscale = s.cscale;
    V.items = s.witems;
    V.rruns = s.wruns_r;
    V.lruns = s.wruns_l;
    V.sums = s.wsums;
    V.zsums = s.wzsums;
    V.ctiles = s.ctiles;
    V.tiles = s.tiles;
    V.x = x;
    V.y = y;
    V.zw = zw;
    V.ldx = ldx;
    V.ldy = ldy;
    V.k = k;
    V.z_rows = s.wz_rows;
    return V;
}

WideCounts wide_counts(const DeviceStore& s) {
    WideCounts c;
    c.right_groups = s.n_wr_groups;
    c.left_groups = s.n_wl_groups;
    c.sums = s.n_wsums;
    c.zsums = s.n_wzsums;
    return c;
}

// Workspace of the narrow passes (partial sums, eight row sets, x in
// clustered order) and of the wide passes (partial sums, row sets).
std::uint64_t narrow_workspace(const DeviceStore& s, std::int64_t k) {
    return (s.z_rows + (kRowSets + 1) * s.order) * std::uint64_t(k) * sizeof(double);
}
std::uint64_t wide_workspace(const DeviceStore& s, std::int64_t k) {
    return (s.wz_rows + std::uint64_t(s.wsets) * kTile) * std::uint64_t(k) * sizeof(double);
}

template <int KB, bool FP32>
void launch_apply(const DeviceStore& s, const double* x, std::int64_t ldx, int k, double* y,
                  std::int64_t ldy, double* zp, hipStream_t stream) {
    const unsigned passes = unsigned((k + KB - 1) / KB);
    const MemberView M = member_view(s, x, ldx, k, y, ldy, zp);
    hipEvent_t* ev = profile_slot(s);
    hipLaunchKernelGGL(apply_permute, dim3(s.n_tiles), dim3(256), 0, stream, M);
    hip_check(hipGetLastError(), "apply permute launch");
    if (ev) hip_check(hipEventRecord(ev[0], stream), "hipEventRecord");
    hipLaunchKernelGGL((apply_pass1<KB, FP32>), dim3(s.n_tiles * kPass1Groups, passes),
                       dim3(64 * kGroupWaves), 0, stream, M);
    hip_check(hipGetLastError(), "apply pass 1 launch");
    if (ev) hip_check(hipEventRecord(ev[1], stream), "hipEventRecord");
    hipLaunchKernelGGL((apply_pass2<KB, FP32>), dim3(s.n_tiles, passes), dim3(64 * kWaves), 0,
                       stream, M);
    hip_check(hipGetLastError(), "apply pass 2 launch");
    if (ev) hip_check(hipEventRecord(ev[2], stream), "hipEventRecord");
}

template <int KB, bool FP32>
void launch_batch_class(const BatchClass& c, const double* x, double* y, double* zp,
                        hipStream_t stream) {
    const unsigned passes = unsigned((c.max_k + KB - 1) / KB);
    const std::uint32_t count = std::uint32_t(c.member_count);
    hipLaunchKernelGGL(batch_permute, dim3(c.total_tiles), dim3(256), 0, stream, c.members,
                       c.first_tile, count, x, y, zp);
    hip_check(hipGetLastError(), "batched permute launch");
    hipLaunchKernelGGL((batch_pass1<KB, FP32>), dim3(c.total_tiles * kPass1Groups, passes),
                       dim3(64 * kGroupWaves), 0, stream, c.members, c.first_tile, count, x, y, zp);
    hip_check(hipGetLastError(), "batched pass 1 launch");
    hipLaunchKernelGGL((batch_pass2<KB, FP32>), dim3(c.total_tiles, passes), dim3(64 * kWaves), 0,
                       stream, c.members, c.first_tile, count, x, y, zp);
    hip_check(hipGetLastError(), "batched pass 2 launch");
}

template <bool FP32>
void dispatch_batch_class(const BatchClass& c, const double* x, double* y, double* zp,
                          hipStream_t st) {
    switch (c.kb) {
        case 1: return launch_batch_class<1, FP32>(c, x, y, zp, st);
        case 2: return launch_batch_class<2, FP32>(c, x, y, zp, st);
        default: return launch_batch_class<4, FP32>(c, x, y, zp, st);
    }
}

template <bool FP32>
void dispatch_narrow(const DeviceStore& s, int kb, const double* x, std::int64_t ldx, int k,
                     double* y, std::int64_t ldy, double* zp, hipStream_t st) {
    switch (kb) {
        case 1: return launch_apply<1, FP32>(s, x, ldx, k, y, ldy, zp, st);
        case 2: return launch_apply<2, FP32>(s, x, ldx, k, y, ldy, zp, st);
        default: return launch_apply<4, FP32>(s, x, ldx, k, y, ldy, zp, st);
    }
}

template <class T> T* upload(const std::vector<T>& host, std::uint64_t& bytes) {
    if (host.empty()) return nullptr;
    T* p = nullptr;
    const std::size_t n = host.size() * sizeof(T);
    hip_check(hipMalloc(reinterpret_cast<void**>(&p), n), "hipMalloc");
    bytes += n;
    hip_check(hipMemcpy(p, host.data(), n, hipMemcpyHostToDevice), "hipMemcpy host-to-device");
    return p;
}

// Relative time of an item for balancing runs, from per-wave clock
// readings of the narrow passes (one right-hand side) fitted over a 2-D
// kernel matrix of order 16384: a fixed part per item (records, staging,
// scans, the pipeline turn) plus a part per component, more for components
// wider than 24 bits (split in FP32 mode).  A diagonal item costs little
// beside its bytes.  Units: hundredths of the per-component time.
std::uint64_t item_cost(const ItemDev& it, const std::vector<std::int32_t>& widths) {
    if (it.kind == kDense) return 7 * std::uint64_t(it.quads);
    std::uint64_t wide = 0;
    for (std::uint32_t j = 0; j < it.count; ++j) wide += widths[it.comp0 + j] > 24 ? 1 : 0;
    const std::uint64_t n = it.count;
    switch (it.kind) {
    case kFused: return 1720 + 164 * n + 94 * wide;
    case kRight: return 1060 + 147 * n + 94 * wide;
    default: return 1340 + 123 * n + 62 * wide;  // kLeft
    }
}

// Tiles of at most kTile indices between consecutive cuts.
std::vector<std::pair<std::uint64_t, std::uint64_t>> make_tiles(std::vector<std::uint64_t> cuts) {
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    std::vector<std::pair<std::uint64_t, std::uint64_t>> tiles;
    for (std::size_t i = 0; i + 1 < cuts.size(); ++i)
        for (std::uint64_t r = cuts[i]; r < cuts[i + 1]; r += kTile)
            tiles.emplace_back(r, std::min<std::uint64_t>(r + kTile, cuts[i + 1]));
    return tiles;
}

// Index of the first tile starting at or after `at`.
std::size_t first_tile(const std::vector<std::pair<std::uint64_t, std::uint64_t>>& tiles,
                       std::uint64_t at) {
    return std::lower_bound(tiles.begin(), tiles.end(), at,
                            [](const auto& t, std::uint64_t r) { return t.first < r; }) -
           tiles.begin();
}

// Cuts `cost` into `parts` contiguous runs whose totals follow `weight`
// (equal when null); out[i] is the first index of run i (out has parts + 1
// entries, offset by `base`).
void split_runs(const std::vector<std::uint64_t>& cost, std::uint32_t parts, std::uint32_t base,
                std::uint32_t* out, const double* weight = nullptr) {
    std::uint64_t total = 0;
    for (auto c : cost) total += c;
    auto weight_of = [&](std::uint32_t w) { return weight ? weight[w] : 1.0; };
    double wsum = 0.0;
    for (std::uint32_t w = 0; w < parts; ++w) wsum += weight_of(w);
    std::uint32_t e = 0;
    std::uint64_t run = 0;
    double wdone = 0.0;
    out[0] = base;
    for (std::uint32_t w = 1; w < parts; ++w) {
        wdone += weight_of(w - 1);
        const std::uint64_t goal = std::uint64_t(double(total) * wdone / wsum);
        while (e < cost.size() && run + cost[e] / 2 < goal) run += cost[e++];
        out[w] = base + e;
    }
    out[parts] = base + std::uint32_t(cost.size());
}

// Among the waves that share a SIMD, issue goes to the oldest first (at
// equal priority), so the wave dispatched first runs fastest and the last
// slowest: with equal work, the pass-1 waves of the four workgroups a CU
// holds took 15, 17, 20 and 22 us (square_exp n=16384, k=1).  The static
// splits therefore give a wave a share of the work by its dispatch rank on
// its SIMD: pass-1 workgroup g has rank (g / kDeviceCUs) mod 4 (workgroups
// go round the CUs in launch order), pass-2 wave w of a tile's workgroup
// rank w / 4.  The weights are 1 / those times, normalised; the waves'
// progress priority (progress_priority) evens out the rest at run time.
constexpr std::uint32_t kDeviceCUs = 256;
double age_weight(std::uint32_t rank) {
    constexpr double kWeight[4] = {1.24, 1.09, 0.93, 0.85};
    return kWeight[rank % 4];
}

// Components of an item wider than 24 bits (split in FP32 mode).
std::uint32_t wide_components(const ItemDev& it, const std::vector<std::int32_t>& widths) {
    std::uint32_t w = 0;
    for (std::uint32_t j = 0; j < it.count; ++j) w += widths[it.comp0 + j] > 24 ? 1u : 0u;
    return w;
}

// Relative time of a wide item (matrix-core passes): per 16 virtual rows of
// a right side block (16 MFMAs per 16 right-hand sides), per 4 of a left
// side block (one MFMA, plus a partial-sum load per column tile), a
// diagonal item as two FP64 steps.
std::uint64_t wide_right_cost(const ItemDev& it, const std::vector<std::int32_t>& widths) {
    const std::uint64_t nv = it.count + wide_components(it, widths);
    return 6 + 16 * ((nv + 15) / 16);
}
std::uint64_t wide_left_cost(const ItemDev& it, const std::vector<std::int32_t>& widths) {
    if (it.kind == kDense) return 10;
    const std::uint64_t nv = it.count + wide_components(it, widths);
    return 8 + 4 * ((nv + 3) / 4);
}

// The wide passes' tables.  Pass R: every right side block (kRight items,
// and the right halves of kFused items, which get partial-sum rows of their
// own after the narrow passes' rows) grouped by column tile; each tile's
// list is cut into workgroups of similar cost (about four per CU over the
// whole store) and each workgroup's share into kWideWaves wave runs.
// Pass L: each row tile's kDense items, the left halves of its kFused items
// and its kLeft items, cut into runs of similar cost (one workgroup each);
// a tile with several runs gets one row set per run.
void build_wide(Plan& plan, const std::vector<ItemDev>& right_items,
                const std::vector<std::vector<ItemDev>>& pass1,
                const std::vector<std::vector<ItemDev>>& pass2,
                const std::vector<std::pair<std::uint64_t, std::uint64_t>>& row_tiles,
                const std::vector<std::pair<std::uint64_t, std::uint64_t>>& col_tiles) {
    const std::vector<std::int32_t>& widths = plan.widths;
    const std::size_t nct = col_tiles.size(), nrt = row_tiles.size();
    auto col_tile_of = [&](std::uint32_t col) {
        const std::size_t t = first_tile(col_tiles, col);
        if (t >= nct || col_tiles[t].first != col)
            throw std::logic_error("device plan: item column is not the start of a column tile");
        return t;
    };
    // quads of the first side block of a kFused item (its right block)
    auto right_quads = [&](const ItemDev& it) {
        std::uint64_t bits = 0;
        for (std::uint32_t j = 0; j < it.count; ++j) bits += std::uint64_t(widths[it.comp0 + j]);
        return std::uint32_t(16 * ((bits + 31) / 32));
    };
    std::vector<std::vector<ItemDev>> vitems(nct), uitems(nrt);
    // Rows of partial sums: the narrow passes' per-component rows of kRight
    // items first (z_rows), then rows in virtual-row order, count + (wide
    // components) per item, the most either precision uses.
    std::uint64_t wz = plan.z_rows;
    auto rows_for = [&](const ItemDev& it) {
        const std::uint64_t r = wz;
        wz += it.count + wide_components(it, widths);
        require(wz < (std::uint64_t{1} << 31), "store too large for one device plan");
        return std::uint32_t(r);
    };
    std::map<std::tuple<std::uint32_t, std::uint32_t, std::uint32_t, std::uint32_t>, std::uint32_t>
        zsum_rows;  // (zrow, zstride, nseg, count) of a kLeft range -> its rows
    for (ItemDev it : right_items) {
        it.nseg = 0;  // partial sums per component
        vitems[col_tile_of(it.col)].push_back(it);
    }
    for (std::size_t t = 0; t < nrt; ++t) {
        for (const auto& it : pass1[t]) {
            if (it.kind == kDense) {
                uitems[t].push_back(it);
                continue;
            }
            const std::uint32_t rq = right_quads(it);
            if (rq == 0 || rq >= it.quads) throw std::logic_error("device plan: fused item sides");
            const std::uint32_t rows = rows_for(it);
            ItemDev v = it;
            v.kind = kRight;
            v.quads = std::uint16_t(rq);
            v.zrow = rows;
            v.nseg = 1;  // partial sums in virtual-row order
            vitems[col_tile_of(it.col)].push_back(v);
            ItemDev u = it;
            u.kind = kLeft;
            u.quad = it.quad + rq;
            u.quads = std::uint16_t(it.quads - rq);
            u.zrow = rows;
            u.zstride = it.count + wide_components(it, widths);  // rows it may read
            u.nseg = 1;
            u.col = u.cols = 0;
            uitems[t].push_back(u);
        }
        for (ItemDev it : pass2[t]) {
            // the column tiles' partial sums of this component range are
            // summed once between the passes, into rows in virtual-row order
            const auto key = std::make_tuple(it.zrow, it.zstride, it.nseg, std::uint32_t(it.count));
            auto found = zsum_rows.find(key);
            if (found == zsum_rows.end()) {
                const std::uint32_t rows = rows_for(it);
                plan.wzsums.push_back(ZSum{it.zrow, it.zstride, it.nseg, it.count, rows, it.comp0, 0, 0});
                found = zsum_rows.emplace(key, rows).first;
            }
            it.zrow = found->second;
            it.zstride = it.count + wide_components(it, widths);
            it.nseg = 1;
            uitems[t].push_back(it);
        }
    }
    plan.wz_rows = wz;
    plan.ctiles.resize(nct);
    for (std::size_t c = 0; c < nct; ++c)
        plan.ctiles[c] = Span{std::uint32_t(col_tiles[c].first),
                              std::uint32_t(col_tiles[c].second - col_tiles[c].first)};

    constexpr std::uint64_t kTargetGroups = 1024;  // about four workgroups per CU
    auto& out = plan.witems;
    // pass R
    std::uint64_t rtotal = 0;
    std::vector<std::vector<std::uint64_t>> rc(nct);
    for (std::size_t c = 0; c < nct; ++c)
        for (const auto& it : vitems[c]) {
            rc[c].push_back(wide_right_cost(it, widths));
            rtotal += rc[c].back();
        }
    // Workgroups per column tile.  All of pass R's workgroups fit on the
    // device at once (kTargetGroups, four per CU), and a workgroup's share
    // follows the dispatch rank on its CU that its size gives it once the
    // workgroups are ordered costliest first (below): the tiles, costliest
    // first, take workgroups of age_weight(rank) units each.
    std::vector<std::uint64_t> tile_cost(nct, 0);
    std::vector<std::size_t> by_cost;
    for (std::size_t c = 0; c < nct; ++c) {
        for (auto v : rc[c]) tile_cost[c] += v;
        if (!vitems[c].empty()) by_cost.push_back(c);
    }
    std::stable_sort(by_cost.begin(), by_cost.end(),
                     [&](std::size_t a, std::size_t b) { return tile_cost[a] > tile_cost[b]; });
    std::vector<std::uint32_t> tile_groups(nct, 0);
    double wsum = 0.0;
    for (std::uint32_t g = 0; g < kTargetGroups; ++g) wsum += age_weight(g / kDeviceCUs);
    double runit = std::max(1.0, double(rtotal) / std::max(wsum, double(nct)));
    for (int attempt = 0; attempt < 32; ++attempt) {
        std::uint64_t placed = 0;
        for (std::size_t c : by_cost) {
            const double unit = runit * age_weight(std::uint32_t(placed / kDeviceCUs));
            const std::uint64_t cap = std::min<std::uint64_t>(64, std::max<std::size_t>(1, (vitems[c].size() + 1) / 2));
            tile_groups[c] = std::uint32_t(std::clamp<std::uint64_t>(
                std::uint64_t(double(tile_cost[c]) / unit + 0.5), 1, cap));
            placed += tile_groups[c];
        }
        if (placed <= std::max<std::uint64_t>(kTargetGroups, by_cost.size())) break;
        runit *= 1.02;
    }
    for (std::size_t c = 0; c < nct; ++c) {
        if (vitems[c].empty()) continue;
        const std::uint32_t groups = tile_groups[c];
        std::vector<std::uint32_t> cut(groups * kWideWaves + 1);
        split_runs(rc[c], groups * kWideWaves, std::uint32_t(out.size()), cut.data());
        out.insert(out.end(), vitems[c].begin(), vitems[c].end());
        for (std::uint32_t w = 0; w < groups * kWideWaves; ++w)
            plan.wruns_r.push_back(WideRun{std::uint32_t(c), cut[w], cut[w + 1], 0});
    }
    // pass L
    std::uint64_t ltotal = 0;
    std::vector<std::vector<std::uint64_t>> lc(nrt);
    for (std::size_t t = 0; t < nrt; ++t)
        for (const auto& it : uitems[t]) {
            lc[t].push_back(wide_left_cost(it, widths));
            ltotal += lc[t].back();
        }
    const std::uint64_t lunit = std::max<std::uint64_t>(1, ltotal / std::max<std::uint64_t>(kTargetGroups, nrt));
    for (std::size_t t = 0; t < nrt; ++t) {
        std::uint64_t cost = 0;
        for (auto v : lc[t]) cost += v;
        const std::uint64_t cap = std::max<std::uint64_t>(1, std::min<std::uint64_t>(8, uitems[t].size()));
        const std::uint32_t runs =
            std::uint32_t(std::clamp<std::uint64_t>((cost + lunit / 2) / lunit, 1, cap));
        // each run is a workgroup; its items are cut again into one share
        // per wave
        std::vector<std::uint32_t> cut(runs * kWideWaves + 1);
        split_runs(lc[t], runs * kWideWaves, std::uint32_t(out.size()), cut.data());
        out.insert(out.end(), uitems[t].begin(), uitems[t].end());
        const std::uint32_t set0 = runs == 1 ? kNoSet : plan.wsets;
        for (std::uint32_t s = 0; s < runs; ++s)
            for (int w = 0; w < kWideWaves; ++w)
                plan.wruns_l.push_back(WideRun{std::uint32_t(t), cut[s * kWideWaves + w],
                                               cut[s * kWideWaves + w + 1],
                                               runs == 1 ? kNoSet : set0 + s});
        if (runs == 1) continue;
        plan.wsums.push_back(WideSum{std::uint32_t(t), set0, runs, 0});
        plan.wsets += runs;
    }
    require(out.size() < (std::uint64_t{1} << 31), "store too large for one device plan");
    // Workgroups in launch order costliest first (by their costliest wave):
    // the oldest workgroups on a CU are issued first and run fastest, and
    // the last ones dispatched, which start as others finish and set the
    // pass's tail, are the shortest.
    auto costliest_first = [&](std::vector<WideRun>& runs, auto item_cost) {
        const std::size_t groups = runs.size() / kWideWaves;
        std::vector<std::uint64_t> cost(groups, 0);
        for (std::size_t g = 0; g < groups; ++g)
            for (int w = 0; w < kWideWaves; ++w) {
                const WideRun& r = runs[g * kWideWaves + w];
                std::uint64_t c = 0;
                for (std::uint32_t i = r.begin; i < r.end; ++i) c += item_cost(out[i], widths);
                cost[g] = std::max(cost[g], c);
            }
        std::vector<std::size_t> order(groups);
        for (std::size_t g = 0; g < groups; ++g) order[g] = g;
        std::stable_sort(order.begin(), order.end(),
                         [&](std::size_t a, std::size_t b) { return cost[a] > cost[b]; });
        std::vector<WideRun> sorted;
        sorted.reserve(runs.size());
        for (std::size_t g : order)
            sorted.insert(sorted.end(), runs.begin() + g * kWideWaves,
                          runs.begin() + (g + 1) * kWideWaves);
        runs = std::move(sorted);
    };
    costliest_first(plan.wruns_r, wide_right_cost);
    costliest_first(plan.wruns_l, wide_left_cost);
}

}  // namespace

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

Plan build_plan(const qh_store_desc& d) {
    require(d.order > 0 && d.perm && d.nblocks > 0 && d.rect && d.is_dense && d.rank,
            "incomplete store description");
    require(d.order < (std::uint64_t{1} << 26),
            "matrix order exceeds the device store limit (2^26)");
    const std::uint64_t n = d.order;

    std::vector<std::int32_t> perm(n);
    {
        std::vector<char> seen(n, 0);
        for (std::uint64_t i = 0; i < n; ++i) {
            require(d.perm[i] < n && !seen[d.perm[i]], "permutation is not a bijection");
            seen[d.perm[i]] = 1;
            perm[i] = static_cast<std::int32_t>(d.perm[i]);
        }
    }
    std::vector<std::uint64_t> row_cuts{0, n}, col_cuts{0, n};
    for (std::uint64_t b = 0; b < d.nblocks; ++b) {
        const std::uint64_t* r = d.rect + 4 * b;
        require(r[0] < r[1] && r[1] <= n && r[2] < r[3] && r[3] <= n, "invalid block rectangle");
        row_cuts.push_back(r[0]);
        row_cuts.push_back(r[1]);
        col_cuts.push_back(r[2]);
        col_cuts.push_back(r[3]);
    }
    const auto row_tiles = make_tiles(row_cuts);
    const auto col_tiles = make_tiles(col_cuts);

    Plan plan;
    plan.order = n;
    std::vector<std::uint32_t>& words = plan.words;
    std::vector<ItemDev> right_items;
    std::vector<std::vector<ItemDev>> pass1(row_tiles.size()), pass2(row_tiles.size());
    plan.block_to_lowrank.assign(d.nblocks, -1);
    std::uint64_t codes = 0, lr_flops = 0, dn_flops = 0, z_rows = 0, dense_values = 0;
    std::uint64_t n_lowrank_all = 0, left_words = 0, right_words = 0;
    std::uint64_t pass_words[2] = {0, 0}, fused_components = 0;
    std::vector<std::vector<std::uint32_t>> dense_quad(d.nblocks);  // per dense block, row tile

    // Starts an item's data on a 16-byte boundary; returns its first quad.
    auto begin_item = [&words]() {
        while (words.size() & 3u) words.push_back(0u);
        require(words.size() / 4 < (std::uint64_t{1} << 32), "store too large for one device plan");
        return std::uint32_t(words.size() / 4);
    };
    auto finish_item = [&words](ItemDev& it) {
        const std::uint64_t quads = (words.size() - 4 * std::uint64_t(it.quad) + 3) / 4;
        if (quads > std::uint64_t(kItemQuads))
            throw std::logic_error("device plan: item larger than the stage buffer");
        it.quads = std::uint16_t(quads);
    };

    for (std::uint64_t b = 0; b < d.nblocks; ++b) {
        const std::uint64_t r0 = d.rect[4 * b], r1 = d.rect[4 * b + 1];
        const std::uint64_t c0 = d.rect[4 * b + 2], c1 = d.rect[4 * b + 3];
        const std::uint64_t m = r1 - r0, nc = c1 - c0;
        const std::size_t rt0 = first_tile(row_tiles, r0), rt1 = first_tile(row_tiles, r1);
        const std::size_t ct0 = first_tile(col_tiles, c0), ct1 = first_tile(col_tiles, c1);
        if (d.is_dense[b]) {
            require(d.dense && d.dense[b], "dense block without values");
            const double* v = d.dense[b];
            for (std::size_t t = rt0; t < rt1; ++t) {
                const std::uint64_t a = row_tiles[t].first - r0;
                const std::uint64_t R = row_tiles[t].second - row_tiles[t].first;
                for (std::uint64_t c = 0; c < nc; c += kDenseColumns) {
                    const std::uint64_t cnt = std::min<std::uint64_t>(kDenseColumns, nc - c);
                    ItemDev it{};
                    it.quad = begin_item();
                    if (c == 0) dense_quad[b].push_back(it.quad);
                    it.kind = kDense;
                    it.count = std::uint8_t(cnt);
                    it.col = std::uint32_t(c0 + c);
                    for (std::uint64_t cc = 0; cc < cnt; ++cc)  // row-major in, column-major out
                        for (std::uint64_t i = 0; i < R; ++i) {
                            std::uint64_t bitsv;
                            std::memcpy(&bitsv, v + (a + i) * nc + c + cc, 8);
                            words.push_back(std::uint32_t(bitsv));
                            words.push_back(std::uint32_t(bitsv >> 32));
                        }
                    finish_item(it);
                    pass1[t].push_back(it);
                }
            }
            dn_flops += 2 * m * nc;
            dense_values += m * nc;
            continue;
        }
        ++n_lowrank_all;
        const std::uint64_t r = d.rank[b];
        require(r <= std::min(m, nc), "factor dimensions do not match block rank");
        if (r == 0) continue;  // contributes nothing
        require(d.widths && d.widths[b] && d.lscale && d.lscale[b] && d.rscale && d.rscale[b] &&
                    d.lwords && d.lwords[b] && d.rwords && d.rwords[b],
                "low-rank block without factors");
        const std::int64_t li = std::int64_t(plan.ranks.size());
        const std::uint32_t comp_base = std::uint32_t(plan.widths.size());
        std::vector<std::uint64_t> loff(r), roff(r), lcnt(r), rcnt(r);
        std::uint64_t lw = 0, rw = 0;
        for (std::uint64_t j = 0; j < r; ++j) {
            const int bits = d.widths[b][j];
            require(bits >= 2 && bits <= 32, "bit width must be in [2, 32]");
            plan.widths.push_back(bits);
            plan.cscale.push_back(d.lscale[b][j] * d.rscale[b][j]);
            plan.rec.lscale.push_back(d.lscale[b][j]);
            plan.rec.rscale.push_back(d.rscale[b][j]);
            loff[j] = lw;
            roff[j] = rw;
            lcnt[j] = (m * bits + 31) / 32;
            rcnt[j] = (nc * bits + 31) / 32;
            lw += lcnt[j];
            rw += rcnt[j];
        }
        left_words += lw;
        right_words += rw;
        plan.left_chunks.emplace_back();
        plan.right_chunks.emplace_back();
        auto& lch = plan.left_chunks.back();
        auto& rch = plan.right_chunks.back();
        const int* wd = d.widths[b];
        // Components [j0, j) of one item with at most `cap` components whose
        // side blocks fit the item: `sides` blocks of 64 lanes x
        // ceil(sum of widths / 32) words.
        auto group_end = [&](std::uint64_t j0, std::uint64_t cap, std::uint64_t sides) {
            std::uint64_t j = j0, bits = 0;
            while (j < r && j - j0 < cap) {
                const std::uint64_t nb = bits + std::uint64_t(wd[j]);
                if (sides * 64 * ((nb + 31) / 32) > 4 * std::uint64_t(kItemQuads)) break;
                bits = nb;
                ++j;
            }
            if (j == j0) throw std::logic_error("device plan: component codes exceed an item");
            return j;
        };
        // Side block of components [j0, j1) of one factor over its rows (or
        // columns) [a, a + n) of the block, n <= 64: lane L holds the codes of
        // row a + L, one after another (component j at bit offset sum of the
        // widths before it), word w of lane L at word 64 w + lane_slot(w, L)
        // of the block (swizzled for right factors).
        auto append_side = [&](const std::uint32_t* const* side_words,
                               const std::vector<std::uint64_t>& off,
                               const std::vector<std::uint64_t>& cnt, std::uint64_t j0,
                               std::uint64_t j1, std::uint64_t a, std::uint64_t n, bool swz) {
            std::uint64_t total = 0;
            for (std::uint64_t j = j0; j < j1; ++j) total += std::uint64_t(wd[j]);
            const std::uint64_t nw = (total + 31) / 32, base = words.size();
            words.resize(base + 64 * nw, 0u);
            for (std::uint64_t L = 0; L < n; ++L) {
                std::uint64_t o = 0;
                for (std::uint64_t j = j0; j < j1; ++j) {
                    const unsigned bits = unsigned(wd[j]);
                    const std::uint32_t* st = side_words[b] + off[j];
                    const std::uint64_t src = (a + L) * bits, w = src >> 5;
                    const unsigned sh = unsigned(src & 31);
                    std::uint64_t pair = st[w];
                    if (sh + bits > 32 && w + 1 < cnt[j]) pair |= std::uint64_t(st[w + 1]) << 32;
                    const std::uint64_t v = (pair >> sh) & ((std::uint64_t{1} << bits) - 1);
                    const std::uint64_t k = o >> 5, s = o & 31;
                    words[base + 64 * k + lane_slot(swz, k, L)] |= std::uint32_t(v << s);
                    if (s + bits > 32)
                        words[base + 64 * (k + 1) + lane_slot(swz, k + 1, L)] |= std::uint32_t(v >> (32 - s));
                    o += bits;
                }
            }
        };

        if (rt1 - rt0 == 1 && ct1 - ct0 == 1) {
            // One row tile and one column tile: both factors in one item.
            const std::uint64_t R = m, Cn = nc;
            for (std::uint64_t j0 = 0; j0 < r;) {
                const std::uint64_t j1 = group_end(j0, kItemComponents, 2);
                ItemDev it{};
                it.quad = begin_item();
                it.kind = kFused;
                it.count = std::uint8_t(j1 - j0);
                it.comp0 = comp_base + std::uint32_t(j0);
                it.col = std::uint32_t(c0);
                it.cols = std::uint32_t(Cn);
                rch.push_back({0u, std::uint32_t(Cn), std::uint32_t(j0), std::uint32_t(j1 - j0),
                               words.size()});
                append_side(d.rwords, roff, rcnt, j0, j1, 0, Cn, true);
                lch.push_back({0u, std::uint32_t(R), std::uint32_t(j0), std::uint32_t(j1 - j0),
                               words.size()});
                append_side(d.lwords, loff, lcnt, j0, j1, 0, R, false);
                finish_item(it);
                pass1[rt0].push_back(it);
                j0 = j1;
            }
            pass_words[0] += lw + rw;
            fused_components += r;
        } else {
            // Column tiles in runs of at most 64: each run has its own partial
            // sums, read by the run's kLeft items (64 pairs per item).
            for (std::size_t v0 = ct0; v0 < ct1; v0 += 64) {
                const std::size_t v1 = std::min(ct1, v0 + 64);
                const std::uint64_t nseg = v1 - v0;
                const std::uint64_t zbase = z_rows;
                for (std::size_t t = v0; t < v1; ++t) {
                    const std::uint64_t a = col_tiles[t].first - c0;
                    const std::uint64_t Cn = col_tiles[t].second - col_tiles[t].first;
                    for (std::uint64_t j0 = 0; j0 < r;) {
                        const std::uint64_t j1 = group_end(j0, kItemComponents, 1);
                        ItemDev it{};
                        it.quad = begin_item();
                        it.kind = kRight;
                        it.count = std::uint8_t(j1 - j0);
                        it.comp0 = comp_base + std::uint32_t(j0);
                        it.col = std::uint32_t(col_tiles[t].first);
                        it.cols = std::uint32_t(Cn);
                        it.zrow = std::uint32_t(zbase + (t - v0) * r + j0);
                        rch.push_back({std::uint32_t(a), std::uint32_t(Cn), std::uint32_t(j0),
                                       std::uint32_t(j1 - j0), words.size()});
                        append_side(d.rwords, roff, rcnt, j0, j1, a, Cn, true);
                        finish_item(it);
                        right_items.push_back(it);
                        j0 = j1;
                    }
                }
                z_rows += nseg * r;
                require(z_rows < (std::uint64_t{1} << 31), "store too large for one device plan");
                for (std::size_t t = rt0; t < rt1; ++t) {
                    const std::uint64_t a = row_tiles[t].first - r0;
                    const std::uint64_t R = row_tiles[t].second - row_tiles[t].first;
                    for (std::uint64_t j0 = 0; j0 < r;) {
                        const std::uint64_t j1 =
                            group_end(j0, std::max<std::uint64_t>(1, 64 / nseg), 1);
                        ItemDev it{};
                        it.quad = begin_item();
                        it.kind = kLeft;
                        it.count = std::uint8_t(j1 - j0);
                        it.comp0 = comp_base + std::uint32_t(j0);
                        it.zrow = std::uint32_t(zbase + j0);
                        it.zstride = std::uint32_t(r);
                        it.nseg = std::uint32_t(nseg);
                        if (v0 == ct0)
                            lch.push_back({std::uint32_t(a), std::uint32_t(R), std::uint32_t(j0),
                                           std::uint32_t(j1 - j0), words.size()});
                        append_side(d.lwords, loff, lcnt, j0, j1, a, R, false);
                        finish_item(it);
                        pass2[t].push_back(it);
                        j0 = j1;
                    }
                }
            }
            pass_words[0] += rw;
            pass_words[1] += lw;
        }
        plan.block_to_lowrank[b] = li;
        plan.comp_base.push_back(comp_base);
        plan.ranks.push_back(std::uint32_t(r));
        codes += (m + nc) * r;
        lr_flops += 2 * r * (m + nc);
    }
    plan.stored_words = words.size();
    words.resize(words.size() + 72, 0u);  // one readable row of 64 words past the last item

    // Reconstruction: one task per (block, row tile, column tile).
    for (std::uint64_t b = 0; b < d.nblocks; ++b) {
        const std::uint64_t r0 = d.rect[4 * b], r1 = d.rect[4 * b + 1];
        const std::uint64_t c0 = d.rect[4 * b + 2], c1 = d.rect[4 * b + 3];
        const std::size_t rt0 = first_tile(row_tiles, r0), rt1 = first_tile(row_tiles, r1);
        const std::size_t ct0 = first_tile(col_tiles, c0), ct1 = first_tile(col_tiles, c1);
        const std::int64_t li = plan.block_to_lowrank[b];
        for (std::size_t t = rt0; t < rt1; ++t)
            for (std::size_t u = ct0; u < ct1; ++u) {
                RecTask k{};
                k.row0 = std::uint32_t(row_tiles[t].first);
                k.rows = std::uint32_t(row_tiles[t].second - row_tiles[t].first);
                k.col0 = std::uint32_t(col_tiles[u].first);
                k.cols = std::uint32_t(col_tiles[u].second - col_tiles[u].first);
                if (d.is_dense[b]) {
                    k.kind = 1;
                    k.quad = dense_quad[b].at(t - rt0);
                    k.bcol0 = std::uint32_t(c0);
                } else if (li >= 0) {
                    k.kind = 2;
                    k.rank = plan.ranks[li];
                    k.comp_base = plan.comp_base[li];
                    k.lchunk = std::uint32_t(plan.rec.chunks.size());
                    // chunk rows and columns count from the block's first
                    for (const auto& ch : plan.left_chunks[li])
                        if (ch.first == k.row0 - r0) plan.rec.chunks.push_back(ch);
                    k.nl = std::uint32_t(plan.rec.chunks.size()) - k.lchunk;
                    k.rchunk = std::uint32_t(plan.rec.chunks.size());
                    for (const auto& ch : plan.right_chunks[li])
                        if (ch.first == k.col0 - c0) plan.rec.chunks.push_back(ch);
                    k.nr = std::uint32_t(plan.rec.chunks.size()) - k.rchunk;
                }
                plan.rec.tasks.push_back(k);
            }
    }

    // Pass assignment: kDense, kFused and kRight items run in pass 1 (kRight
    // in any run, the others in their tile's runs), kLeft items in pass 2.
    auto cost_of = [&](const ItemDev& it) { return item_cost(it, plan.widths); };
    const std::size_t ntiles = row_tiles.size();
    std::uint64_t right_load = 0;
    for (const auto& it : right_items) right_load += cost_of(it);
    const std::uint64_t pass_fused[2] = {fused_components, 0};
    const std::uint64_t pass_dense[2] = {dense_values, 0};

    // Item table: right-factor items, then per tile its pass-1 and pass-2 items.
    std::vector<ItemDev>& items = plan.items;
    items = right_items;
    plan.tiles.resize(ntiles);
    std::vector<std::uint64_t> run_cost;  // pass-1 cost of every (tile, run)
    for (std::size_t t = 0; t < ntiles; ++t) {
        TileDev& T = plan.tiles[t];
        T.row0 = std::uint32_t(row_tiles[t].first);
        T.rows = std::uint32_t(row_tiles[t].second - row_tiles[t].first);
        for (int pass = 0; pass < 2; ++pass) {
            const auto& list = pass == 0 ? pass1[t] : pass2[t];
            std::vector<std::uint64_t> cost;
            for (const auto& it : list) cost.push_back(cost_of(it));
            std::uint32_t* split = pass == 0 ? T.split1 : T.split2;
            double weight[kWaves];
            for (int w = 0; w < kWaves; ++w) weight[w] = pass == 0 ? 1.0 : age_weight(std::uint32_t(w) / 4);
            split_runs(cost, kWaves, std::uint32_t(items.size()), split, weight);
            if (pass == 0)
                for (int w = 0; w < kWaves; ++w) {
                    std::uint64_t c = 0;
                    for (std::uint32_t i = split[w]; i < split[w + 1]; ++i)
                        c += cost[i - split[0]];
                    run_cost.push_back(c);
                }
            items.insert(items.end(), list.begin(), list.end());
        }
    }
    require(items.size() < (std::uint64_t{1} << 31), "store too large for one device plan");
    // The right-factor items fill the pass-1 runs up to totals that follow
    // the runs' dispatch-rank weights: run r takes items (in order) while
    // the cost of runs 0..r stays under their share of the whole, so that
    // rounding does not pile up in the last run.
    std::uint64_t all = 0;
    for (auto c : run_cost) all += c;
    all += right_load;
    std::vector<double> run_weight(run_cost.size());
    double wsum = 0.0;
    for (std::size_t r = 0; r < run_cost.size(); ++r) {
        const std::uint32_t group = std::uint32_t(r / kGroupWaves);
        run_weight[r] = age_weight(group / kDeviceCUs);
        wsum += run_weight[r];
    }
    std::uint32_t next = 0;
    std::uint64_t done = 0;  // cost of the runs filled so far
    double wdone = 0.0;
    for (std::size_t r = 0; r < run_cost.size(); ++r) {
        TileDev& T = plan.tiles[r / kWaves];
        const int w = int(r % kWaves);
        T.rsplit[w] = next;
        done += run_cost[r];
        wdone += run_weight[r];
        const std::uint64_t goal = std::uint64_t(double(all) * wdone / wsum);
        const bool last = r + 1 == run_cost.size();
        while (next < right_items.size() &&
               (last || done + cost_of(right_items[next]) / 2 <= goal)) {
            done += cost_of(right_items[next]);
            ++next;
        }
        T.rsplit[w + 1] = next;
    }
    plan.z_rows = z_rows;
    // each wave's run and its first two records
    plan.heads.assign(2 * ntiles * std::size_t(kWaves), WaveHead{});
    for (std::size_t t = 0; t < ntiles; ++t) {
        const TileDev& T = plan.tiles[t];
        for (int w = 0; w < kWaves; ++w) {
            for (int pass = 0; pass < 2; ++pass) {
                WaveHead& h = plan.heads[(pass * ntiles + t) * kWaves + std::size_t(w)];
                if (pass == 0) {
                    h.b1 = T.rsplit[w], h.e1 = T.rsplit[w + 1], h.b2 = T.split1[w], h.e2 = T.split1[w + 1];
                } else {
                    h.b1 = h.e1 = 0, h.b2 = T.split2[w], h.e2 = T.split2[w + 1];
                }
                const std::uint32_t n1 = h.e1 - h.b1, total = n1 + (h.e2 - h.b2);
                for (std::uint32_t i = 0; i < 2 && i < total; ++i)
                    h.first[i] = items[i < n1 ? h.b1 + i : h.b2 + (i - n1)];
            }
        }
    }
    build_wide(plan, right_items, pass1, pass2, row_tiles, col_tiles);

    qh_store_stats& st = plan.stats;
    st.order = n;
    st.nblocks = d.nblocks;
    std::uint64_t ndense = 0;
    for (std::uint64_t b = 0; b < d.nblocks; ++b) ndense += d.is_dense[b] ? 1 : 0;
    st.ndense = ndense;
    st.nlowrank = n_lowrank_all;
    st.ncomponents = plan.widths.size();
    st.code_words = left_words + right_words;
    st.left_code_words = left_words;
    st.right_code_words = right_words;
    st.codes = codes;
    st.dense_values = dense_values;
    st.lowrank_flops_per_rhs = lr_flops;
    st.dense_flops_per_rhs = dn_flops;
    st.tiles = plan.tiles.size();
    st.items = items.size();
    std::uint64_t right_runs = 0;
    for (const auto& T : plan.tiles)
        for (int w = 0; w < kWaves; ++w) right_runs += T.rsplit[w + 1] > T.rsplit[w] ? 1 : 0;
    st.right_waves = right_runs;
    st.pass_code_words[0] = pass_words[0];
    st.pass_code_words[1] = pass_words[1];
    st.fused_components = fused_components;
    for (int p = 0; p < 2; ++p) {
        st.pass_fused_components[p] = pass_fused[p];
        st.pass_dense_values[p] = pass_dense[p];
    }
    plan.perm = std::move(perm);
    return plan;
}

void upload_rec_tables(const DeviceStore& s);

void reconstruct(const DeviceStore& s, double* out, std::int64_t ld, void* stream) {
    require(out != nullptr, "null output pointer");
    require(ld >= std::int64_t(s.order), "leading dimension is smaller than the matrix order");
    DeviceScope scope(s.device);
    upload_rec_tables(s);
    if (s.rec.host.tasks.empty()) return;
    hipLaunchKernelGGL(reconstruct_tiles, dim3(unsigned(s.rec.host.tasks.size())), dim3(256), 0,
                       static_cast<hipStream_t>(stream), s.rec.tasks, s.rec.chunks, s.words,
                       s.widths, s.rec.lscale, s.rec.rscale, s.perm, out, ld);
    hip_check(hipGetLastError(), "reconstruct launch");
}

// The reconstruction tables (per-component left and right scales among
// them) on the device, uploaded on first use.
void upload_rec_tables(const DeviceStore& s) {
    std::lock_guard<std::mutex> guard(s.rec.lock);
    if (s.rec.uploaded) return;
    std::uint64_t bytes = 0;
    s.rec.tasks = upload(s.rec.host.tasks, bytes);
    s.rec.chunks = upload(s.rec.host.chunks, bytes);
    s.rec.lscale = upload(s.rec.host.lscale, bytes);
    s.rec.rscale = upload(s.rec.host.rscale, bytes);
    s.rec.uploaded = true;
}

FactorPlan* create_factor_plan(const DeviceStore& s, const std::uint64_t* blocks, const int* sides,
                               const std::uint64_t* offsets, std::uint64_t count,
                               std::uint64_t out_floats) {
    require(count == 0 || (blocks && sides && offsets), "null factor slot list");
    std::vector<FactorTask> tasks;
    for (std::uint64_t i = 0; i < count; ++i) {
        const std::uint64_t block = blocks[i];
        const int side = sides[i];
        require(block < s.block_to_lowrank.size(), "block index out of range");
        require(side == 0 || side == 1, "side must be 0 (left) or 1 (right)");
        const std::int64_t li = s.block_to_lowrank[block];
        require(li >= 0, "factor slot of a dense or rank-0 block");
        const auto& chunks = side == 0 ? s.left_chunks[li] : s.right_chunks[li];
        std::uint32_t rows = 0;
        for (const auto& c : chunks) rows = std::max(rows, c.first + c.count);
        const std::uint32_t rank = s.ranks[li];
        require(offsets[i] <= out_floats && std::uint64_t(rows) * rank <= out_floats - offsets[i],
                "factor slot past the end of the output");
        for (const auto& c : chunks)
            tasks.push_back(FactorTask{c.word, offsets[i], c.first, c.count, c.j0, c.jn,
                                       std::uint32_t(s.comp_base[li]), std::uint32_t(side), rows,
                                       rank});
    }
    require(tasks.size() < (std::uint64_t{1} << 31), "too many factor chunks");
    DeviceScope scope(s.device);
    upload_rec_tables(s);
    auto* p = new FactorPlan;
    p->store = &s;
    p->count = std::uint32_t(tasks.size());
    try {
        std::uint64_t bytes = 0;
        p->tasks = upload(tasks, bytes);
    } catch (...) {
        delete p;
        throw;
    }
    s.pins.fetch_add(1);  // the store refuses destruction while the plan exists
    return p;
}

void factor_plan_decode(const FactorPlan& p, float* out, void* stream) {
    require(out != nullptr, "null output pointer");
    if (p.count == 0) return;
    const DeviceStore& s = *p.store;
    DeviceScope scope(s.device);
    hipLaunchKernelGGL(decode_factors, dim3(p.count), dim3(256), 0, static_cast<hipStream_t>(stream),
                       p.tasks, s.words, s.widths, s.rec.lscale, s.rec.rscale, out);
    hip_check(hipGetLastError(), "factor decode launch");
}

void destroy_factor_plan(FactorPlan* p) {
    if (!p) return;
    if (p->tasks) {
        DeviceScope scope(p->store->device);
        (void)hipFree(p->tasks);
    }
    p->store->pins.fetch_sub(1);
    delete p;
}

DeviceStore* create_device_store(const qh_store_desc& d, int device) {
    const int count = device_count();
    if (count == 0) throw std::runtime_error("no CUDA device is available (no HIP device found)");
    require(device >= 0 && device < count, "invalid device ordinal");
    Plan plan = build_plan(d);
    auto* s = new DeviceStore;
    s->device = device;
    s->order = plan.order;
    try {
        DeviceScope scope(device);
        std::uint64_t bytes = 0;
        s->perm = upload(plan.perm, bytes);
        s->words = upload(plan.words, bytes);
        s->widths = upload(plan.widths, bytes);
        s->cscale = upload(plan.cscale, bytes);
        s->items = upload(plan.items, bytes);
        s->tiles = upload(plan.tiles, bytes);
        s->heads = upload(plan.heads, bytes);
        s->witems = upload(plan.witems, bytes);
        s->wruns_r = upload(plan.wruns_r, bytes);
        s->wruns_l = upload(plan.wruns_l, bytes);
        s->wsums = upload(plan.wsums, bytes);
        s->wzsums = upload(plan.wzsums, bytes);
        s->ctiles = upload(plan.ctiles, bytes);
        s->device_bytes = bytes;
    } catch (...) {
        destroy_device_store(s);
        throw;
    }
    s->n_words = plan.stored_words;
    s->n_comp = plan.widths.size();
    s->n_tiles = std::uint32_t(plan.tiles.size());
    s->n_items = std::uint32_t(plan.items.size());
    s->z_rows = plan.z_rows;
    s->n_wr_groups = std::uint32_t(plan.wruns_r.size() / kWideWaves);
    s->n_wl_groups = std::uint32_t(plan.wruns_l.size() / kWideWaves);
    s->n_wsums = std::uint32_t(plan.wsums.size());
    s->n_wzsums = std::uint32_t(plan.wzsums.size());
    s->wsets = plan.wsets;
    s->wz_rows = plan.wz_rows;
    s->block_to_lowrank = std::move(plan.block_to_lowrank);
    s->left_chunks = std::move(plan.left_chunks);
    s->right_chunks = std::move(plan.right_chunks);
    s->comp_base = std::move(plan.comp_base);
    s->ranks = std::move(plan.ranks);
    s->stats = plan.stats;
    s->rec.host = std::move(plan.rec);
    s->stats.device_bytes = s->device_bytes;
    s->stats.device = device;
    return s;
}

// Bounds-checked element access for the host emulation.
template <class V> const typename V::value_type& at(const V& v, std::uint64_t i) {
    if (i >= v.size()) throw std::out_of_range("device plan index out of range");
    return v[i];
}

void emulate_plan(const Plan& pl, const double* x, std::int64_t k, double* y) {
    const std::uint64_t n = pl.order;
    const std::uint64_t K = std::uint64_t(k);
    std::vector<double> zp((pl.z_rows + n) * K, 0.0);
    auto check = [](bool ok, const char* what) {
        if (!ok) throw std::out_of_range(what);
    };
    // Word `w` of an item's data: inside what the wave stages, plus the one
    // word past the data that a decode may read.
    auto word = [&](const ItemDev& it, std::uint64_t w) {
        check(w < 4 * std::uint64_t(kItemQuads + 1), "stage read out of range");
        return at(pl.words, 4 * std::uint64_t(it.quad) + w);
    };
    // Code of lane `lane` at bit offset `o` of the side block starting at
    // word `side` of the item (word w of the lane at 64 w + lane).
    auto code_at = [&](const ItemDev& it, std::uint64_t side, std::uint64_t o, std::uint32_t lane,
                       std::uint32_t bits, bool swz) {
        check(lane < 64, "lane out of range");
        const std::uint64_t k = o >> 5, sh = o & 31, k1 = (o + bits - 1) >> 5;
        const std::uint64_t last = side + 64 * k1 + lane_slot(swz, k1, lane);
        check(last < 4 * std::uint64_t(it.quads), "code outside the item data");
        std::uint64_t pair = word(it, side + 64 * k + lane_slot(swz, k, lane));
        if (sh + bits > 32)
            pair |= std::uint64_t(word(it, side + 64 * (k + 1) + lane_slot(swz, k + 1, lane))) << 32;
        const std::uint64_t raw = (pair >> sh) & ((std::uint64_t{1} << bits) - 1);
        const std::uint64_t sign = std::uint64_t{1} << (bits - 1);
        return double(std::int64_t(raw ^ sign) - std::int64_t(sign));
    };
    // every wave's head record: its run bounds and first two records
    check(pl.heads.size() == 2 * pl.tiles.size() * std::size_t(kWaves), "wave head count");
    for (std::size_t t = 0; t < pl.tiles.size(); ++t)
        for (int w = 0; w < kWaves; ++w)
            for (int pass = 0; pass < 2; ++pass) {
                const TileDev& T = pl.tiles[t];
                const WaveHead& h = pl.heads[(pass * pl.tiles.size() + t) * kWaves + std::size_t(w)];
                const bool bounds =
                    pass == 0 ? h.b1 == T.rsplit[w] && h.e1 == T.rsplit[w + 1] && h.b2 == T.split1[w] &&
                                    h.e2 == T.split1[w + 1]
                              : h.b1 == 0 && h.e1 == 0 && h.b2 == T.split2[w] && h.e2 == T.split2[w + 1];
                check(bounds, "wave head bounds");
                const std::uint32_t n1 = h.e1 - h.b1, total = n1 + (h.e2 - h.b2);
                for (std::uint32_t i = 0; i < 2; ++i) {
                    const ItemDev want = i < total ? at(pl.items, i < n1 ? h.b1 + i : h.b2 + (i - n1)) : ItemDev{};
                    check(std::memcmp(&want, &h.first[i], sizeof(ItemDev)) == 0, "wave head record");
                }
            }
    auto check_item = [&](const ItemDev& it) {
        check(it.quads <= kItemQuads, "item size");
        check(it.count >= 1 && it.count <= kItemComponents, "item component count");
        check(4 * (std::uint64_t(it.quad) + it.quads) <= pl.words.size(), "item data range");
    };
    auto xval = [&](std::uint32_t col, std::uint64_t p) {
        return x[std::uint64_t(at(pl.perm, col)) * K + p];
    };
    // Words of a side block of the item's components (64 lanes).
    auto side_words = [&](const ItemDev& it) {
        std::uint64_t total = 0;
        for (std::uint32_t g = 0; g < it.count; ++g) total += at(pl.widths, it.comp0 + g);
        return 64 * ((total + 31) / 32);
    };
    // Right-factor sums of components [0, count) of an item whose right side
    // block starts at word 0: z[g][p].  Returns the start of the next block.
    auto right_sums = [&](const ItemDev& it, std::vector<double>& z) {
        check(it.cols >= 1 && it.cols <= std::uint32_t(kTile), "item columns");
        z.assign(std::uint64_t(it.count) * K, 0.0);
        std::uint64_t o = 0;
        for (std::uint32_t g = 0; g < it.count; ++g) {
            const std::uint32_t bits = at(pl.widths, it.comp0 + g);
            for (std::uint64_t p = 0; p < K; ++p) {
                double s = 0.0;
                for (std::uint32_t c = 0; c < it.cols; ++c)
                    s += code_at(it, 0, o, c, bits, true) * xval(it.col + c, p);
                z[g * K + p] = s;
            }
            o += bits;
        }
        return side_words(it);
    };
    // Left-factor rows: acc[i][p] += sum_g code(i, g) * cscale_g * z[g][p].
    auto left_rows = [&](const ItemDev& it, std::uint64_t side, std::uint32_t R,
                         const std::vector<double>& z, std::vector<double>& acc) {
        std::uint64_t o = 0;
        for (std::uint32_t g = 0; g < it.count; ++g) {
            const std::uint32_t bits = at(pl.widths, it.comp0 + g);
            const double sc = at(pl.cscale, it.comp0 + g);
            for (std::uint32_t i = 0; i < R; ++i)
                for (std::uint64_t p = 0; p < K; ++p)
                    acc[i * K + p] += code_at(it, side, o, i, bits, false) * (sc * z[g * K + p]);
            o += bits;
        }
    };
    std::vector<double> z;
    // Pass 1, the right-factor runs of every tile.
    for (const TileDev& T : pl.tiles)
        for (std::uint32_t i = T.rsplit[0]; i < T.rsplit[kWaves]; ++i) {
            for (int w = 0; w < kWaves; ++w)
                check(T.rsplit[w] <= T.rsplit[w + 1], "right runs out of order");
            const ItemDev& it = at(pl.items, i);
            check_item(it);
            check(it.kind == kRight, "right wave holds another item kind");
            right_sums(it, z);
            for (std::uint32_t g = 0; g < it.count; ++g)
                for (std::uint64_t p = 0; p < K; ++p)
                    zp.at((std::uint64_t(it.zrow) + g) * K + p) = z[g * K + p];
        }
    // Pass 1 and pass 2 tiles, waves added in order.
    for (int pass = 0; pass < 2; ++pass)
        for (const TileDev& T : pl.tiles) {
            const std::uint32_t R = T.rows;
            check(R >= 1 && R <= std::uint32_t(kTile), "tile rows");
            const std::uint32_t* split = pass == 0 ? T.split1 : T.split2;
            std::vector<double> total(std::uint64_t(R) * K, 0.0);
            if (pass == 1)
                for (std::uint32_t i = 0; i < R; ++i)
                    for (std::uint64_t p = 0; p < K; ++p)
                        total[i * K + p] = zp.at((pl.z_rows + T.row0 + i) * K + p);
            for (int w = 0; w < kWaves; ++w) {
                std::vector<double> acc(std::uint64_t(R) * K, 0.0);
                for (std::uint32_t e = split[w]; e < split[w + 1]; ++e) {
                    const ItemDev& it = at(pl.items, e);
                    check_item(it);
                    if (it.kind == kDense) {
                        check(it.count <= kDenseColumns, "dense item columns");
                        for (std::uint32_t c = 0; c < it.count; ++c)
                            for (std::uint32_t i = 0; i < R; ++i) {
                                const std::uint64_t wv = 2 * (std::uint64_t(c) * R + i);
                                const std::uint64_t bitsv =
                                    std::uint64_t(word(it, wv)) | (std::uint64_t(word(it, wv + 1)) << 32);
                                check(wv + 2 <= 4 * std::uint64_t(it.quads), "dense value outside the item");
                                double v;
                                std::memcpy(&v, &bitsv, 8);
                                for (std::uint64_t p = 0; p < K; ++p)
                                    acc[i * K + p] += v * xval(it.col + c, p);
                            }
                    } else if (it.kind == kFused) {
                        const std::uint64_t roff = right_sums(it, z);
                        left_rows(it, roff, R, z, acc);
                    } else if (pass == 1 && it.kind == kLeft) {
                        check(std::uint64_t(it.count) * it.nseg <= 64 && it.nseg >= 1, "left item pairs");
                        z.assign(std::uint64_t(it.count) * K, 0.0);
                        for (std::uint32_t g = 0; g < it.count; ++g)
                            for (std::uint32_t s = 0; s < it.nseg; ++s)
                                for (std::uint64_t p = 0; p < K; ++p)
                                    z[g * K + p] += zp.at((std::uint64_t(it.zrow) +
                                                           std::uint64_t(s) * it.zstride + g) * K + p);
                        check(std::uint64_t(it.zrow) + std::uint64_t(it.nseg - 1) * it.zstride +
                                  it.count <= pl.z_rows, "left item reads past its partial sums");
                        left_rows(it, 0, R, z, acc);
                    } else {
                        throw std::out_of_range("kRight item in a tile's run");
                    }
                }
                for (std::size_t i = 0; i < total.size(); ++i) total[i] += acc[i];
            }
            for (std::uint32_t i = 0; i < R; ++i)
                for (std::uint64_t p = 0; p < K; ++p) {
                    if (pass == 0)
                        zp.at((pl.z_rows + T.row0 + i) * K + p) = total[i * K + p];
                    else
                        y[std::uint64_t(at(pl.perm, T.row0 + i)) * K + p] = total[i * K + p];
                }
        }
}

// Virtual rows of components [comp0, comp0 + count) in the order the wide
// kernels use (build_vrows in wide_passes.hip): FP64, one per component in
// order; FP32 split, the components wider than 24 bits first, each as a
// high then a low row, then the others in order.  Entries: component j and
// part (0 whole, 1 high, 2 low).
std::vector<std::pair<std::uint32_t, int>> vrow_list(const std::vector<std::int32_t>& widths,
                                                     std::uint32_t comp0, std::uint32_t count,
                                                     bool fp32) {
    std::vector<std::pair<std::uint32_t, int>> out;
    if (fp32)
        for (std::uint32_t j = 0; j < count; ++j)
            if (widths.at(comp0 + j) > 24) {
                out.push_back({j, 1});
                out.push_back({j, 2});
            }
    for (std::uint32_t j = 0; j < count; ++j)
        if (!fp32 || widths.at(comp0 + j) <= 24) out.push_back({j, 0});
    return out;
}

// The wide passes' tables checked on the host: replays pass R, the
// partial-sum reductions, pass L and the row-set sums with the kernels'
// index arithmetic for one precision's virtual-row order, every access
// bounds-checked, and the structure they rely on (each item in exactly one
// run, a pass-R workgroup's waves on one column tile, codes inside their
// side block, partial sums written before they are read).  y = H x in FP64.
void emulate_wide(const Plan& pl, const double* x, std::int64_t k, double* y, bool fp32) {
    auto check = [](bool ok, const char* what) {
        if (!ok) throw std::out_of_range(what);
    };
    const std::uint64_t n = pl.order, K = std::uint64_t(k);
    check(pl.wruns_r.size() % kWideWaves == 0, "pass R runs are not whole workgroups");
    std::vector<int> seen(pl.witems.size(), 0);
    std::vector<double> z(pl.wz_rows * K, 0.0);
    std::vector<char> written(pl.wz_rows, 0);
    auto put = [&](std::uint64_t row, std::uint64_t p, double v) {
        check(row < pl.wz_rows, "partial-sum row out of range");
        z[row * K + p] = v;
        written[row] = 1;
    };
    // Code of lane L ending at bit E (width b) of an item's side block, as
    // the kernels read it (two words, the one past the block reads as zero).
    auto code_at = [&](const ItemDev& it, std::uint32_t L, std::uint32_t E, std::uint32_t b,
                       bool swz) {
        const std::uint64_t nw = it.quads / 16u;
        check(E >= b && E <= 32 * nw, "code outside its side block");
        const std::uint64_t kk = E >> 5, sh = E & 31;
        auto word = [&](std::uint64_t w) -> std::uint64_t {
            if (w >= nw) return 0;  // the kernels read a zero or slack row, shifted out
            return at(pl.words, 4ull * it.quad + 64 * w + lane_slot(swz, w, L));
        };
        const std::uint64_t pair = (word(kk) << 32) | (kk ? word(kk - 1) : 0);
        const std::uint64_t field = (pair >> sh) & 0xFFFFFFFFull;  // the 32 bits ending at E
        const std::int32_t v = std::int32_t(std::uint32_t(field));
        return double(v >> (32 - b));
    };
    auto item_ok = [&](const ItemDev& it) {
        check(it.quads <= kItemQuads && it.quads % 16 == 0, "item size");
        check(4ull * (std::uint64_t(it.quad) + it.quads) <= pl.words.size(), "item data range");
        check(it.count >= 1 && it.count <= kItemComponents, "item component count");
        check(std::uint64_t(it.comp0) + it.count <= pl.widths.size(), "item components");
    };
    auto end_bits = [&](const ItemDev& it) {
        std::vector<std::uint32_t> E(it.count);
        std::uint32_t e = 0;
        for (std::uint32_t j = 0; j < it.count; ++j) E[j] = e += std::uint32_t(pl.widths[it.comp0 + j]);
        return E;
    };
    // pass R
    for (std::size_t g = 0; g < pl.wruns_r.size() / kWideWaves; ++g) {
        const std::uint32_t tile = pl.wruns_r[g * kWideWaves].tile;
        const Span& span = at(pl.ctiles, tile);
        check(span.count >= 1 && span.count <= std::uint32_t(kTile) && span.first + span.count <= n,
              "column tile");
        for (int w = 0; w < kWideWaves; ++w) {
            const WideRun& run = pl.wruns_r[g * kWideWaves + w];
            check(run.tile == tile, "a pass-R workgroup spans two column tiles");
            check(run.begin <= run.end && run.end <= pl.witems.size(), "pass R run range");
            for (std::uint32_t i = run.begin; i < run.end; ++i) {
                const ItemDev& it = pl.witems[i];
                ++seen[i];
                item_ok(it);
                check(it.kind == kRight && it.col == span.first && it.cols <= span.count &&
                          it.nseg <= 1,
                      "pass R item is not a right side block of its column tile");
                const auto E = end_bits(it);
                std::vector<double> zs(std::uint64_t(it.count) * K);
                for (std::uint32_t j = 0; j < it.count; ++j) {
                    const std::uint32_t b = std::uint32_t(pl.widths[it.comp0 + j]);
                    const double sc = pl.cscale[it.comp0 + j];
                    for (std::uint64_t p = 0; p < K; ++p) {
                        double sum = 0.0;
                        for (std::uint32_t c = 0; c < span.count; ++c) {
                            const double cv = c < it.cols ? code_at(it, c, E[j], b, true) : 0.0;
                            sum += cv * x[std::uint64_t(at(pl.perm, span.first + c)) * K + p];
                        }
                        zs[j * K + p] = sc * sum;
                    }
                }
                if (it.nseg == 0) {  // per component
                    for (std::uint32_t j = 0; j < it.count; ++j)
                        for (std::uint64_t p = 0; p < K; ++p)
                            put(std::uint64_t(it.zrow) + j, p, zs[j * K + p]);
                } else {  // in virtual-row order
                    const auto vl = vrow_list(pl.widths, it.comp0, it.count, fp32);
                    for (std::size_t v = 0; v < vl.size(); ++v)
                        for (std::uint64_t p = 0; p < K; ++p)
                            put(std::uint64_t(it.zrow) + v, p, zs[vl[v].first * K + p]);
                }
            }
        }
    }
    // partial sums over a block's column tiles, into rows in virtual-row order
    for (const ZSum& zr : pl.wzsums) {
        check(zr.nseg >= 1 && zr.count >= 1, "partial-sum reduction shape");
        const auto vl = vrow_list(pl.widths, zr.comp0, zr.count, fp32);
        for (std::size_t v = 0; v < vl.size(); ++v)
            for (std::uint64_t p = 0; p < K; ++p) {
                double s = 0.0;
                for (std::uint32_t sg = 0; sg < zr.nseg; ++sg) {
                    const std::uint64_t row =
                        std::uint64_t(zr.zrow) + std::uint64_t(sg) * zr.zstride + vl[v].first;
                    check(row < pl.z_rows && written[row], "reduction reads an unwritten partial sum");
                    s += z[row * K + p];
                }
                check(std::uint64_t(zr.vzrow) + v >= pl.z_rows, "reduction overwrites partial sums");
                put(std::uint64_t(zr.vzrow) + v, p, s);
            }
    }
    // pass L: workgroups of kWideWaves wave runs on one row tile and set
    check(pl.wruns_l.size() % kWideWaves == 0, "pass L runs are not whole workgroups");
    std::vector<double> sets(std::uint64_t(pl.wsets) * kTile * K, 0.0);
    std::vector<int> tile_runs(pl.tiles.size(), 0), direct(pl.tiles.size(), 0);
    for (std::size_t g = 0; g < pl.wruns_l.size() / kWideWaves; ++g) {
      const WideRun& run = pl.wruns_l[g * kWideWaves];
      const TileDev& T = at(pl.tiles, run.tile);
      const std::uint32_t R = T.rows;
      check(R >= 1 && R <= std::uint32_t(kTile), "tile rows");
      ++tile_runs[run.tile];
      std::vector<double> acc(std::uint64_t(kTile) * K, 0.0);
      for (int wv = 0; wv < kWideWaves; ++wv) {
        const WideRun& wr = pl.wruns_l[g * kWideWaves + wv];
        check(wr.tile == run.tile && wr.set == run.set, "a pass-L workgroup spans two runs");
        // the kernel walks a workgroup's wave runs as one range
        check(wv == 0 || wr.begin == pl.wruns_l[g * kWideWaves + wv - 1].end,
              "the wave runs of a pass-L workgroup are not one range");
        check(wr.begin <= wr.end && wr.end <= pl.witems.size(), "pass L run range");
        for (std::uint32_t i = wr.begin; i < wr.end; ++i) {
            const ItemDev& it = pl.witems[i];
            ++seen[i];
            if (it.kind == kDense) {
                check(it.count >= 1 && it.count <= kDenseColumns, "dense item columns");
                check(std::uint64_t(it.col) + it.count <= n, "dense item column range");
                check(4ull * it.quad + 2ull * it.count * R <= pl.words.size(), "dense item data");
                for (std::uint32_t c = 0; c < it.count; ++c)
                    for (std::uint32_t r = 0; r < R; ++r) {
                        const std::uint64_t wv = 4ull * it.quad + 2ull * (std::uint64_t(c) * R + r);
                        const std::uint64_t bitsv = std::uint64_t(pl.words[wv]) |
                                                    (std::uint64_t(pl.words[wv + 1]) << 32);
                        double v;
                        std::memcpy(&v, &bitsv, 8);
                        for (std::uint64_t p = 0; p < K; ++p)
                            acc[r * K + p] += v * x[std::uint64_t(at(pl.perm, it.col + c)) * K + p];
                    }
                continue;
            }
            item_ok(it);
            check(it.kind == kLeft && it.nseg == 1, "pass L item kind");
            check(it.quads / 16u <= 16u, "left side block rows");
            const auto E = end_bits(it);
            const auto vl = vrow_list(pl.widths, it.comp0, it.count, fp32);
            check(vl.size() <= it.zstride && it.zstride <= 2u * it.count, "virtual rows of the item");
            check(std::uint64_t(it.zrow) + it.zstride <= pl.wz_rows, "partial-sum rows of the item");
            for (std::size_t v = 0; v < vl.size(); ++v) {
                const std::uint32_t j = vl[v].first;
                const std::uint32_t b = std::uint32_t(pl.widths[it.comp0 + j]);
                const std::uint64_t row = std::uint64_t(it.zrow) + v;
                check(written[row], "pass L reads an unwritten partial sum");
                for (std::uint32_t r = 0; r < R; ++r) {
                    // the part of the code this row carries, the high part with its 2^16
                    const double cv = code_at(it, r, E[j], b, false);
                    double a = cv;
                    if (vl[v].second != 0) {
                        const std::int64_t c = std::int64_t(cv);
                        const std::int64_t lo = std::int64_t(std::int16_t(std::uint16_t(c & 0xFFFF)));
                        a = vl[v].second == 2 ? double(lo) : double(c - lo);
                    }
                    for (std::uint64_t p = 0; p < K; ++p) acc[r * K + p] += a * z[row * K + p];
                }
            }
        }
      }
        for (std::uint32_t r = 0; r < R; ++r)
            for (std::uint64_t p = 0; p < K; ++p) {
                if (run.set == kNoSet) {
                    ++direct[run.tile];
                    y[std::uint64_t(at(pl.perm, T.row0 + r)) * K + p] = acc[r * K + p];
                } else {
                    check(run.set < pl.wsets, "row set index");
                    sets[(std::uint64_t(run.set) * kTile + r) * K + p] = acc[r * K + p];
                }
            }
    }
    for (std::size_t t = 0; t < pl.tiles.size(); ++t)
        check(tile_runs[t] >= 1 && (direct[t] == 0 || tile_runs[t] == 1),
              "row tile without runs, or several runs writing y");
    for (const WideSum& ws : pl.wsums) {
        const TileDev& T = at(pl.tiles, ws.tile);
        check(direct[ws.tile] == 0 && std::uint64_t(ws.set0) + ws.count <= pl.wsets &&
                  int(ws.count) == tile_runs[ws.tile],
              "row-set sum");
        for (std::uint32_t r = 0; r < T.rows; ++r)
            for (std::uint64_t p = 0; p < K; ++p) {
                double v = 0.0;
                for (std::uint32_t s = 0; s < ws.count; ++s)
                    v += sets[(std::uint64_t(ws.set0 + s) * kTile + r) * K + p];
                y[std::uint64_t(at(pl.perm, T.row0 + r)) * K + p] = v;
            }
    }
    for (int c : seen) check(c == 1, "a wide item is not in exactly one run");
}

void destroy_device_store(DeviceStore* s) {
    if (!s) return;
    if (s->pins.load() > 0)
        throw std::invalid_argument(
            "store is in use by a batch, factor plan or server: release that first");
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(s->device);
    for (void* e : s->profile.events) (void)hipEventDestroy(static_cast<hipEvent_t>(e));
    for (void* p : {static_cast<void*>(s->perm), static_cast<void*>(s->words),
                    static_cast<void*>(s->widths), static_cast<void*>(s->cscale),
                    static_cast<void*>(s->items), static_cast<void*>(s->tiles),
                    static_cast<void*>(s->heads),
                    static_cast<void*>(s->witems), static_cast<void*>(s->wruns_r),
                    static_cast<void*>(s->wruns_l), static_cast<void*>(s->wsums),
                    static_cast<void*>(s->wzsums),
                    static_cast<void*>(s->ctiles), static_cast<void*>(s->rec.tasks), static_cast<void*>(s->rec.chunks),
                    static_cast<void*>(s->rec.lscale), static_cast<void*>(s->rec.rscale)})
        if (p) (void)hipFree(p);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete s;
}

namespace {
// Columns per launch chain of the narrow passes: all k, or the most (a
// multiple of kb) whose workspace stays addressable by the kernels' 32-bit
// buffer offsets.
std::int64_t narrow_columns(const DeviceStore& s, std::int64_t k, int kb) {
    if (narrow_workspace(s, k) < kOutOfRange) return k;
    const std::uint64_t per = narrow_workspace(s, 1);
    const std::int64_t cols = std::int64_t((kOutOfRange - 1) / per) / kb * kb;
    require(cols >= kb, "store too large for the narrow passes (use 16 or more right-hand sides per pass)");
    return cols;
}
}  // namespace

std::uint64_t apply_workspace_bytes(const DeviceStore& s, std::int64_t k) {
    // enough for either path, whichever the right-hand-side block selects
    return std::max(narrow_workspace(s, narrow_columns(s, k, 4)), wide_workspace(s, k));
}

void apply(const DeviceStore& s, const double* x, std::int64_t ldx, std::int64_t k, double* y,
           std::int64_t ldy, int precision, void* stream, void* workspace,
           std::uint64_t workspace_bytes) {
    require(k > 0 && k <= std::int64_t(std::numeric_limits<int>::max()),
            "right-hand-side shape mismatch");
    require(ldx >= k && ldy >= k, "leading dimension smaller than the right-hand-side count");
    require(x != nullptr && y != nullptr, "null right-hand side or result pointer");
    require(precision == QH_PRECISION_FP64 || precision == QH_PRECISION_FP32SPLIT,
            "unknown precision mode");
    const bool fp32 = precision == QH_PRECISION_FP32SPLIT;
    const int kb = rhs_block(s, k);
    const bool wide = is_wide(kb);
    // the narrow passes run in column chunks whose workspace and x rows stay
    // below their 32-bit buffer offsets; the wide passes address 64-bit
    const std::int64_t cols = wide ? k : narrow_columns(s, k, kb);
    const std::uint64_t need = wide ? wide_workspace(s, k) : narrow_workspace(s, cols);
    require(workspace == nullptr || workspace_bytes >= need, "workspace is too small");
    require(wide || ((s.order - 1) * std::uint64_t(ldx) + std::uint64_t(cols)) * sizeof(double) <
                        kOutOfRange,
            "right-hand-side leading dimension too large for the narrow passes");
    DeviceScope scope(s.device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    double* zp = static_cast<double*>(workspace);
    bool owned = false;
    if (!zp && need) {
        hip_check(hipMallocAsync(reinterpret_cast<void**>(&zp), need, st), "hipMallocAsync");
        owned = true;
    }
    try {
        if (wide) {
            launch_wide(wide_view(s, x, ldx, int(k), y, ldy, zp), wide_counts(s), kb, fp32, st,
                        profile_slot(s));
        } else {
            for (std::int64_t c0 = 0; c0 < k; c0 += cols) {
                const int kc = int(std::min(cols, k - c0));
                if (fp32)
                    dispatch_narrow<true>(s, kb, x + c0, ldx, kc, y + c0, ldy, zp, st);
                else
                    dispatch_narrow<false>(s, kb, x + c0, ldx, kc, y + c0, ldy, zp, st);
            }
        }
    } catch (...) {
        if (owned) (void)hipFreeAsync(zp, st);
        throw;
    }
    if (owned) hip_check(hipFreeAsync(zp, st), "hipFreeAsync");
}

// Workspace of qh_apply_noalloc: one array in every device's memory, made
// when the code object is loaded and held until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: its
// device address comes from hipGetSymbolAddress, once per device.
static __device__ __align__(256) std::uint8_t scratch[kScratchBytes];

namespace {
std::mutex scratch_lock;
std::map<int, void*> scratch_ptr;  // device ordinal -> address of that device's array

void* scratch_address(int device) {
    std::lock_guard<std::mutex> guard(scratch_lock);
    const auto it = scratch_ptr.find(device);
    if (it != scratch_ptr.end()) return it->second;
    DeviceScope scope(device);
    void* ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
    scratch_ptr[device] = ptr;
    return ptr;
}
}  // namespace

void apply_noalloc(const DeviceStore& s, const double* x, std::int64_t ldx, std::int64_t k,
                   double* y, std::int64_t ldy, int precision, void* stream) {
    require(k > 0 && k <= std::int64_t(std::numeric_limits<int>::max()),
            "right-hand-side shape mismatch");
    const std::uint64_t need = apply_workspace_bytes(s, k);
    if (need > kScratchBytes)
        throw std::invalid_argument("apply needs " + std::to_string(need) +
                                    " workspace bytes, more than the static scratch holds (" +
                                    std::to_string(kScratchBytes) + "); use qh_apply");
    void* ws = scratch_address(s.device);
    apply(s, x, ldx, k, y, ldy, precision, stream, ws, kScratchBytes);
}

void profile_begin(DeviceStore& s, std::uint64_t max_applies) {
    require(max_applies > 0 && max_applies < (std::uint64_t{1} << 20), "invalid profile size");
    DeviceScope scope(s.device);
    auto& p = s.profile;
    std::lock_guard<std::mutex> guard(p.lock);
    while (p.events.size() < 3 * max_applies) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        p.events.push_back(e);
    }
    p.used = 0;
    p.armed = true;
}

qh_profile profile_end(DeviceStore& s) {
    DeviceScope scope(s.device);
    auto& p = s.profile;
    std::lock_guard<std::mutex> guard(p.lock);
    qh_profile out{};
    out.applies = p.used;
    for (std::uint64_t a = 0; a < p.used; ++a) {
        auto* ev = reinterpret_cast<hipEvent_t*>(p.events.data() + 3 * a);
        hip_check(hipEventSynchronize(ev[2]), "hipEventSynchronize");
        float ms = 0.0f;
        hip_check(hipEventElapsedTime(&ms, ev[0], ev[1]), "hipEventElapsedTime");
        out.pass_ms[0] += ms;
        hip_check(hipEventElapsedTime(&ms, ev[1], ev[2]), "hipEventElapsedTime");
        out.pass_ms[1] += ms;
    }
    p.armed = false;
    p.used = 0;
    return out;
}

void decode_codes(const DeviceStore& s, std::uint64_t block, int side, std::int32_t* out,
                  void* stream) {
    require(block < s.block_to_lowrank.size(), "block index out of range");
    require(side == 0 || side == 1, "side must be 0 (left) or 1 (right)");
    require(out != nullptr, "null output pointer");
    const std::int64_t li = s.block_to_lowrank[block];
    if (li < 0) return;  // dense or rank 0: nothing to decode
    const auto& chunks = side == 0 ? s.left_chunks[li] : s.right_chunks[li];
    if (chunks.empty()) return;
    std::uint32_t rows = 0;
    for (const auto& c : chunks) rows = std::max(rows, c.first + c.count);
    DeviceScope scope(s.device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    ChunkRef* dev = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void**>(&dev), chunks.size() * sizeof(ChunkRef)),
              "hipMalloc");
    hipError_t e = hipMemcpy(dev, chunks.data(), chunks.size() * sizeof(ChunkRef),
                             hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(decode_side, dim3(unsigned(chunks.size())), dim3(64), 0, st, s.words,
                           dev, s.widths + s.comp_base[li], rows, side == 0 ? 0 : 1, out);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    (void)hipFree(dev);
    hip_check(e, "decode launch");
}

// ---------------------------------------------------------------------------
// Batches

void set_rhs_block(DeviceStore& s, std::int64_t k, int kb) {
    require(k > 0, "right-hand-side shape mismatch");
    require(kb == 0 || kb == 1 || kb == 2 || kb == 4 || kb == 16 || kb == 32 || kb == 64,
            "right-hand sides per pass must be 1, 2, 4, 16, 32 or 64 (0 = default)");
    std::unique_lock<std::shared_mutex> guard(s.rhs_lock);
    if (kb == 0) s.rhs_block.erase(k);
    else s.rhs_block[k] = kb;
}

Batch* create_batch(const DeviceStore* const* stores, const std::int64_t* k, std::uint64_t count) {
    require(count > 0 && stores && k, "empty batch");
    auto* b = new Batch;
    try {
        b->device = stores[0] ? stores[0]->device : 0;
        std::uint64_t off = 0, zoff = 0;
        std::vector<int> kbs;
        for (std::uint64_t i = 0; i < count; ++i) {
            require(stores[i] != nullptr, "null store in batch");
            require(stores[i]->device == b->device, "batch members must share one device");
            require(k[i] > 0 && k[i] <= std::int64_t(std::numeric_limits<int>::max()),
                    "right-hand-side shape mismatch");
            const int kb = rhs_block(*stores[i], k[i]);
            const std::uint64_t ws =
                is_wide(kb) ? wide_workspace(*stores[i], k[i]) : narrow_workspace(*stores[i], k[i]);
            // the narrow passes address a member's x and workspace with 32-bit offsets
            require(is_wide(kb) || (ws < kOutOfRange &&
                                    stores[i]->order * std::uint64_t(k[i]) * sizeof(double) < kOutOfRange),
                    "batch member too large for the narrow passes (use 16 or more right-hand "
                    "sides per pass)");
            kbs.push_back(kb);
            b->stores.push_back(stores[i]);
            b->k.push_back(k[i]);
            b->offset.push_back(off);
            b->zoffset.push_back(zoff);
            off += stores[i]->order * std::uint64_t(k[i]);
            zoff += ws / sizeof(double);
        }
        b->total = off;
        b->z_total = zoff;
        for (const DeviceStore* s : b->stores) s->pins.fetch_add(1);
        DeviceScope scope(b->device);
        for (int kb : {1, 2, 4}) {
            BatchClass c;
            c.kb = kb;
            std::vector<MemberView> views;
            std::vector<std::uint32_t> ft{0};
            for (std::uint64_t i = 0; i < count; ++i) {
                if (kbs[i] != kb) continue;
                const DeviceStore& s = *b->stores[i];
                // x, y and the workspace as element offsets from the call's bases
                views.push_back(member_view(
                    s, reinterpret_cast<const double*>(std::uintptr_t(b->offset[i])), b->k[i],
                    int(b->k[i]), reinterpret_cast<double*>(std::uintptr_t(b->offset[i])),
                    b->k[i], reinterpret_cast<double*>(std::uintptr_t(b->zoffset[i]))));
                require(std::uint64_t(ft.back()) + s.n_tiles < (std::uint64_t{1} << 31),
                        "batch too large for one launch");
                ft.push_back(ft.back() + s.n_tiles);
                c.max_k = std::max(c.max_k, b->k[i]);
            }
            if (views.empty()) continue;
            c.member_count = views.size();
            c.total_tiles = ft.back();
            std::uint64_t bytes = 0;
            c.members = upload(views, bytes);
            c.first_tile = upload(ft, bytes);
            b->classes.push_back(c);
        }
        for (int kc : {16, 32, 64}) {
            WideBatchClass c;
            c.kc = kc;
            std::vector<WideView> views;
            std::vector<std::uint32_t> fr{0}, fl{0}, fs{0}, fz{0};
            for (std::uint64_t i = 0; i < count; ++i) {
                if (kbs[i] != kc) continue;
                const DeviceStore& s = *b->stores[i];
                // x and y as offsets in doubles, the workspace in bytes
                views.push_back(wide_view(
                    s, reinterpret_cast<const double*>(std::uintptr_t(b->offset[i])), b->k[i],
                    int(b->k[i]), reinterpret_cast<double*>(std::uintptr_t(b->offset[i])),
                    b->k[i], reinterpret_cast<void*>(std::uintptr_t(b->zoffset[i] * sizeof(double)))));
                require(std::uint64_t(fl.back()) + s.n_wl_groups + s.n_wr_groups + s.n_wsums <
                            (std::uint64_t{1} << 31),
                        "batch too large for one launch");
                fr.push_back(fr.back() + s.n_wr_groups);
                fl.push_back(fl.back() + s.n_wl_groups);
                fs.push_back(fs.back() + s.n_wsums);
                fz.push_back(fz.back() + s.n_wzsums);
                c.max_k = std::max(c.max_k, b->k[i]);
            }
            if (views.empty()) continue;
            c.member_count = views.size();
            c.right_groups = fr.back();
            c.left_groups = fl.back();
            c.sums = fs.back();
            c.zsums = fz.back();
            std::uint64_t bytes = 0;
            c.members = upload(views, bytes);
            c.first_right = upload(fr, bytes);
            c.first_left = upload(fl, bytes);
            c.first_sum = upload(fs, bytes);
            c.first_zsum = upload(fz, bytes);
            b->wclasses.push_back(c);
        }
    } catch (...) {
        destroy_batch(b);
        throw;
    }
    return b;
}

void destroy_batch(Batch* b) {
    if (!b) return;
    if (b->pins.load() > 0)
        throw std::invalid_argument("batch is in use by a graph: destroy the graph first");
    for (const DeviceStore* s : b->stores) s->pins.fetch_sub(1);
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    for (auto& c : b->classes)
        for (void* p : {static_cast<void*>(c.members), static_cast<void*>(c.first_tile)})
            if (p) (void)hipFree(p);
    for (auto& c : b->wclasses)
        for (void* p : {static_cast<void*>(c.members), static_cast<void*>(c.first_right),
                        static_cast<void*>(c.first_left), static_cast<void*>(c.first_sum),
                        static_cast<void*>(c.first_zsum)})
            if (p) (void)hipFree(p);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete b;
}

void batch_apply(const Batch& b, const double* x, double* y, int precision, void* stream,
                 void* workspace, std::uint64_t workspace_bytes) {
    require(x != nullptr && y != nullptr, "null right-hand side or result pointer");
    require(precision == QH_PRECISION_FP64 || precision == QH_PRECISION_FP32SPLIT,
            "unknown precision mode");
    const std::uint64_t need = b.z_total * sizeof(double);
    require(workspace == nullptr || workspace_bytes >= need, "workspace is too small");
    DeviceScope scope(b.device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    double* zp = static_cast<double*>(workspace);
    bool owned = false;
    if (!zp && need) {
        hip_check(hipMallocAsync(reinterpret_cast<void**>(&zp), need, st), "hipMallocAsync");
        owned = true;
    }
    const bool fp32 = precision == QH_PRECISION_FP32SPLIT;
    try {
        for (const auto& c : b.classes) {
            if (fp32)
                dispatch_batch_class<true>(c, x, y, zp, st);
            else
                dispatch_batch_class<false>(c, x, y, zp, st);
        }
        for (const auto& c : b.wclasses) launch_batch_wide(c, x, y, zp, fp32, st);
    } catch (...) {
        if (owned) (void)hipFreeAsync(zp, st);
        throw;
    }
    if (owned) hip_check(hipFreeAsync(zp, st), "hipFreeAsync");
}

void* batch_graph(const Batch& b, const double* x, double* y, int precision, void* workspace,
                  std::uint64_t workspace_bytes) {
    require(workspace != nullptr || b.z_total == 0, "graph capture needs a workspace");
    DeviceScope scope(b.device);
    hipStream_t st = nullptr;
    hip_check(hipStreamCreate(&st), "hipStreamCreate");
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        try {
            batch_apply(b, x, y, precision, st, workspace, workspace_bytes);
        } catch (...) {
            (void)hipStreamEndCapture(st, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipStreamDestroy(st);
            throw;
        }
        e = hipStreamEndCapture(st, &graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipStreamDestroy(st);
    hip_check(e, "graph capture");
    return exec;
}

void graph_launch(void* exec, void* stream) {
    require(exec != nullptr, "null graph");
    hip_check(hipGraphLaunch(static_cast<hipGraphExec_t>(exec), static_cast<hipStream_t>(stream)),
              "hipGraphLaunch");
}

void graph_destroy(void* exec) {
    if (exec) (void)hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec));
}

}  // namespace qhamd
