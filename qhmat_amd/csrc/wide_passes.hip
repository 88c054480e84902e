// Wide passes of the MI355X apply: 16, 32 or 64 right-hand sides per pass on
// the matrix cores (v_mfma_f32_16x16x4f32 in FP32-split mode,
// v_mfma_f64_16x16x4f64 in FP64 mode and for the diagonal blocks).  Layout
// and pass structure in device_store.hpp ("wide passes").
//
// Each code is read from HBM once per apply and decoded once per pass: a
// decoded code is the A operand of one MFMA per 16 right-hand sides.  x of a
// column tile is read once per pass-R workgroup into LDS, and the partial
// sums z of a block's column tile make one round trip through the workspace.
//
// Split precision.  In FP32 mode a component wider than 24 bits is split
// into two exact FP32 halves (high: the code's upper part, low: its
// sign-extended low 16 bits) that take two "virtual rows" of the item,
// interleaved (high at an even row, low after it); other components take
// one.  Pass R joins the two halves' sums as 2^16 high + low in FP64; pass L
// multiplies the high half by 2^16 zs.  Only the wide components cost a
// second MFMA slot.
//
// Determinism: a column's result depends only on its own inputs, on the
// plan (runs, item order) and on the path; every sum has a fixed order and
// one writer, and no floating-point atomics are used.
//
// Reference: src/qhmat/core.cpp:360-390 (the FP64 loop nest these passes
// reorder) and src/qhmat/cuda.cu:64-164 (the reference's device path).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

#include "device_common.hpp"
#include "device_store.hpp"
#include "wide_passes.hpp"

namespace qhamd {
namespace {

// Per-virtual-row fields packed in one word: end bit E of the component's
// code in every lane's stream (bits 16-26), component index in the item
// (8-15), part (6-7) and width b (0-5).  The dummy row (E = 32, b = 1) decodes
// the top bit of a lane's first word; its products are never used.
enum Part : std::uint32_t { kWhole = 0, kHigh = 1, kLow = 2 };
constexpr std::uint32_t kDummyRow = (32u << 16) | 1u;

__device__ __forceinline__ std::uint32_t vrow_pack(std::uint32_t E, std::uint32_t j,
                                                   std::uint32_t part, std::uint32_t b) {
    return (E << 16) | (j << 8) | (part << 6) | b;
}
__device__ __forceinline__ std::uint32_t vrow_end(std::uint32_t m) { return m >> 16; }
__device__ __forceinline__ std::uint32_t vrow_comp(std::uint32_t m) { return (m >> 8) & 0xFFu; }
__device__ __forceinline__ std::uint32_t vrow_part(std::uint32_t m) { return (m >> 6) & 3u; }
__device__ __forceinline__ std::uint32_t vrow_bits(std::uint32_t m) { return m & 63u; }

// Set bits of `mask` in the lanes below this one.
__device__ __forceinline__ std::uint32_t lanes_below(std::uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi(std::uint32_t(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(std::uint32_t(mask), 0u));
}

// Virtual rows of an item (lanes = its components j < count, `bits` their
// widths): the end bits are the inclusive prefix sum of the widths in item
// order.  FP64 mode: row j is component j.  FP32 mode: the components wider
// than 24 bits first, each as a high and a low row, then the others in
// order.  Writes vrow[0 .. V) and returns V.
template <bool FP32>
__device__ __forceinline__ std::uint32_t build_vrows(std::uint32_t bits, std::uint32_t count,
                                                     int lane, std::uint32_t* vrow) {
    const bool comp = std::uint32_t(lane) < count;
    const std::uint32_t E = wave_scan(comp ? bits : 0u, lane);
    if constexpr (!FP32) {
        if (comp) vrow[lane] = vrow_pack(E, std::uint32_t(lane), kWhole, bits);
        return count;
    } else {
        const bool wide = comp && bits > 24u;
        const std::uint64_t wmask = __ballot(wide);
        const std::uint64_t nmask = __ballot(comp && !wide);
        const std::uint32_t W = std::uint32_t(__builtin_popcountll(wmask));
        if (wide) {
            const std::uint32_t v = 2u * lanes_below(wmask);
            vrow[v] = vrow_pack(E, std::uint32_t(lane), kHigh, bits);
            vrow[v + 1] = vrow_pack(E, std::uint32_t(lane), kLow, bits);
        } else if (comp) {
            vrow[2u * W + lanes_below(nmask)] = vrow_pack(E, std::uint32_t(lane), kWhole, bits);
        }
        return count + W;
    }
}

// Virtual row v of an item with nv rows, read without a branch (the read is
// unconditional, at an index kept inside the 128-row table; the dummy row
// is selected past nv).
__device__ __forceinline__ std::uint32_t vrow_at(const std::uint32_t* vrow, std::uint32_t v,
                                                 std::uint32_t nv) {
    const std::uint32_t m = vrow[min(v, 127u)];
    return v < nv ? m : kDummyRow;
}

// The part of a code a virtual row carries: the whole code, or its split
// into h 2^16 + l with l the sign-extended low 16 bits (both exact in FP32;
// a code of small magnitude stays in l).
__device__ __forceinline__ int code_part(int code, std::uint32_t part) {
    const int lo = __builtin_amdgcn_sbfe(code, 0u, 16u);
    const int hi = (code >> 16) + ((code >> 15) & 1);
    return part == kLow ? lo : (part == kHigh ? hi : code);
}

// Codes of columns 16 q .. 16 q + 15 of the virtual row with packed fields m
// in a staged right side block (word w of column c at side[64 w + (c xor
// 4 (w mod 16))]): two aligned four-word reads per four columns.  The head
// row below the block and the slack row past it are readable.
// SPLIT: the rows may carry split parts; otherwise every row is whole.
template <bool SPLIT>
__device__ __forceinline__ void decode_right16(const std::uint32_t* side, std::uint32_t m, int q,
                                               int (&code)[16]) {
    const std::uint32_t E = vrow_end(m);
    const int kk = int(E >> 5);
    const std::uint32_t sh = 32u - vrow_bits(m), part = vrow_part(m);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int c4 = 16 * q + 4 * u;
        const quad_t hi = *reinterpret_cast<const quad_t*>(side + 64 * kk + (c4 ^ (4 * (kk & 15))));
        const quad_t lo =
            *reinterpret_cast<const quad_t*>(side + 64 * (kk - 1) + (c4 ^ (4 * ((kk - 1) & 15))));
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int c = static_cast<int>(__builtin_amdgcn_alignbit(hi[i], lo[i], E)) >> sh;
            code[4 * u + i] = SPLIT ? code_part(c, part) : c;
        }
    }
}

template <int KC, bool FP32> struct RightShape {
    using Real = real_t<FP32>;
    static constexpr int NS = KC / 16;                  // 16-wide slices of right-hand sides
    static constexpr int XS = FP32 ? 68 : 66;           // row stride of the x tile (elements)
    static constexpr int kXBytes = KC * XS * int(sizeof(Real));
    static constexpr int kVrow = kLdsHead + kStageBytes;  // per wave: head, stage, rows, scales
    static constexpr int kScale = kVrow + 128 * 4;
    static constexpr int kWaveBytes = kScale + 64 * 8;
    static constexpr int kBytes = kXBytes + kWideWaves * kWaveBytes;
};

// Registers of the item a wave fetches next.
struct RightFetch {
    quad_t q[kItemQuads / 64];
    std::uint32_t bits;  // lanes = components
    double scale;        // lanes = components
};

__device__ __forceinline__ void right_fetch(const WideView& V, const ItemDev& f, int lane,
                                            RightFetch& F) {
    const __amdgpu_buffer_rsrc_t data = buffer(V.words + 4ull * f.quad, std::uint32_t(f.quads) * 16u);
#pragma unroll
    for (int t = 0; t < kItemQuads / 64; ++t)
        F.q[t] = __builtin_amdgcn_raw_buffer_load_b128(data, std::uint32_t(lane + 64 * t) * 16u, 0, 0);
    F.bits = __builtin_amdgcn_raw_buffer_load_b32(buffer(V.widths + f.comp0, f.count * 4u),
                                                  std::uint32_t(lane) * 4u, 0, 0);
    F.scale = load_f64(buffer(V.cscale + f.comp0, f.count * 8u), std::uint32_t(lane) * 8u);
}

// Item record i as eight wave-uniform dwords (so that buffer descriptors
// built from it stay in scalar registers), sizes clamped to the stage.
__device__ __forceinline__ ItemDev load_item(const ItemDev* items, std::uint32_t i) {
    const std::uint32_t* w = reinterpret_cast<const std::uint32_t*>(items + i);
    std::uint32_t d[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) d[t] = __builtin_amdgcn_readfirstlane(w[t]);
    ItemDev f;
    f.quad = d[0];
    f.quads = std::uint16_t(min(d[1] & 0xFFFFu, std::uint32_t(kItemQuads)));
    f.kind = std::uint8_t((d[1] >> 16) & 0xFFu);
    f.count = std::uint8_t(min(d[1] >> 24, std::uint32_t(kItemComponents)));
    f.comp0 = d[2];
    f.col = d[3];
    f.cols = d[4];
    f.zrow = d[5];
    f.zstride = d[6];
    f.nseg = d[7];
    return f;
}

// Item i of a run ending at `end`, or an empty item past it.
__device__ __forceinline__ ItemDev item_or_empty(const WideView& V, std::uint32_t i,
                                                 std::uint32_t end) {
    if (i < end) return load_item(V.items, i);
    return ItemDev{};
}

// Pass R, workgroup g (kWideWaves waves on one column tile), right-hand sides
// [p0, p0 + KC).
template <int KC, bool FP32>
__device__ __forceinline__ void right_group(const WideView& V, std::uint32_t g, int p0,
                                            std::uint8_t* lds) {
    using S = RightShape<KC, FP32>;
    using Real = typename S::Real;
    using Mm = Mma<FP32>;
    const int tid = int(threadIdx.x);
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63, q = lane >> 4, r = lane & 15;
    const std::uint32_t tile = V.rruns[g * kWideWaves].tile;
    const Span span = V.ctiles[tile];
    const std::uint32_t ncols = min(span.count, std::uint32_t(kTile));
    const int k = V.k;

    // x of the tile's columns, transposed: xs[p][c] (zero past the tile's
    // columns and past k); every load is issued before the first is used
    Real* xs = reinterpret_cast<Real*>(lds);
    {
        constexpr int kPer = 64 * KC / (64 * kWideWaves);
        std::int64_t row[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int c = (tid + 64 * kWideWaves * i) / KC;
            row[i] = std::uint32_t(c) < ncols
                         ? std::int64_t(V.perm[span.first + std::uint32_t(c)]) * V.ldx : -1;
        }
        double v[kPer];
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            // an unconditional load at a valid address (row 0, column k - 1
            // stand in), zero selected where there is no element
            const int p = (tid + 64 * kWideWaves * i) % KC;
            const bool ok = row[i] >= 0 && p0 + p < k;
            const double got = V.x[(row[i] >= 0 ? row[i] : 0) + min(p0 + p, k - 1)];
            v[i] = ok ? got : 0.0;
        }
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int e = tid + 64 * kWideWaves * i;
            xs[(e % KC) * S::XS + e / KC] = Real(v[i]);
        }
    }
    __syncthreads();

    std::uint8_t* wl = lds + S::kXBytes + wave * S::kWaveBytes;
    quad_t* stage_q = reinterpret_cast<quad_t*>(wl + kLdsHead);
    const std::uint32_t* stage = reinterpret_cast<const std::uint32_t*>(wl + kLdsHead);
    std::uint32_t* vrow = reinterpret_cast<std::uint32_t*>(wl + S::kVrow);
    double* scl = reinterpret_cast<double*>(wl + S::kScale);
    Real* zout = static_cast<Real*>(V.zw);

    const WideRun run = V.rruns[g * kWideWaves + wave];
    const std::uint32_t begin = __builtin_amdgcn_readfirstlane(run.begin);
    const std::uint32_t end = __builtin_amdgcn_readfirstlane(run.end);
    ItemDev f = item_or_empty(V, begin, end);
    RightFetch F;
    right_fetch(V, f, lane, F);
    for (std::uint32_t i = begin; i < end; ++i) {
        // issue priority by progress in FP32-split mode (sphere k=64: 359 -> 349
        // us per pass R); FP64 pass R was slower with it (506 -> 514 us)
        if constexpr (FP32) progress_priority(i - begin, end - begin);
        // stage item i, then fetch item i + 1 while it is decoded
#pragma unroll
        for (int t = 0; t < kItemQuads / 64; ++t)
            if (std::uint32_t(lane + 64 * t) < f.quads) stage_q[lane + 64 * t] = F.q[t];
        scl[lane] = F.scale;
        const std::uint32_t bits = F.bits;
        const ItemDev cur = f;
        f = item_or_empty(V, i + 1, end);
        right_fetch(V, f, lane, F);
        wave_lds_sync();
        const std::uint32_t nv = build_vrows<FP32>(bits, cur.count, lane, vrow);
        wave_lds_sync();

        // rows 2 W .. nv - 1 are whole components: their groups skip the
        // split handling
        const std::uint32_t split_rows = 2u * (nv - cur.count);
        for (std::uint32_t g0 = 0; g0 < nv; g0 += 16) {
            const std::uint32_t va = g0 + std::uint32_t(a_row<FP32>(r));
            const std::uint32_t m = vrow_at(vrow, va, nv);
            int code[16];
            if (g0 < split_rows)
                decode_right16<true>(stage, m, q, code);
            else
                decode_right16<false>(stage, m, q, code);
            Real a[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) a[j] = Real(code[j]);
            // this lane's output rows (virtual rows g0 + 4 q + t): part, the
            // component's scale and the element offset of right-hand side 0
            std::uint32_t part[4];
            double sc[4];
            std::uint64_t at[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const std::uint32_t v = g0 + 4u * std::uint32_t(q) + std::uint32_t(t);
                const std::uint32_t mv = vrow[min(v, 127u)];
                part[t] = v < nv ? vrow_part(mv) : kLow;  // past the rows: nothing to write
                const std::uint32_t j = vrow_comp(mv);
                sc[t] = scl[j];
                // per component, or in virtual-row order (a split
                // component's two rows both carry its sum)
                at[t] = (std::uint64_t(cur.zrow) + (cur.nseg ? v : j)) * std::uint64_t(k) + p0 + r;
            }
#pragma unroll 2
            for (int n = 0; n < S::NS; ++n) {
                const Real* xr = xs + (16 * n + r) * S::XS + 16 * q;
                typename Mm::acc d = Mm::zero();
#pragma unroll
                for (int j = 0; j < 16; ++j) d = Mm::mad(a[j], xr[j], d);
                if (p0 + 16 * n + r < k) {
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (part[t] == kLow) continue;
                        // a high row sits at an even t, its low row at t + 1
                        double z = double(d[t]);
                        if (part[t] == kHigh) z = z * 65536.0 + double(d[(t + 1) & 3]);
                        const Real zs = Real(sc[t] * z);
                        Real* o = zout + at[t] + 16 * n;
                        o[0] = zs;
                        if (cur.nseg && part[t] == kHigh) o[k] = zs;
                    }
                }
            }
        }
        wave_lds_sync();  // the next item overwrites the stage and the rows
    }
}


// Pass L.  A workgroup owns one run of a row tile's items and one row group
// of the tile: RB = 4 / NS of its four 16-row blocks, for KC = 16 NS
// right-hand sides.  Its kWideWaves waves each take a share of the run
// (their own wave runs) and form the group's rows for all KC right-hand
// sides: per K-step a lane decodes the codes of its RB rows (one per row
// block) and each decoded code is the A operand of NS MFMAs, one per
// 16-column slice, while each B operand serves the RB row blocks.  The
// waves' FP64 sums are added in a fixed order at the end.  A wave stages
// only its group's rows of an item's side block (the words of lanes
// 16 rb0 .. 16 (rb0 + RB) - 1 of every word row), so the row groups of one
// run read disjoint bytes.  The row groups of a run are neighbouring
// workgroups of one XCD (xcd_order), so the partial sums they all read stay
// in its L2.
template <int KC, bool FP32> struct LeftShape {
    static constexpr int NS = KC / 16;     // 16-column slices per wave
    static constexpr int RB = 4 / NS;      // 16-row blocks per wave
    static constexpr int NG = NS;          // row groups (workgroups per run)
    static constexpr int W = 16 * RB;      // words of a staged word row
    static constexpr int kQuads = RB;      // 16-byte loads per lane per item (16 word rows)
    // per wave: a head row, 16 word rows, a slack row, then the virtual rows
    static constexpr int kHead = W * 4;
    static constexpr int kVrow = kHead + 17 * W * 4;
    static constexpr int kWaveBytes = kVrow + 128 * 4;
    // the final sums, one 16-wide slice at a time: [wave][row block][t][lane]
    static constexpr int kSumBytes = kWideWaves * RB * 4 * 64 * 8;
    static constexpr int kBytes = kWideWaves * kWaveBytes > kSumBytes ? kWideWaves * kWaveBytes
                                                                       : kSumBytes;
    static_assert(kWaveBytes % 16 == 0, "LDS alignment");
};

// What a pass-L wave fetches of an item one item ahead: its row group's
// words of the item's left side block (RB 16-byte groups per lane), the
// component widths (lanes = components) and the B operands of its first
// four K-steps.  A diagonal or empty item fetches nothing (every load out of
// range).
template <int KC, bool FP32> struct LeftFetch {
    quad_t q[LeftShape<KC, FP32>::kQuads];
    std::uint32_t bits;
    real_t<FP32> b[4][KC / 16];
};

// B operands of K-steps 4 c .. 4 c + 3 of a kLeft item: lane (q, r) holds,
// for step s and slice n, the partial sum of virtual row 4 (4 c + s) + q
// for right-hand side p0 + 16 n + r.  The item's rows are in virtual-row
// order, so the addresses need only the record: one lane offset per chunk,
// the step as a uniform offset and the slice as an immediate.  The loads are
// unconditional: a row past the item lies past the buffer's bytes and reads
// zero (its A operand is zero too), and a right-hand side past k reads the
// next row's first values or zero (an MFMA result column depends on its own
// B column only, and those columns are never written).
template <int KC, bool FP32>
__device__ __forceinline__ void left_b(const WideView& V, const ItemDev& f, std::uint32_t c,
                                       int q, int r, int p0, real_t<FP32> (&b)[4][KC / 16]) {
    using Real = real_t<FP32>;
    constexpr std::uint32_t esz = sizeof(Real);
    const std::uint32_t k = std::uint32_t(V.k);
    const bool u = f.kind == kLeft && f.count > 0;
    const __amdgpu_buffer_rsrc_t zb = buffer(static_cast<const Real*>(V.zw) + std::uint64_t(f.zrow) * k,
                                             u ? f.zstride * k * esz : 0u);
    const std::uint32_t lane_off = ((16u * c + std::uint32_t(q)) * k + std::uint32_t(p0 + r)) * esz;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int step_off = int(4u * std::uint32_t(s) * k * esz);
#pragma unroll
        for (int n = 0; n < KC / 16; ++n) {
            const std::uint32_t off = lane_off + 16u * esz * std::uint32_t(n);
            if constexpr (FP32) {
                b[s][n] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(zb, off, step_off, 0));
            } else {
                const auto v = __builtin_amdgcn_raw_buffer_load_b64(zb, off, step_off, 0);
                b[s][n] = __longlong_as_double(static_cast<long long>(
                    (static_cast<unsigned long long>(v[1]) << 32) | static_cast<unsigned long long>(v[0])));
            }
        }
    }
}

// The lane's 16-byte group t of its row group's words: word row w, words
// 4 cq .. 4 cq + 3 of the group's W words (lanes 16 rb0 + 4 cq ..).
template <int KC, bool FP32>
__device__ __forceinline__ void left_group_word(int lane, int t, std::uint32_t& w, std::uint32_t& cq) {
    using S = LeftShape<KC, FP32>;
    const std::uint32_t idx = std::uint32_t(lane + 64 * t);
    w = idx / std::uint32_t(S::W / 4);
    cq = idx % std::uint32_t(S::W / 4);
}

template <int KC, bool FP32>
__device__ __forceinline__ void left_fetch(const WideView& V, const ItemDev& f, int lane, int p0,
                                           int rb0, LeftFetch<KC, FP32>& F) {
    using S = LeftShape<KC, FP32>;
    const bool u = f.kind == kLeft && f.count > 0;
    const __amdgpu_buffer_rsrc_t data =
        buffer(V.words + 4ull * f.quad, u ? std::uint32_t(f.quads) * 16u : 0u);
#pragma unroll
    for (int t = 0; t < S::kQuads; ++t) {
        std::uint32_t w, cq;
        left_group_word<KC, FP32>(lane, t, w, cq);
        const std::uint32_t word = 64u * w + 16u * std::uint32_t(rb0) + 4u * cq;
        F.q[t] = __builtin_amdgcn_raw_buffer_load_b128(data, word * 4u, 0, 0);
    }
    F.bits = __builtin_amdgcn_raw_buffer_load_b32(buffer(V.widths + f.comp0, u ? f.count * 4u : 0u),
                                                  std::uint32_t(lane) * 4u, 0, 0);
    left_b<KC, FP32>(V, f, 0, lane >> 4, lane & 15, p0, F.b);
}

// One fetched item of a pass-L wave run, added to the wave's sums of its
// row group (sums[i][n] slot t: row 16 (rb0 + i) + 4 q + t, right-hand side
// p0 + 16 n + r).
template <int KC, bool FP32>
__device__ __forceinline__ void left_item(const WideView& V, const ItemDev& f,
                                          const LeftFetch<KC, FP32>& F, std::uint32_t R, int lane,
                                          int p0, int rb0, std::uint8_t* wl,
                                          f64x4 (&sums)[LeftShape<KC, FP32>::RB][KC / 16]) {
    using Real = real_t<FP32>;
    using S = LeftShape<KC, FP32>;
    constexpr int NS = KC / 16, RB = S::RB;
    const int q = lane >> 4, r = lane & 15;
    const int k = V.k;
    if (f.count == 0) return;  // past the end of the run
    if (f.kind == kDense) {
        // R x cnt FP64 values, column-major; lane (q, r) supplies row
        // 16 rb + a_row(r) of column 4 s + q and x of that column (buffer
        // loads past the item's values and columns read zero; rows past the
        // tile read its last row and are never written)
        const std::uint32_t cnt = min(std::uint32_t(f.count), std::uint32_t(kDenseColumns));
        const __amdgpu_buffer_rsrc_t dv = buffer(V.words + 4ull * f.quad, cnt * R * 8u);
        const __amdgpu_buffer_rsrc_t pc = buffer(V.perm + f.col, cnt * 4u);
        double b[kDenseColumns / 4][NS];
#pragma unroll
        for (int s = 0; s < kDenseColumns / 4; ++s) {
            const std::uint32_t c = 4u * std::uint32_t(s) + std::uint32_t(q);
            const std::int64_t xrow =
                std::int64_t(__builtin_amdgcn_raw_buffer_load_b32(pc, c * 4u, 0, 0)) * V.ldx;
#pragma unroll
            for (int n = 0; n < NS; ++n) b[s][n] = V.x[xrow + min(p0 + 16 * n + r, k - 1)];
        }
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            const std::uint32_t row =
                min(16u * std::uint32_t(rb0 + i) + std::uint32_t(a_row<false>(r)), R - 1u);
#pragma unroll
            for (int s = 0; s < kDenseColumns / 4; ++s) {
                const std::uint32_t c = 4u * std::uint32_t(s) + std::uint32_t(q);
                const double a = load_f64(dv, (c * R + row) * 8u);
#pragma unroll
                for (int n = 0; n < NS; ++n) sums[i][n] = Mma<false>::mad(a, b[s][n], sums[i][n]);
            }
        }
        return;
    }
    // the group's words: word row w of the block at stage[W w .. W w + W - 1]
    quad_t* stage_q = reinterpret_cast<quad_t*>(wl + S::kHead);
    const std::uint32_t* stage = reinterpret_cast<const std::uint32_t*>(wl + S::kHead);
    std::uint32_t* vrow = reinterpret_cast<std::uint32_t*>(wl + S::kVrow);
    const std::uint32_t rows16 = std::uint32_t(f.quads) / 16u;  // word rows of the block
#pragma unroll
    for (int t = 0; t < S::kQuads; ++t) {
        std::uint32_t w, cq;
        left_group_word<KC, FP32>(lane, t, w, cq);
        if (w < rows16) stage_q[w * std::uint32_t(S::W / 4) + cq] = F.q[t];
    }
    wave_lds_sync();
    const std::uint32_t nv = build_vrows<FP32>(F.bits, f.count, lane, vrow);
    wave_lds_sync();
    const std::uint32_t arow = FP32 ? std::uint32_t(r) : std::uint32_t(a_row<false>(r));
    typename Mma<FP32>::acc acc[RB][NS];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int n = 0; n < NS; ++n) acc[i][n] = Mma<FP32>::zero();
    // A: the codes of rows 16 (rb0 + i) + arow (lanes past the tile's rows
    // read the zero words of the block's padding lanes); a high part carries
    // its 2^16 (exact); rows past the item are zero.  K-steps from row 2 W
    // on are whole components (no split handling), and only the last one
    // can reach past the rows.
    const std::uint32_t split_rows = 2u * (nv - std::uint32_t(f.count));
    auto run_chunk = [&](std::uint32_t c, const Real (&b)[4][NS]) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const std::uint32_t v0 = 16u * c + 4u * std::uint32_t(s);
            const std::uint32_t v = v0 + std::uint32_t(q);
            const std::uint32_t m = vrow_at(vrow, v, nv);
            const std::uint32_t E = vrow_end(m);
            const std::uint32_t kk = E >> 5, sh = 32u - vrow_bits(m), part = vrow_part(m);
            const std::uint32_t* hi = stage + std::uint32_t(S::W) * kk + arow;
            const std::uint32_t* lo = stage + (int(S::W) * (int(kk) - 1) + int(arow));
            const bool split = v0 < split_rows, edge = v0 + 4u > nv;
#pragma unroll
            for (int i = 0; i < RB; ++i) {
                int code = static_cast<int>(__builtin_amdgcn_alignbit(hi[16 * i], lo[16 * i], E)) >> sh;
                Real a;
                if (split) {
                    code = code_part(code, part);
                    a = part == kHigh ? Real(code) * Real(65536) : Real(code);
                } else {
                    a = Real(code);
                }
                if (edge) a = v < nv ? a : Real(0);
#pragma unroll
                for (int n = 0; n < NS; ++n) acc[i][n] = Mma<FP32>::mad(a, b[s][n], acc[i][n]);
            }
        }
    };
    // chunk 0 came with the fetch; later chunks are requested where they
    // are used
    const std::uint32_t nchunks = (nv + 15u) / 16u;
    run_chunk(0, F.b);
    for (std::uint32_t c = 1; c < nchunks; ++c) {
        Real b1[4][NS];
        left_b<KC, FP32>(V, f, c, q, r, p0, b1);
        run_chunk(c, b1);
    }
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int n = 0; n < NS; ++n)
#pragma unroll
            for (int t = 0; t < 4; ++t) sums[i][n][t] += double(acc[i][n][t]);
    wave_lds_sync();  // the next item overwrites the stage and the rows
}

template <int KC, bool FP32>
__device__ __forceinline__ void left_group(const WideView& V, std::uint32_t g, std::uint32_t rg,
                                           int p0, std::uint8_t* lds) {
    using S = LeftShape<KC, FP32>;
    constexpr int NS = S::NS, RB = S::RB;
    const int tid = int(threadIdx.x);
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63, q = lane >> 4, r = lane & 15;
    const int rb0 = int(rg) * RB;
    const WideRun run = V.lruns[g * kWideWaves + wave];
    const std::uint32_t begin = __builtin_amdgcn_readfirstlane(run.begin);
    const std::uint32_t end = __builtin_amdgcn_readfirstlane(run.end);
    const TileDev* __restrict__ T = V.tiles + run.tile;
    const std::uint32_t R = min(T->rows, std::uint32_t(kTile));
    const std::uint32_t row0 = T->row0;
    const int k = V.k;
    std::uint8_t* wl = lds + wave * S::kWaveBytes;

    f64x4 sums[RB][NS];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int n = 0; n < NS; ++n) sums[i][n] = Mma<false>::zero();

    // Two fetch buffers alternate, so no register of a load in flight is
    // copied; both halves of the loop always run (an odd count ends with
    // an empty item).
    LeftFetch<KC, FP32> FA, FB;
    ItemDev fa = item_or_empty(V, begin, end);
    left_fetch<KC, FP32>(V, fa, lane, p0, rb0, FA);
    for (std::uint32_t i = begin; i < end; i += 2) {
        const ItemDev fb = item_or_empty(V, i + 1, end);
        left_fetch<KC, FP32>(V, fb, lane, p0, rb0, FB);
        left_item<KC, FP32>(V, fa, FA, R, lane, p0, rb0, wl, sums);
        fa = item_or_empty(V, i + 2, end);
        left_fetch<KC, FP32>(V, fa, lane, p0, rb0, FA);
        left_item<KC, FP32>(V, fb, FB, R, lane, p0, rb0, wl, sums);
    }

    // The four waves' sums, one slice at a time, through LDS: wave w adds
    // slot t = w of the group's row blocks of every wave as (s0 + s1) +
    // (s2 + s3) and writes rows 16 (rb0 + i) + 4 q + w of right-hand side
    // p0 + 16 n + r.
    double* part = reinterpret_cast<double*>(lds);
    double* sets = reinterpret_cast<double*>(static_cast<std::uint8_t*>(V.zw) +
                                             V.z_rows * std::uint64_t(k) * sizeof(double));
    for (int n = 0; n < NS; ++n) {
        __syncthreads();  // the stage (the first slice) or the previous slice is done with
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
            for (int t = 0; t < 4; ++t) part[((wave * RB + i) * 4 + t) * 64 + lane] = sums[i][n][t];
        __syncthreads();
        const int p = p0 + 16 * n + r;
#pragma unroll
        for (int i = 0; i < RB; ++i) {
            double v[kWideWaves];
#pragma unroll
            for (int w = 0; w < kWideWaves; ++w) v[w] = part[((w * RB + i) * 4 + wave) * 64 + lane];
            const double sum = (v[0] + v[1]) + (v[2] + v[3]);
            const std::uint32_t row = 16u * std::uint32_t(rb0 + i) + 4u * std::uint32_t(q) + std::uint32_t(wave);
            if (row < R && p < k) {
                if (run.set == kNoSet)
                    V.y[std::int64_t(V.perm[row0 + row]) * V.ldy + p] = sum;
                else
                    sets[(std::uint64_t(run.set) * kTile + row) * std::uint64_t(k) + std::uint64_t(p)] = sum;
            }
        }
    }
}

// sum_{s < count} p[s stride], added in order from zero, with the loads
// issued four at a time (the same rounding as the plain loop).
template <class T>
__device__ __forceinline__ double ordered_sum(const T* p, std::uint64_t stride,
                                              std::uint32_t count) {
    double v = 0.0;
    std::uint32_t s = 0;
    for (; s + 4u <= count; s += 4u) {
        const T a0 = p[s * stride], a1 = p[(s + 1u) * stride];
        const T a2 = p[(s + 2u) * stride], a3 = p[(s + 3u) * stride];
        v = (((v + double(a0)) + double(a1)) + double(a2)) + double(a3);
    }
    for (; s < count; ++s) v += double(p[s * stride]);
    return v;
}

// A block's partial sums over its column tiles, summed in order into rows
// in the components' virtual-row order (a split component's two rows both
// carry the sum), one thread per (row, right-hand side).  Wave 0 lays out the
// rows from the widths first, as build_vrows does.
template <bool FP32>
__device__ __forceinline__ void zsum_group(const WideView& V, std::uint32_t g, int p0, int kc,
                                           std::uint32_t* vrow) {
    using Real = real_t<FP32>;
    const ZSum zs = V.zsums[g];
    const std::uint32_t count = min(zs.count, std::uint32_t(kItemComponents));
    const std::uint32_t w = min(std::uint32_t(kc), std::uint32_t(V.k) - std::uint32_t(p0));
    __shared__ std::uint32_t nv_shared;
    if (threadIdx.x < 64) {
        const int lane = int(threadIdx.x);
        const std::uint32_t bits = std::uint32_t(lane) < count ? std::uint32_t(V.widths[zs.comp0 + lane]) : 0u;
        const std::uint32_t nv = build_vrows<FP32>(bits, count, lane, vrow);
        if (lane == 0) nv_shared = nv;
    }
    __syncthreads();
    const std::uint32_t nv = nv_shared;
    const std::uint32_t k = std::uint32_t(V.k);
    Real* z = static_cast<Real*>(V.zw);
    for (std::uint32_t e = threadIdx.x; e < nv * w; e += blockDim.x) {
        const std::uint32_t v = e / w, p = std::uint32_t(p0) + e % w;
        const std::uint32_t j = vrow_comp(vrow[v]);
        const double s = ordered_sum(z + (std::uint64_t(zs.zrow) + j) * k + p,
                                     std::uint64_t(zs.zstride) * k, zs.nseg);
        z[(std::uint64_t(zs.vzrow) + v) * k + p] = Real(s);
    }
}

// A tile's row sets summed in order into y: element `e` of the tile's rows
// x right-hand sides (row-major), one per thread, 256 per workgroup.
constexpr std::uint32_t kSumChunk = 256;
__device__ __forceinline__ void sum_group(const WideView& V, std::uint32_t g, std::uint32_t e) {
    const WideSum ws = V.sums[g];
    const TileDev* __restrict__ T = V.tiles + ws.tile;
    const std::uint32_t R = min(T->rows, std::uint32_t(kTile));
    const std::uint32_t k = std::uint32_t(V.k);
    const double* sets = reinterpret_cast<const double*>(static_cast<const std::uint8_t*>(V.zw) +
                                                         V.z_rows * std::uint64_t(k) * sizeof(double));
    if (e >= R * k) return;
    const std::uint32_t i = e / k, p = e - i * k;
    const double v = ordered_sum(sets + (std::uint64_t(ws.set0) * kTile + i) * k + p,
                                 std::uint64_t(kTile) * k, ws.count);
    V.y[std::int64_t(V.perm[T->row0 + i]) * V.ldy + p] = v;
}
// Workgroups per row-set sum for k right-hand sides.
inline unsigned sum_chunks(int k) { return unsigned((kTile * std::uint32_t(k) + kSumChunk - 1) / kSumChunk); }

// Waves per SIMD the register allocation must allow: pass R; pass L, whose
// two fetch buffers grow with the right-hand sides per pass.
template <bool FP32> constexpr int wide_occupancy() { return FP32 ? 4 : 3; }
// Pass L takes as many right-hand sides per launch slice as pass R in FP32
// mode; FP64 mode keeps 16 (a wave's FP64 item sums and B operands for more
// columns would hold it to two waves per SIMD, which measured slower than
// decoding once per 16 columns).
template <bool FP32> constexpr int left_kc(int kc) { return FP32 ? kc : 16; }
template <int LKC, bool FP32> constexpr int left_occupancy() { return FP32 ? (LKC == 16 ? 4 : 3) : 3; }
template <int LKC, bool FP32> constexpr int batch_left_occupancy() { return LKC == 16 ? 3 : 2; }

// Pass R and pass L take (workgroup, slice) pairs in XCD order (xcd_order):
// the slices of one workgroup's work (right-hand-side slices, pass L's row
// groups) run on one XCD one after another, so the codes they all read are
// fetched into one L2.
template <int KC, bool FP32>
__global__ __launch_bounds__(64 * kWideWaves, wide_occupancy<FP32>()) void wide_right(
    WideView V, std::uint32_t groups, std::uint32_t slices) {
    __shared__ __align__(16) std::uint8_t lds[RightShape<KC, FP32>::kBytes];
    std::uint32_t g, s;
    xcd_order(blockIdx.x, slices, g, s);
    if (g >= groups) return;  // the grid's padding to a multiple of 8
    right_group<KC, FP32>(V, g, int(s) * KC, lds);
}

template <int LKC, bool FP32>
__global__ __launch_bounds__(64 * kWideWaves, (left_occupancy<LKC, FP32>())) void wide_left(
    WideView V, std::uint32_t groups, std::uint32_t slices) {
    __shared__ __align__(16) std::uint8_t lds[LeftShape<LKC, FP32>::kBytes];
    // slice s: row group s mod NG of right-hand-side slice s / NG
    std::uint32_t g, s;
    xcd_order(blockIdx.x, slices, g, s);
    if (g >= groups) return;
    constexpr std::uint32_t NG = LeftShape<LKC, FP32>::NG;
    left_group<LKC, FP32>(V, g, s % NG, int(s / NG) * LKC, lds);
}

__global__ __launch_bounds__(kSumChunk) void wide_sum(WideView V) {
    sum_group(V, blockIdx.x, blockIdx.y * kSumChunk + threadIdx.x);
}

template <bool FP32>
__global__ __launch_bounds__(256) void wide_zsum(WideView V, int kc) {
    __shared__ std::uint32_t vrow[128];
    zsum_group<FP32>(V, blockIdx.x, int(blockIdx.y) * kc, kc, vrow);
}

// Batched: member m of a launch owns global workgroups [first[m], first[m+1]).
__device__ __forceinline__ std::uint32_t owner(const std::uint32_t* __restrict__ first,
                                               std::uint32_t count, std::uint32_t index) {
    std::uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
        const std::uint32_t mid = (lo + hi) >> 1;
        if (first[mid] <= index) lo = mid;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ WideView bind(const WideView* __restrict__ members, std::uint32_t m,
                                         const double* x, double* y, std::uint8_t* zw) {
    WideView V = members[m];
    V.x = x + reinterpret_cast<std::uintptr_t>(V.x);
    V.y = y + reinterpret_cast<std::uintptr_t>(V.y);
    V.zw = zw + reinterpret_cast<std::uintptr_t>(V.zw);
    return V;
}

template <int KC, bool FP32>
__global__ __launch_bounds__(64 * kWideWaves, wide_occupancy<FP32>()) void batch_wide_right(
    const WideView* __restrict__ members, const std::uint32_t* __restrict__ first,
    std::uint32_t count, const double* x, double* y, std::uint8_t* zw, std::uint32_t slices) {
    __shared__ __align__(16) std::uint8_t lds[RightShape<KC, FP32>::kBytes];
    std::uint32_t g, s;
    xcd_order(blockIdx.x, slices, g, s);
    if (g >= first[count]) return;  // the grid's padding to a multiple of 8
    const std::uint32_t m = owner(first, count, g);
    const WideView V = bind(members, m, x, y, zw);
    const int p0 = int(s) * KC;
    if (p0 >= V.k) return;  // the whole workgroup leaves together
    right_group<KC, FP32>(V, g - first[m], p0, lds);
}

template <int LKC, bool FP32>
__global__ __launch_bounds__(64 * kWideWaves, (batch_left_occupancy<LKC, FP32>())) void batch_wide_left(
    const WideView* __restrict__ members, const std::uint32_t* __restrict__ first,
    std::uint32_t count, const double* x, double* y, std::uint8_t* zw, std::uint32_t slices) {
    __shared__ __align__(16) std::uint8_t lds[LeftShape<LKC, FP32>::kBytes];
    std::uint32_t g, s;
    xcd_order(blockIdx.x, slices, g, s);
    if (g >= first[count]) return;
    const std::uint32_t m = owner(first, count, g);
    const WideView V = bind(members, m, x, y, zw);
    constexpr std::uint32_t NG = LeftShape<LKC, FP32>::NG;
    const int p0 = int(s / NG) * LKC;
    if (p0 >= V.k) return;
    left_group<LKC, FP32>(V, g - first[m], s % NG, p0, lds);
}

template <bool FP32>
__global__ __launch_bounds__(256) void batch_wide_zsum(const WideView* __restrict__ members,
                                                       const std::uint32_t* __restrict__ first,
                                                       std::uint32_t count, const double* x,
                                                       double* y, std::uint8_t* zw, int kc) {
    const std::uint32_t m = owner(first, count, blockIdx.x);
    const WideView V = bind(members, m, x, y, zw);
    __shared__ std::uint32_t vrow[128];
    const int p0 = int(blockIdx.y) * kc;
    if (p0 >= V.k) return;
    zsum_group<FP32>(V, blockIdx.x - first[m], p0, kc, vrow);
}

__global__ __launch_bounds__(kSumChunk) void batch_wide_sum(const WideView* __restrict__ members,
                                                            const std::uint32_t* __restrict__ first,
                                                            std::uint32_t count, const double* x,
                                                            double* y, std::uint8_t* zw) {
    const std::uint32_t m = owner(first, count, blockIdx.x);
    const WideView V = bind(members, m, x, y, zw);
    sum_group(V, blockIdx.x - first[m], blockIdx.y * kSumChunk + threadIdx.x);
}

template <int KC, bool FP32>
void launch_wide_kc(const WideView& V, const WideCounts& c, hipStream_t st, hipEvent_t* ev) {
    const unsigned passes = unsigned((V.k + KC - 1) / KC);
    if (ev) hip_check(hipEventRecord(ev[0], st), "hipEventRecord");
    if (c.right_groups) {
        hipLaunchKernelGGL((wide_right<KC, FP32>), dim3(xcd_grid(c.right_groups, passes)),
                           dim3(64 * kWideWaves), 0, st, V, c.right_groups, passes);
        hip_check(hipGetLastError(), "wide pass R launch");
    }
    if (c.zsums) {
        hipLaunchKernelGGL((wide_zsum<FP32>), dim3(c.zsums, passes), dim3(256), 0, st, V, KC);
        hip_check(hipGetLastError(), "wide partial-sum reduction launch");
    }
    if (ev) hip_check(hipEventRecord(ev[1], st), "hipEventRecord");
    constexpr int LKC = left_kc<FP32>(KC);
    const unsigned lslices = unsigned((V.k + LKC - 1) / LKC) * LeftShape<LKC, FP32>::NG;
    hipLaunchKernelGGL((wide_left<LKC, FP32>), dim3(xcd_grid(c.left_groups, lslices)),
                       dim3(64 * kWideWaves), 0, st, V, c.left_groups, lslices);
    hip_check(hipGetLastError(), "wide pass L launch");
    if (c.sums) {
        hipLaunchKernelGGL(wide_sum, dim3(c.sums, sum_chunks(V.k)), dim3(kSumChunk), 0, st, V);
        hip_check(hipGetLastError(), "wide row-set sum launch");
    }
    if (ev) hip_check(hipEventRecord(ev[2], st), "hipEventRecord");
}

template <int KC, bool FP32>
void launch_batch_wide_kc(const WideBatchClass& c, const double* x, double* y, std::uint8_t* zw,
                          hipStream_t st) {
    const unsigned passes = unsigned((c.max_k + KC - 1) / KC);
    const std::uint32_t n = std::uint32_t(c.member_count);
    if (c.right_groups) {
        hipLaunchKernelGGL((batch_wide_right<KC, FP32>),
                           dim3(xcd_grid(c.right_groups, passes)),
                           dim3(64 * kWideWaves), 0, st, c.members, c.first_right, n, x, y, zw, passes);
        hip_check(hipGetLastError(), "batched wide pass R launch");
    }
    if (c.zsums) {
        hipLaunchKernelGGL((batch_wide_zsum<FP32>), dim3(c.zsums, passes), dim3(256), 0, st,
                           c.members, c.first_zsum, n, x, y, zw, KC);
        hip_check(hipGetLastError(), "batched wide partial-sum reduction launch");
    }
    constexpr int LKC = left_kc<FP32>(KC);
    const unsigned lslices = unsigned((c.max_k + LKC - 1) / LKC) * LeftShape<LKC, FP32>::NG;
    hipLaunchKernelGGL((batch_wide_left<LKC, FP32>),
                       dim3(xcd_grid(c.left_groups, lslices)),
                       dim3(64 * kWideWaves), 0, st, c.members, c.first_left, n, x, y, zw, lslices);
    hip_check(hipGetLastError(), "batched wide pass L launch");
    if (c.sums) {
        hipLaunchKernelGGL(batch_wide_sum, dim3(c.sums, sum_chunks(c.max_k)), dim3(kSumChunk), 0,
                           st, c.members, c.first_sum, n, x, y, zw);
        hip_check(hipGetLastError(), "batched wide row-set sum launch");
    }
}

}  // namespace

void launch_wide(const WideView& V, const WideCounts& c, int kc, bool fp32, void* stream,
                 void* events) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipEvent_t* ev = static_cast<hipEvent_t*>(events);
    if (fp32) {
        switch (kc) {
            case 16: return launch_wide_kc<16, true>(V, c, st, ev);
            case 32: return launch_wide_kc<32, true>(V, c, st, ev);
            default: return launch_wide_kc<64, true>(V, c, st, ev);
        }
    }
    switch (kc) {
        case 16: return launch_wide_kc<16, false>(V, c, st, ev);
        case 32: return launch_wide_kc<32, false>(V, c, st, ev);
        default: return launch_wide_kc<64, false>(V, c, st, ev);
    }
}

void launch_batch_wide(const WideBatchClass& c, const double* x, double* y, void* zw, bool fp32,
                       void* stream) {
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::uint8_t* w = static_cast<std::uint8_t*>(zw);
    if (fp32) {
        switch (c.kc) {
            case 16: return launch_batch_wide_kc<16, true>(c, x, y, w, st);
            case 32: return launch_batch_wide_kc<32, true>(c, x, y, w, st);
            default: return launch_batch_wide_kc<64, true>(c, x, y, w, st);
        }
    }
    switch (c.kc) {
        case 16: return launch_batch_wide_kc<16, false>(c, x, y, w, st);
        case 32: return launch_batch_wide_kc<32, false>(c, x, y, w, st);
        default: return launch_batch_wide_kc<64, false>(c, x, y, w, st);
    }
}

}  // namespace qhamd
