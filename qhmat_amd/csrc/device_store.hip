// MI355X (gfx950) apply of a qhmat H-matrix, y = H x.
//
// The packed store is uploaded once in the tile-major layout described in
// device_store.hpp.  An apply is two launches on the caller's stream, both
// deterministic (fixed summation order, one writer per output element, no
// floating-point atomics):
//
//  1. right-factor pass: one wave per (low-rank block, column tile).  The
//     wave streams the tile's chunk into LDS with 16-byte loads and the
//     tile's x rows (gathered through the permutation) next to it; four
//     lanes share each component, decode their codes from LDS
//     (v_alignbit_b32 + v_bfe_i32), multiply by x and are reduced in the
//     quad.  Output: partial dot products
//         zp[zbase + s*rank + j][p] = sum_{c in tile s} codeR(c, j) x[c, p].
//  2. row-tile pass: one workgroup of eight waves per row tile.  The waves
//     split the tile's blocks in insertion order; a wave streams each
//     block's chunk for the tile into LDS and every lane (one row) decodes
//     its code of each component and accumulates
//         y[row, p] += sum_j codeL(row, j) * (sL_j sR_j sum_s zp[.][p]),
//     plus D[row, :] x for the diagonal block.  The eight partial rows are
//     added in wave order and written once to y[perm[row], p].
//
// FP64 mode computes in FP64 throughout.  FP32 split mode converts codes of
// at most 24 bits and x to FP32 (exact for the codes), splits wider codes
// into a signed high half and an unsigned 16-bit low half accumulated apart,
// sums each block's contribution to a row in FP32 and everything across
// blocks, the scales and the diagonal in FP64.
//
// Reference: src/qhmat/cuda.cu:64-164 (one thread per output entry scanning
// every block, store uploaded per call) and core.cpp:360-390 (FP64 loop nest).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "device_store.hpp"

namespace qhamd {
namespace {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Makes `device` current for the scope and restores the caller's device.
struct DeviceScope {
    int previous = -1;
    explicit DeviceScope(int device) {
        hip_check(hipGetDevice(&previous), "hipGetDevice");
        if (previous != device) hip_check(hipSetDevice(device), "hipSetDevice");
    }
    ~DeviceScope() {
        int now = -1;
        if (hipGetDevice(&now) == hipSuccess && now != previous && previous >= 0)
            (void)hipSetDevice(previous);
    }
};

void require(bool ok, const char* message) {
    if (!ok) throw std::invalid_argument(message);
}

// ---------------------------------------------------------------------------
// Device helpers

// Code number `idx` of the stream at `seg` (LDS or global).  The word pair is
// always read: a code that does not straddle takes all its bits from the low
// word, and one word past every staged stream is kept readable.
__device__ __forceinline__ int decode_code(const std::uint32_t* seg, std::uint32_t idx,
                                           std::uint32_t bits) {
    const std::uint32_t pos = idx * bits;
    const std::uint32_t w = pos >> 5;
    const std::uint32_t raw = __builtin_amdgcn_alignbit(seg[w + 1], seg[w], pos & 31u);
    return bits == 32 ? static_cast<int>(raw) : __builtin_amdgcn_sbfe(raw, 0u, bits);
}

__host__ __device__ __forceinline__ std::uint32_t stream_words(std::uint32_t count,
                                                               std::uint32_t bits) {
    return (count * bits + 31u) >> 5;
}

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

// LDS written by some lanes of a wave and read by others: keep the compiler
// from moving accesses across this point and wait for the LDS queue.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Copies `quads` 16-byte groups from global (16-byte aligned) to LDS.  Up to
// four loads per lane are in flight before the first LDS write waits.
// (A native clang vector type: an array of HIP's union-based uint4 is kept
// in private scratch memory by the compiler.)
typedef std::uint32_t quad_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void stage_quads(std::uint32_t* __restrict__ lds,
                                            const std::uint32_t* __restrict__ src,
                                            std::uint32_t quads, int lane) {
    const quad_t* s = reinterpret_cast<const quad_t*>(src);
    quad_t* d = reinterpret_cast<quad_t*>(lds);
    for (std::uint32_t q0 = 0; q0 < quads; q0 += 256) {
        const std::uint32_t q = q0 + lane;
        const quad_t zero = {0u, 0u, 0u, 0u};
        const quad_t v0 = q < quads ? s[q] : zero;
        const quad_t v1 = q + 64 < quads ? s[q + 64] : zero;
        const quad_t v2 = q + 128 < quads ? s[q + 128] : zero;
        const quad_t v3 = q + 192 < quads ? s[q + 192] : zero;
        if (q < quads) d[q] = v0;
        if (q + 64 < quads) d[q + 64] = v1;
        if (q + 128 < quads) d[q + 128] = v2;
        if (q + 192 < quads) d[q + 192] = v3;
    }
}

// ---------------------------------------------------------------------------
// Pass 1: right factor, one wave per (block, column tile).
template <bool FP32> using real_t = typename std::conditional<FP32, float, double>::type;

template <int KB, bool FP32>
__device__ __forceinline__ void right_factor_unit(const MemberView& M, std::uint32_t unit, int p0,
                                                  std::uint32_t* __restrict__ stage,
                                                  real_t<FP32> (*xs)[KB]) {
    using Real = real_t<FP32>;
    const std::int32_t* __restrict__ widths = M.widths;
    const std::uint32_t* __restrict__ words = M.words;
    const std::int32_t* __restrict__ perm = M.perm;
    const double* __restrict__ x = M.x;
    double* __restrict__ zp = M.zp;
    const std::int64_t ldx = M.ldx;
    const int k = M.k;

    const UnitDev u = M.units[unit];
    const BlockDev B = M.blocks[u.block];
    const int lane = threadIdx.x & 63;
    const int kk = min(KB, k - p0);
    const std::uint32_t C = u.cols;

    // Request everything before waiting on anything: the first pass's widths,
    // the permutation rows, the whole chunk when it fits the stage buffer,
    // then the x rows.
    const bool whole = u.words + 5 <= std::uint32_t(kStageWords);
    const std::uint32_t delta0 = std::uint32_t(u.chunk & 3u);
    const std::uint32_t jn0 = min(std::uint32_t(kPassComponents), B.rank);
    std::uint32_t wl = lane < int(jn0) ? std::uint32_t(widths[B.comp_base + lane]) : 0u;
    const std::int64_t xrow = lane < int(C) ? std::int64_t(perm[u.col0 + lane]) * ldx + p0 : 0;
    if (whole) stage_quads(stage, words + (u.chunk - delta0), (u.words + delta0 + 1 + 3) >> 2, lane);
    if (lane < int(C)) {
#pragma unroll
        for (int p = 0; p < KB; ++p) xs[lane][p] = p < kk ? Real(x[xrow + p]) : Real(0);
    }

    const int g = lane >> 2;  // component slot of this lane in a pass
    const int q = lane & 3;   // columns q, q+4, q+8, ... of the tile
    std::uint32_t done = 0;   // chunk words of the components already processed
    for (std::uint32_t j0 = 0; j0 < B.rank; j0 += kPassComponents) {
        const std::uint32_t jn = min(std::uint32_t(kPassComponents), B.rank - j0);
        // widths and stream offsets of this pass's components (prefix scan)
        if (j0 > 0) wl = lane < int(jn) ? std::uint32_t(widths[B.comp_base + j0 + lane]) : 0u;
        const std::uint32_t nwl = stream_words(C, wl);
        std::uint32_t incl = nwl;
#pragma unroll
        for (int off = 1; off < kPassComponents; off <<= 1) {
            const std::uint32_t t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        const std::uint32_t total = __shfl(incl, kPassComponents - 1, 64);
        const std::uint32_t bits = __shfl(wl, g & (kPassComponents - 1), 64);
        const std::uint32_t seg = __shfl(incl - nwl, g & (kPassComponents - 1), 64);

        std::uint32_t base = delta0 + done;  // first word of this pass in the stage buffer
        if (!whole) {
            const std::uint64_t first = u.chunk + done;
            base = std::uint32_t(first & 3u);
            stage_quads(stage, words + (first - base), (total + base + 1 + 3) >> 2, lane);
        }
        wave_lds_sync();

        double tot[KB];
        const bool mine = g < int(jn);
        const std::uint32_t* sw = stage + base + seg;
        if (!FP32 || bits <= 24) {
            Real part[KB];
#pragma unroll
            for (int p = 0; p < KB; ++p) part[p] = Real(0);
            if (mine) {
                for (std::uint32_t c = q; c < C; c += 4) {
                    const Real code = Real(decode_code(sw, c, bits));
#pragma unroll
                    for (int p = 0; p < KB; ++p) part[p] = fma(code, xs[c][p], part[p]);
                }
            }
#pragma unroll
            for (int p = 0; p < KB; ++p) {
                Real v = part[p];
                v += __shfl_xor(v, 1, 64);
                v += __shfl_xor(v, 2, 64);
                tot[p] = double(v);
            }
        } else {
            // 25..32-bit codes: signed high half and unsigned low 16 bits,
            // both exact in FP32, summed apart and joined in FP64.
            float hi[KB], lo[KB];
#pragma unroll
            for (int p = 0; p < KB; ++p) hi[p] = lo[p] = 0.0f;
            for (std::uint32_t c = q; c < C; c += 4) {
                const int code = decode_code(sw, c, bits);
                const float h = float(code >> 16), l = float(code & 0xFFFF);
#pragma unroll
                for (int p = 0; p < KB; ++p) {
                    hi[p] = fmaf(h, float(xs[c][p]), hi[p]);
                    lo[p] = fmaf(l, float(xs[c][p]), lo[p]);
                }
            }
#pragma unroll
            for (int p = 0; p < KB; ++p) {
                float h = hi[p], l = lo[p];
                h += __shfl_xor(h, 1, 64);
                h += __shfl_xor(h, 2, 64);
                l += __shfl_xor(l, 1, 64);
                l += __shfl_xor(l, 2, 64);
                tot[p] = double(h) * 65536.0 + double(l);
            }
        }
        if (mine && q == 0) {
            double* out = zp + (std::uint64_t(u.zrow) + j0 + g) * std::uint64_t(k) + p0;
#pragma unroll
            for (int p = 0; p < KB; ++p)
                if (p < kk) out[p] = tot[p];
        }
        done += total;
        if (!whole) wave_lds_sync();  // the next pass restages
    }
}

// kUnitWaves units per workgroup, one per wave (a one-wave workgroup would
// leave the per-CU workgroup limit, not the waves, bounding occupancy).
template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kUnitWaves) void right_factor_pass(MemberView M) {
    __shared__ __align__(16) std::uint32_t stage[kUnitWaves][kStageWords + 8];
    __shared__ real_t<FP32> xs[kUnitWaves][kTile][KB];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const std::uint32_t unit = blockIdx.x * kUnitWaves + wave;
    if (unit >= M.units_count) return;
    right_factor_unit<KB, FP32>(M, unit, blockIdx.y * KB, stage[wave], xs[wave]);
}

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
// as offsets (xoff, yoff, zoff) from the call's base pointers.
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
__global__ __launch_bounds__(64 * kUnitWaves) void right_factor_batch(
    const MemberView* __restrict__ members, const std::uint32_t* __restrict__ first,
    std::uint32_t count, std::uint32_t total, const double* x, double* y, double* zp) {
    __shared__ __align__(16) std::uint32_t stage[kUnitWaves][kStageWords + 8];
    __shared__ real_t<FP32> xs[kUnitWaves][kTile][KB];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const std::uint32_t unit = blockIdx.x * kUnitWaves + wave;
    if (unit >= total) return;
    const std::uint32_t m = locate_member(first, count, unit);
    const MemberView M = bind_member(members, m, x, y, zp);
    const int p0 = blockIdx.y * KB;
    if (p0 >= M.k) return;
    right_factor_unit<KB, FP32>(M, unit - first[m], p0, stage[wave], xs[wave]);
}

// ---------------------------------------------------------------------------
// Pass 2: row tiles, kRowWaves waves per tile.
template <int KB, bool FP32>
__device__ __forceinline__ void row_tile_unit(const MemberView& M, std::uint32_t tile_index,
                                              int p0,
                                              std::uint32_t (*stage_all)[kStageWords + 8],
                                              double (*partial)[2][64]) {
    const EntryDev* __restrict__ entries = M.entries;
    const BlockDev* __restrict__ blocks = M.blocks;
    const DenseDev* __restrict__ denses = M.denses;
    const std::int32_t* __restrict__ widths = M.widths;
    const double* __restrict__ lscale = M.lscale;
    const double* __restrict__ rscale = M.rscale;
    const std::uint32_t* __restrict__ words = M.words;
    const double* __restrict__ dvals = M.dense;
    const std::int32_t* __restrict__ perm = M.perm;
    const double* __restrict__ zp = M.zp;
    const double* __restrict__ x = M.x;
    double* __restrict__ y = M.y;
    const std::int64_t ldx = M.ldx, ldy = M.ldy;
    const int k = M.k;

    // Fields read through the pointer: a local copy indexed by the wave
    // number would be placed in private (scratch) memory.
    const TileDev* __restrict__ tile = M.tiles + tile_index;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int kk = min(KB, k - p0);
    const std::uint32_t R = tile->rows;
    const bool active = lane < int(R);
    std::uint32_t* stage = stage_all[wave];

    double acc[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) acc[p] = 0.0;

    const std::uint32_t e_begin = tile->split[wave], e_end = tile->split[wave + 1];
    for (std::uint32_t e = e_begin; e < e_end; ++e) {
        const EntryDev ent = entries[e];
        if (ent.what & kDenseBit) {
            // Up to kDenseCols columns of a diagonal block (column-major, one
            // coalesced column per lane-row step): every load is issued before
            // the first FMA; the x rows are staged in LDS and read as broadcasts.
            const DenseDev D = denses[ent.what & ~kDenseBit];
            const std::uint32_t cbeg = ent.words, cnt = ent.cols;
            double* xsd = reinterpret_cast<double*>(stage);
            const double* col = dvals + D.off + std::uint64_t(cbeg) * D.m + ent.row + lane;
            double dv[kDenseCols];
#pragma unroll
            for (int i = 0; i < kDenseCols; ++i)
                dv[i] = (active && i < int(cnt)) ? col[std::uint64_t(i) * D.m] : 0.0;
            if (lane < int(cnt)) {
                const double* xr = x + std::int64_t(perm[D.col0 + cbeg + lane]) * ldx + p0;
#pragma unroll
                for (int p = 0; p < KB; ++p) xsd[lane * KB + p] = p < kk ? xr[p] : 0.0;
            }
            wave_lds_sync();
#pragma unroll
            for (int i = 0; i < kDenseCols; ++i) {
                if (i < int(cnt)) {
#pragma unroll
                    for (int p = 0; p < KB; ++p) acc[p] = fma(dv[i], xsd[i * KB + p], acc[p]);
                }
            }
            wave_lds_sync();
            continue;
        }
        const BlockDev B = blocks[ent.what];
        float blk[KB];
#pragma unroll
        for (int p = 0; p < KB; ++p) blk[p] = 0.0f;
        // Lanes own components: each loads its width, scales and right-factor
        // sums at once; the component loop then reads them with readlane.
        // The whole chunk is staged in the same breath when it fits.
        const bool whole = ent.words + 5 <= std::uint32_t(kStageWords) && B.rank <= 64;
        std::uint64_t first = ent.chunk;
        std::uint32_t base = std::uint32_t(first & 3u);
        for (std::uint32_t j0 = 0; j0 < B.rank;) {
            const std::uint32_t jn = min(64u, B.rank - j0);
            const bool mine = lane < int(jn);
            const std::uint32_t cj = B.comp_base + j0 + lane;
            const std::uint32_t bits_l = mine ? std::uint32_t(widths[cj]) : 0u;
            const double s_l = mine ? lscale[cj] * rscale[cj] : 0.0;
            double z_l[KB];
#pragma unroll
            for (int p = 0; p < KB; ++p) z_l[p] = 0.0;
            if (mine) {
                for (std::uint32_t sg = 0; sg < B.nseg; ++sg) {
                    const double* zr = zp + (std::uint64_t(B.zbase) + sg * B.rank + j0 + lane) *
                                                std::uint64_t(k) + p0;
#pragma unroll
                    for (int p = 0; p < KB; ++p)
                        if (p < kk) z_l[p] += zr[p];
                }
            }
            if (whole) stage_quads(stage, words + (first - base), (ent.words + base + 1 + 3) >> 2,
                                   lane);
            const std::uint32_t nw_l = mine ? stream_words(R, bits_l) : 0u;
            std::uint32_t incl = nw_l;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const std::uint32_t t = __shfl_up(incl, off, 64);
                if (lane >= off) incl += t;
            }
            std::uint32_t take = jn;
            if (!whole) {  // the components whose streams fit the stage buffer
                const bool fits = mine && incl + base + 5 <= std::uint32_t(kStageWords);
                take = std::uint32_t(__popcll(__ballot(fits)));
            }
            const std::uint32_t total = __builtin_amdgcn_readlane(incl, take - 1);
            if (!whole)
                stage_quads(stage, words + (first - base), (total + base + 1 + 3) >> 2, lane);
            const std::uint32_t off_l = incl - nw_l;
            real_t<FP32> zs_l[KB];
#pragma unroll
            for (int p = 0; p < KB; ++p) zs_l[p] = real_t<FP32>(s_l * z_l[p]);
            wave_lds_sync();
            for (std::uint32_t j = 0; j < take; ++j) {
                const std::uint32_t bits = __builtin_amdgcn_readlane(bits_l, j);
                const std::uint32_t off = __builtin_amdgcn_readlane(off_l, j);
                const int code = active ? decode_code(stage + base + off, lane, bits) : 0;
                if (!FP32) {
                    const double cd = double(code);
#pragma unroll
                    for (int p = 0; p < KB; ++p)
                        acc[p] = fma(cd, double(lane_value(zs_l[p], j)), acc[p]);
                } else if (bits <= 24) {
                    const float cf = float(code);
#pragma unroll
                    for (int p = 0; p < KB; ++p)
                        blk[p] = fmaf(cf, float(lane_value(zs_l[p], j)), blk[p]);
                } else {
                    const float h = float(code >> 16), l = float(code & 0xFFFF);
#pragma unroll
                    for (int p = 0; p < KB; ++p) {
                        const float zf = float(lane_value(zs_l[p], j));
                        blk[p] = fmaf(h, zf * 65536.0f, blk[p]);
                        blk[p] = fmaf(l, zf, blk[p]);
                    }
                }
            }
            first += total;
            j0 += take;
            wave_lds_sync();  // the next round (or entry) restages
            base = std::uint32_t(first & 3u);
        }
        if (FP32) {
#pragma unroll
            for (int p = 0; p < KB; ++p) acc[p] += double(blk[p]);
        }
    }

    // Add the waves' partial rows in wave order, two right-hand sides at a time.
    const std::int64_t yrow = active ? std::int64_t(perm[tile->row0 + lane]) * ldy + p0 : 0;
#pragma unroll
    for (int pc = 0; pc < KB; pc += 2) {
        __syncthreads();
        if (wave > 0) {
            partial[wave - 1][0][lane] = acc[pc];
            partial[wave - 1][1][lane] = pc + 1 < KB ? acc[pc + 1] : 0.0;
        }
        __syncthreads();
        if (wave == 0 && active) {
            double a0 = acc[pc], a1 = pc + 1 < KB ? acc[pc + 1] : 0.0;
#pragma unroll
            for (int w = 0; w < kRowWaves - 1; ++w) {
                a0 += partial[w][0][lane];
                a1 += partial[w][1][lane];
            }
            if (pc < kk) y[yrow + pc] = a0;
            if (pc + 1 < kk) y[yrow + pc + 1] = a1;
        }
    }
}

template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kRowWaves) void row_tile_pass(MemberView M) {
    __shared__ __align__(16) std::uint32_t stage_all[kRowWaves][kStageWords + 8];
    __shared__ double partial[kRowWaves - 1][2][64];
    row_tile_unit<KB, FP32>(M, blockIdx.x, blockIdx.y * KB, stage_all, partial);
}

template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kRowWaves) void row_tile_batch(
    const MemberView* __restrict__ members, const std::uint32_t* __restrict__ first,
    std::uint32_t count, const double* x, double* y, double* zp) {
    __shared__ __align__(16) std::uint32_t stage_all[kRowWaves][kStageWords + 8];
    __shared__ double partial[kRowWaves - 1][2][64];
    const std::uint32_t m = locate_member(first, count, blockIdx.x);
    const MemberView M = bind_member(members, m, x, y, zp);
    const int p0 = blockIdx.y * KB;
    if (p0 >= M.k) return;  // the whole workgroup leaves together
    row_tile_unit<KB, FP32>(M, blockIdx.x - first[m], p0, stage_all, partial);
}

// Test hook: codes of one factor side of one block, from its tile chunks,
// written component-major in the block's own row order.
__global__ void decode_side(const std::uint32_t* __restrict__ words,
                            const ChunkRef* __restrict__ chunks,
                            const std::int32_t* __restrict__ widths, std::uint32_t rank,
                            std::uint32_t rows, std::int32_t* __restrict__ out) {
    const ChunkRef ch = chunks[blockIdx.x];
    std::uint64_t off = ch.chunk;
    for (std::uint32_t j = 0; j < rank; ++j) {
        const std::uint32_t bits = widths[j];
        for (std::uint32_t i = threadIdx.x; i < ch.count; i += blockDim.x)
            out[std::uint64_t(j) * rows + ch.first + i] = decode_code(words + off, i, bits);
        off += stream_words(ch.count, bits);
    }
}

MemberView member_view(const DeviceStore& s, const double* x, std::int64_t ldx, int k, double* y,
                       std::int64_t ldy, double* zp) {
    MemberView M{};
    M.perm = s.perm;
    M.words = s.words;
    M.widths = s.widths;
    M.lscale = s.lscale;
    M.rscale = s.rscale;
    M.dense = s.dense;
    M.blocks = s.blocks;
    M.denses = s.denses;
    M.units = s.units;
    M.tiles = s.tiles;
    M.entries = s.entries;
    M.x = x;
    M.y = y;
    M.zp = zp;
    M.ldx = ldx;
    M.ldy = ldy;
    M.k = k;
    M.units_count = s.n_units;
    M.tiles_count = s.n_tiles;
    return M;
}

// Event slots of the next profiled apply, or null when not profiling.
hipEvent_t* profile_slot(const DeviceStore& s) {
    auto& p = s.profile;
    if (!p.armed || 3 * (p.used + 1) > p.events.size()) return nullptr;
    return reinterpret_cast<hipEvent_t*>(p.events.data() + 3 * p.used++);
}

template <int KB, bool FP32>
void launch_apply(const DeviceStore& s, const double* x, std::int64_t ldx, int k, double* y,
                  std::int64_t ldy, double* zp, hipStream_t stream) {
    const unsigned passes = unsigned((k + KB - 1) / KB);
    const MemberView M = member_view(s, x, ldx, k, y, ldy, zp);
    hipEvent_t* ev = profile_slot(s);
    if (ev) hip_check(hipEventRecord(ev[0], stream), "hipEventRecord");
    if (s.n_units) {
        hipLaunchKernelGGL((right_factor_pass<KB, FP32>),
                           dim3((s.n_units + kUnitWaves - 1) / kUnitWaves, passes),
                           dim3(64 * kUnitWaves), 0, stream, M);
        hip_check(hipGetLastError(), "right-factor pass launch");
    }
    if (ev) hip_check(hipEventRecord(ev[1], stream), "hipEventRecord");
    hipLaunchKernelGGL((row_tile_pass<KB, FP32>), dim3(s.n_tiles, passes), dim3(64 * kRowWaves),
                       0, stream, M);
    hip_check(hipGetLastError(), "row-tile pass launch");
    if (ev) hip_check(hipEventRecord(ev[2], stream), "hipEventRecord");
}

template <int KB, bool FP32>
void launch_batch_class(const BatchClass& c, const double* x, double* y, double* zp,
                        hipStream_t stream) {
    const unsigned passes = unsigned((c.max_k + KB - 1) / KB);
    const std::uint32_t count = std::uint32_t(c.member_count);
    if (c.total_units) {
        hipLaunchKernelGGL((right_factor_batch<KB, FP32>),
                           dim3((c.total_units + kUnitWaves - 1) / kUnitWaves, passes),
                           dim3(64 * kUnitWaves), 0, stream, c.members, c.first_unit, count,
                           c.total_units, x, y, zp);
        hip_check(hipGetLastError(), "batched right-factor pass launch");
    }
    hipLaunchKernelGGL((row_tile_batch<KB, FP32>), dim3(c.total_tiles, passes),
                       dim3(64 * kRowWaves), 0, stream, c.members, c.first_tile, count, x, y, zp);
    hip_check(hipGetLastError(), "batched row-tile pass launch");
}

template <bool FP32>
void dispatch_batch_class(const BatchClass& c, const double* x, double* y, double* zp,
                          hipStream_t st) {
    switch (c.kb) {
        case 1: return launch_batch_class<1, FP32>(c, x, y, zp, st);
        case 2: return launch_batch_class<2, FP32>(c, x, y, zp, st);
        case 4: return launch_batch_class<4, FP32>(c, x, y, zp, st);
        case 8: return launch_batch_class<8, FP32>(c, x, y, zp, st);
        default: return launch_batch_class<16, FP32>(c, x, y, zp, st);
    }
}

template <bool FP32>
void dispatch_rhs(const DeviceStore& s, const double* x, std::int64_t ldx, int k, double* y,
                  std::int64_t ldy, double* zp, hipStream_t st) {
    if (k <= 1) return launch_apply<1, FP32>(s, x, ldx, k, y, ldy, zp, st);
    if (k <= 2) return launch_apply<2, FP32>(s, x, ldx, k, y, ldy, zp, st);
    if (k <= 4) return launch_apply<4, FP32>(s, x, ldx, k, y, ldy, zp, st);
    if (k <= 8) return launch_apply<8, FP32>(s, x, ldx, k, y, ldy, zp, st);
    return launch_apply<16, FP32>(s, x, ldx, k, y, ldy, zp, st);
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

// Relative cost of a row-tile entry, used to balance the waves of a tile:
// a fixed latency share plus bytes loaded per lane plus decode work.
constexpr std::uint64_t kEntryCost = 40;

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

}  // namespace

void append_codes(const std::uint32_t* stream, std::uint64_t stream_words_count,
                  std::uint64_t first, std::uint64_t count, int bits,
                  std::vector<std::uint32_t>& out) {
    const std::uint64_t bit0 = first * std::uint64_t(bits);
    const std::uint64_t nbits = count * std::uint64_t(bits);
    const std::uint64_t nw = (nbits + 31) / 32;
    for (std::uint64_t o = 0; o < nw; ++o) {
        const std::uint64_t src = bit0 + 32 * o;
        const std::uint64_t w = src >> 5;
        const unsigned sh = unsigned(src & 31);
        std::uint64_t pair = stream[w];
        if (sh && w + 1 < stream_words_count) pair |= std::uint64_t(stream[w + 1]) << 32;
        std::uint32_t v = std::uint32_t(pair >> sh);
        const std::uint64_t left = nbits - 32 * o;
        if (left < 32) v &= (std::uint32_t{1} << left) - 1u;
        out.push_back(v);
    }
}

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
    Plan* s = &plan;
    plan.order = n;
    std::vector<std::uint32_t> words;
    std::vector<std::int32_t> widths;
    std::vector<double> lsc, rsc, dvals;
    std::vector<BlockDev> blocks;
    std::vector<DenseDev> denses;
    std::vector<UnitDev> units;
    std::vector<std::vector<EntryDev>> per_tile(row_tiles.size());
    std::vector<std::vector<std::uint64_t>> per_cost(row_tiles.size());
    s->block_to_lowrank.assign(d.nblocks, -1);
    std::uint64_t codes = 0, lr_flops = 0, dn_flops = 0, z_rows = 0;
    std::uint64_t n_lowrank_all = 0, left_words = 0, right_words = 0;

    auto pad4 = [&words] {
        while (words.size() & 3u) words.push_back(0u);
    };
    for (std::uint64_t b = 0; b < d.nblocks; ++b) {
        const std::uint64_t r0 = d.rect[4 * b], r1 = d.rect[4 * b + 1];
        const std::uint64_t c0 = d.rect[4 * b + 2], c1 = d.rect[4 * b + 3];
        const std::uint64_t m = r1 - r0, nc = c1 - c0;
        const std::size_t t0 = first_tile(row_tiles, r0);
        if (d.is_dense[b]) {
            require(d.dense && d.dense[b], "dense block without values");
            const std::uint32_t di = std::uint32_t(denses.size());
            denses.push_back({std::uint32_t(r0), std::uint32_t(c0), std::uint32_t(m),
                              std::uint32_t(nc), dvals.size()});
            const double* v = d.dense[b];
            const std::size_t at = dvals.size();
            dvals.resize(at + m * nc);
            for (std::uint64_t i = 0; i < m; ++i)  // row-major in, column-major out
                for (std::uint64_t c = 0; c < nc; ++c) dvals[at + c * m + i] = v[i * nc + c];
            for (std::size_t t = t0; t < row_tiles.size() && row_tiles[t].first < r1; ++t)
                for (std::uint64_t c = 0; c < nc; c += kDenseCols) {
                    const std::uint64_t cnt = std::min<std::uint64_t>(kDenseCols, nc - c);
                    per_tile[t].push_back({kDenseBit | di, std::uint32_t(row_tiles[t].first - r0),
                                           std::uint32_t(c), std::uint32_t(cnt), 0u});
                    per_cost[t].push_back(kEntryCost + 8 * cnt);  // 8 bytes per lane per column
                }
            dn_flops += 2 * m * nc;
            continue;
        }
        ++n_lowrank_all;
        const std::uint64_t r = d.rank[b];
        require(r <= std::min(m, nc), "factor dimensions do not match block rank");
        if (r == 0) continue;  // contributes nothing
        require(d.widths && d.widths[b] && d.lscale && d.lscale[b] && d.rscale && d.rscale[b] &&
                    d.lwords && d.lwords[b] && d.rwords && d.rwords[b],
                "low-rank block without factors");
        const std::uint32_t li = std::uint32_t(blocks.size());
        BlockDev B{std::uint32_t(r), std::uint32_t(widths.size()), std::uint32_t(z_rows), 0};
        std::vector<std::uint64_t> loff(r), roff(r), lcnt(r), rcnt(r);
        std::uint64_t lw = 0, rw = 0;
        for (std::uint64_t j = 0; j < r; ++j) {
            const int bits = d.widths[b][j];
            require(bits >= 2 && bits <= 32, "bit width must be in [2, 32]");
            widths.push_back(bits);
            lsc.push_back(d.lscale[b][j]);
            rsc.push_back(d.rscale[b][j]);
            loff[j] = lw;
            roff[j] = rw;
            lcnt[j] = (m * bits + 31) / 32;
            rcnt[j] = (nc * bits + 31) / 32;
            lw += lcnt[j];
            rw += rcnt[j];
        }
        left_words += lw;
        right_words += rw;
        // left factor: one chunk per row tile
        s->left_chunks.emplace_back();
        for (std::size_t t = t0; t < row_tiles.size() && row_tiles[t].first < r1; ++t) {
            const std::uint64_t a = row_tiles[t].first - r0;
            const std::uint64_t rows = row_tiles[t].second - row_tiles[t].first;
            const std::uint64_t at = words.size();
            for (std::uint64_t j = 0; j < r; ++j)
                append_codes(d.lwords[b] + loff[j], lcnt[j], a, rows, d.widths[b][j], words);
            const std::uint64_t chunk_words = words.size() - at;
            pad4();
            s->left_chunks.back().push_back({std::uint32_t(a), std::uint32_t(rows), at});
            require(chunk_words + 8 < (std::uint64_t{1} << 32), "block chunk too large");
            per_tile[t].push_back({li, std::uint32_t(a), std::uint32_t(chunk_words), 0u, at});
            per_cost[t].push_back(kEntryCost + 4 * r + chunk_words / 16);  // bytes per lane + decode
        }
        // right factor: one chunk (and one work unit) per column tile
        s->right_chunks.emplace_back();
        for (std::size_t t = first_tile(col_tiles, c0);
             t < col_tiles.size() && col_tiles[t].first < c1; ++t) {
            const std::uint64_t a = col_tiles[t].first - c0;
            const std::uint64_t cols = col_tiles[t].second - col_tiles[t].first;
            const std::uint64_t at = words.size();
            for (std::uint64_t j = 0; j < r; ++j)
                append_codes(d.rwords[b] + roff[j], rcnt[j], a, cols, d.widths[b][j], words);
            const std::uint64_t chunk_words = words.size() - at;
            require(chunk_words + 8 < (std::uint64_t{1} << 32), "block chunk too large");
            pad4();
            s->right_chunks.back().push_back({std::uint32_t(a), std::uint32_t(cols), at});
            units.push_back({li, std::uint32_t(col_tiles[t].first), std::uint32_t(cols),
                             std::uint32_t(z_rows + B.nseg * r), std::uint32_t(chunk_words), 0u,
                             at});
            ++B.nseg;
        }
        z_rows += std::uint64_t(B.nseg) * r;
        s->block_to_lowrank[b] = li;
        blocks.push_back(B);
        codes += (m + nc) * r;
        lr_flops += 2 * r * (m + nc);
    }
    require(z_rows < (std::uint64_t{1} << 31), "store too large for one device plan");
    const std::uint64_t stored_words = words.size();
    words.resize(words.size() + 8, 0u);  // reads one word past the last chunk

    std::vector<TileDev> tiles(row_tiles.size());
    std::vector<EntryDev> entries;
    for (std::size_t t = 0; t < row_tiles.size(); ++t) {
        TileDev& T = tiles[t];
        T.row0 = std::uint32_t(row_tiles[t].first);
        T.rows = std::uint32_t(row_tiles[t].second - row_tiles[t].first);
        const auto& list = per_tile[t];
        const auto& cost = per_cost[t];
        const std::uint32_t base = std::uint32_t(entries.size());
        entries.insert(entries.end(), list.begin(), list.end());
        // Contiguous split of the entries into kRowWaves runs of similar cost.
        std::uint64_t total = 0;
        for (auto c : cost) total += c;
        std::uint32_t e = 0;
        std::uint64_t run = 0;
        T.split[0] = base;
        for (int w = 1; w < kRowWaves; ++w) {
            const std::uint64_t goal = total * std::uint64_t(w) / kRowWaves;
            while (e < list.size() && run + cost[e] / 2 < goal) run += cost[e++];
            T.split[w] = base + e;
        }
        T.split[kRowWaves] = base + std::uint32_t(list.size());
    }

    qh_store_stats& st = plan.stats;
    st.order = n;
    st.nblocks = d.nblocks;
    st.ndense = denses.size();
    st.nlowrank = n_lowrank_all;
    st.ncomponents = widths.size();
    st.code_words = left_words + right_words;
    st.left_code_words = left_words;
    st.right_code_words = right_words;
    st.codes = codes;
    st.dense_values = dvals.size();
    st.lowrank_flops_per_rhs = lr_flops;
    st.dense_flops_per_rhs = dn_flops;
    st.tiles = tiles.size();
    st.items = units.size();
    st.entries = entries.size();

    plan.perm = std::move(perm);
    plan.words = std::move(words);
    plan.widths = std::move(widths);
    plan.lscale = std::move(lsc);
    plan.rscale = std::move(rsc);
    plan.dense = std::move(dvals);
    plan.blocks = std::move(blocks);
    plan.denses = std::move(denses);
    plan.units = std::move(units);
    plan.tiles = std::move(tiles);
    plan.entries = std::move(entries);
    plan.z_rows = z_rows;
    plan.stored_words = stored_words;
    return plan;
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
        s->lscale = upload(plan.lscale, bytes);
        s->rscale = upload(plan.rscale, bytes);
        s->dense = upload(plan.dense, bytes);
        s->blocks = upload(plan.blocks, bytes);
        s->denses = upload(plan.denses, bytes);
        s->units = upload(plan.units, bytes);
        s->tiles = upload(plan.tiles, bytes);
        s->entries = upload(plan.entries, bytes);
        s->device_bytes = bytes;
    } catch (...) {
        destroy_device_store(s);
        throw;
    }
    s->n_words = plan.stored_words;
    s->n_comp = plan.widths.size();
    s->n_dense_values = plan.dense.size();
    s->n_lowrank = std::uint32_t(plan.blocks.size());
    s->n_dense = std::uint32_t(plan.denses.size());
    s->n_units = std::uint32_t(plan.units.size());
    s->n_tiles = std::uint32_t(plan.tiles.size());
    s->n_entries = std::uint32_t(plan.entries.size());
    s->z_rows = plan.z_rows;
    s->block_to_lowrank = std::move(plan.block_to_lowrank);
    s->left_chunks = std::move(plan.left_chunks);
    s->right_chunks = std::move(plan.right_chunks);
    for (const auto& B : plan.blocks) {
        s->comp_base.push_back(B.comp_base);
        s->ranks.push_back(B.rank);
    }
    s->stats = plan.stats;
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
    std::vector<double> zp(pl.z_rows * std::uint64_t(k), 0.0);
    auto code_at = [&](std::uint64_t seg_word, std::uint32_t idx, std::uint32_t bits) {
        const std::uint32_t pos = idx * bits;
        const std::uint64_t w = seg_word + (pos >> 5);
        const std::uint64_t pair = std::uint64_t(at(pl.words, w)) |
                                   (std::uint64_t(at(pl.words, w + 1)) << 32);
        const std::uint64_t raw = (pair >> (pos & 31u)) & ((std::uint64_t{1} << bits) - 1);
        const std::uint64_t sign = std::uint64_t{1} << (bits - 1);
        return double(std::int64_t(raw ^ sign) - std::int64_t(sign));
    };
    // Pass 1: every unit, the components in passes of kPassComponents.
    for (const UnitDev& u : pl.units) {
        const BlockDev& B = at(pl.blocks, u.block);
        if (u.cols == 0 || u.cols > std::uint32_t(kTile)) throw std::out_of_range("unit width");
        std::uint64_t all = 0;
        for (std::uint32_t j = 0; j < B.rank; ++j)
            all += stream_words(u.cols, at(pl.widths, B.comp_base + j));
        if (all != u.words) throw std::out_of_range("unit word count");
        if (u.words + 5 <= std::uint32_t(kStageWords))  // staged whole, one word past the end
            at(pl.words, (u.chunk & ~std::uint64_t(3)) + 4 * ((u.words + (u.chunk & 3) + 4) >> 2) - 1);
        std::uint64_t done = 0;
        for (std::uint32_t j0 = 0; j0 < B.rank; j0 += kPassComponents) {
            const std::uint32_t jn = std::min<std::uint32_t>(kPassComponents, B.rank - j0);
            std::uint64_t total = 0;
            for (std::uint32_t g = 0; g < jn; ++g)
                total += stream_words(u.cols, at(pl.widths, B.comp_base + j0 + g));
            const std::uint64_t first = u.chunk + done;
            const std::uint64_t delta = first & 3u;
            const std::uint64_t quads = (total + delta + 1 + 3) >> 2;
            if (4 * quads > std::uint64_t(kStageWords + 8)) throw std::out_of_range("stage size");
            at(pl.words, first - delta + 4 * quads - 1);  // staged range is readable
            std::uint64_t seg = 0;
            for (std::uint32_t g = 0; g < jn; ++g) {
                const std::uint32_t j = j0 + g;
                const std::uint32_t bits = at(pl.widths, B.comp_base + j);
                for (std::int64_t p = 0; p < k; ++p) {
                    double sum = 0.0;
                    for (std::uint32_t c = 0; c < u.cols; ++c)
                        sum += code_at(first + seg, c, bits) *
                               x[std::int64_t(at(pl.perm, u.col0 + c)) * k + p];
                    zp.at((std::uint64_t(u.zrow) + j) * std::uint64_t(k) + p) = sum;
                }
                seg += stream_words(u.cols, bits);
            }
            done += total;
        }
    }
    // Pass 2: every tile, every wave's entries.
    for (const TileDev& T : pl.tiles) {
        if (T.rows == 0 || T.rows > std::uint32_t(kTile)) throw std::out_of_range("tile rows");
        for (std::int64_t p = 0; p < k; ++p)
            for (std::uint32_t i = 0; i < T.rows; ++i) {
                double acc = 0.0;
                for (int w = 0; w < kRowWaves; ++w)
                    for (std::uint32_t e = T.split[w]; e < T.split[w + 1]; ++e) {
                        const EntryDev& ent = at(pl.entries, e);
                        if (ent.what & kDenseBit) {
                            const DenseDev& D = at(pl.denses, ent.what & ~kDenseBit);
                            if (ent.cols == 0 || ent.cols > std::uint32_t(kDenseCols) ||
                                ent.words + ent.cols > D.n)
                                throw std::out_of_range("dense entry columns");
                            for (std::uint32_t c = ent.words; c < ent.words + ent.cols; ++c)
                                acc += at(pl.dense, D.off + std::uint64_t(c) * D.m + ent.row + i) *
                                       x[std::int64_t(at(pl.perm, D.col0 + c)) * k + p];
                            continue;
                        }
                        const BlockDev& B = at(pl.blocks, ent.what);
                        std::uint64_t all = 0;
                        for (std::uint32_t j = 0; j < B.rank; ++j)
                            all += stream_words(T.rows, at(pl.widths, B.comp_base + j));
                        if (all != ent.words) throw std::out_of_range("entry word count");
                        at(pl.words, (ent.chunk & ~std::uint64_t(3)) +
                                         4 * ((ent.words + (ent.chunk & 3) + 4) >> 2) - 1);
                        std::uint64_t off = ent.chunk;
                        for (std::uint32_t j = 0; j < B.rank; ++j) {
                            const std::uint32_t cj = B.comp_base + j;
                            const std::uint32_t bits = at(pl.widths, cj);
                            double z = 0.0;
                            for (std::uint32_t sg = 0; sg < B.nseg; ++sg)
                                z += zp.at((std::uint64_t(B.zbase) + sg * B.rank + j) * k + p);
                            acc += code_at(off, i, bits) * (at(pl.lscale, cj) * at(pl.rscale, cj) * z);
                            off += stream_words(T.rows, bits);
                        }
                    }
                y[std::int64_t(at(pl.perm, T.row0 + i)) * k + p] = acc;
            }
    }
    (void)n;
}

void destroy_device_store(DeviceStore* s) {
    if (!s) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(s->device);
    for (void* e : s->profile.events) (void)hipEventDestroy(static_cast<hipEvent_t>(e));
    for (void* p : {static_cast<void*>(s->perm), static_cast<void*>(s->words),
                    static_cast<void*>(s->widths), static_cast<void*>(s->lscale),
                    static_cast<void*>(s->rscale), static_cast<void*>(s->dense),
                    static_cast<void*>(s->blocks), static_cast<void*>(s->denses),
                    static_cast<void*>(s->units), static_cast<void*>(s->tiles),
                    static_cast<void*>(s->entries)})
        if (p) (void)hipFree(p);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete s;
}

std::uint64_t apply_workspace_bytes(const DeviceStore& s, std::int64_t k) {
    return s.z_rows * std::uint64_t(k) * sizeof(double);
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
    const std::uint64_t need = apply_workspace_bytes(s, k);
    require(workspace == nullptr || workspace_bytes >= need, "workspace is too small");
    DeviceScope scope(s.device);
    hipStream_t st = static_cast<hipStream_t>(stream);
    double* zp = static_cast<double*>(workspace);
    bool owned = false;
    if (!zp && need) {
        hip_check(hipMallocAsync(reinterpret_cast<void**>(&zp), need, st), "hipMallocAsync");
        owned = true;
    }
    try {
        if (precision == QH_PRECISION_FP64)
            dispatch_rhs<false>(s, x, ldx, int(k), y, ldy, zp, st);
        else
            dispatch_rhs<true>(s, x, ldx, int(k), y, ldy, zp, st);
    } catch (...) {
        if (owned) (void)hipFreeAsync(zp, st);
        throw;
    }
    if (owned) hip_check(hipFreeAsync(zp, st), "hipFreeAsync");
}

void profile_begin(DeviceStore& s, std::uint64_t max_applies) {
    require(max_applies > 0 && max_applies < (std::uint64_t{1} << 20), "invalid profile size");
    DeviceScope scope(s.device);
    auto& p = s.profile;
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
                           dev, s.widths + s.comp_base[li], s.ranks[li], rows, out);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    (void)hipFree(dev);
    hip_check(e, "decode launch");
}

// ---------------------------------------------------------------------------
// Batches

namespace {
int rhs_class(std::int64_t k) {
    return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : k <= 8 ? 8 : 16;
}
}  // namespace

Batch* create_batch(const DeviceStore* const* stores, const std::int64_t* k, std::uint64_t count) {
    require(count > 0 && stores && k, "empty batch");
    auto* b = new Batch;
    try {
        b->device = stores[0] ? stores[0]->device : 0;
        std::uint64_t off = 0, zoff = 0;
        for (std::uint64_t i = 0; i < count; ++i) {
            require(stores[i] != nullptr, "null store in batch");
            require(stores[i]->device == b->device, "batch members must share one device");
            require(k[i] > 0 && k[i] <= std::int64_t(std::numeric_limits<int>::max()),
                    "right-hand-side shape mismatch");
            b->stores.push_back(stores[i]);
            b->k.push_back(k[i]);
            b->offset.push_back(off);
            b->zoffset.push_back(zoff);
            off += stores[i]->order * std::uint64_t(k[i]);
            zoff += stores[i]->z_rows * std::uint64_t(k[i]);
        }
        b->total = off;
        b->z_total = zoff;
        DeviceScope scope(b->device);
        for (int kb : {1, 2, 4, 8, 16}) {
            BatchClass c;
            c.kb = kb;
            std::vector<MemberView> views;
            std::vector<std::uint32_t> fu{0}, ft{0};
            for (std::uint64_t i = 0; i < count; ++i) {
                if (rhs_class(b->k[i]) != kb) continue;
                const DeviceStore& s = *b->stores[i];
                // x, y and the workspace as element offsets from the call's bases
                views.push_back(member_view(
                    s, reinterpret_cast<const double*>(std::uintptr_t(b->offset[i])), b->k[i],
                    int(b->k[i]), reinterpret_cast<double*>(std::uintptr_t(b->offset[i])),
                    b->k[i], reinterpret_cast<double*>(std::uintptr_t(b->zoffset[i]))));
                require(std::uint64_t(fu.back()) + s.n_units < (std::uint64_t{1} << 31) &&
                            std::uint64_t(ft.back()) + s.n_tiles < (std::uint64_t{1} << 31),
                        "batch too large for one launch");
                fu.push_back(fu.back() + s.n_units);
                ft.push_back(ft.back() + s.n_tiles);
                c.max_k = std::max(c.max_k, b->k[i]);
            }
            if (views.empty()) continue;
            c.member_count = views.size();
            c.total_units = fu.back();
            c.total_tiles = ft.back();
            std::uint64_t bytes = 0;
            c.members = upload(views, bytes);
            c.first_unit = upload(fu, bytes);
            c.first_tile = upload(ft, bytes);
            b->classes.push_back(c);
        }
    } catch (...) {
        destroy_batch(b);
        throw;
    }
    return b;
}

void destroy_batch(Batch* b) {
    if (!b) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->device);
    for (auto& c : b->classes)
        for (void* p : {static_cast<void*>(c.members), static_cast<void*>(c.first_unit),
                        static_cast<void*>(c.first_tile)})
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
    try {
        for (const auto& c : b.classes) {
            if (precision == QH_PRECISION_FP64)
                dispatch_batch_class<false>(c, x, y, zp, st);
            else
                dispatch_batch_class<true>(c, x, y, zp, st);
        }
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
