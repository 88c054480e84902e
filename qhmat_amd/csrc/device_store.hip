// MI355X (gfx950) apply of a qhmat H-matrix, y = H x.
//
// The packed store is uploaded once as a list of work items (layout in
// device_store.hpp).  An apply is three launches on the caller's stream, all
// deterministic (fixed summation order, one writer per output element, no
// floating-point atomics):
//
//  pass 1  four workgroups of four waves per row tile, each wave running one
//          of the tile's 16 item runs: the tile's diagonal (kDense) and
//          fused (kFused: z and U z inside the item) contributions, summed
//          in a fixed tree per workgroup into one of four row-sum sets of
//          the workspace, and a share of the kRight items: partial sums
//          z[c-tile][j] = sum_{c in tile} codeR(c, j) x[c] into the workspace.
//  pass 2  the same shape over the tile's kLeft items (U z with z summed
//          over the block's column tiles), into four more row-sum sets.
//  combine one small workgroup per row tile: the eight sets in a fixed tree,
//          written once to y[perm[row]].
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

// Slot of lane L in word row w of a side block.  Right-factor blocks are
// read by lanes that sit on different rows at the same columns (a lane per
// component), so their rows are swizzled: lane L at L xor 4 (w mod 16),
// which keeps every aligned group of four columns together and puts the
// same columns of different rows on different banks.  Left-factor blocks
// are read a whole row at a time and keep L.
__host__ __device__ __forceinline__ std::uint32_t lane_slot(bool swz, std::uint64_t w,
                                                            std::uint64_t L) {
    return std::uint32_t(swz ? (L ^ (4u * (w & 15u))) : L);
}

// Code of lane `lane` at bit offset `o` of a side block at `side` (global
// memory).  The next row of the block is read only when the code straddles
// into it.
__device__ __forceinline__ int decode_side_code(const std::uint32_t* side, std::uint32_t lane,
                                                std::uint32_t o, std::uint32_t bits, bool swz) {
    const std::uint32_t k = o >> 5, sh = o & 31u;
    const std::uint32_t lo = side[64u * k + lane_slot(swz, k, lane)];
    const std::uint32_t hi =
        sh + bits > 32u ? side[64u * (k + 1u) + lane_slot(swz, k + 1u, lane)] : 0u;
    const std::uint32_t raw = __builtin_amdgcn_alignbit(hi, lo, sh);
    return bits == 32 ? static_cast<int>(raw) : __builtin_amdgcn_sbfe(raw, 0u, bits);
}


// LDS written by some lanes of a wave and read by others: keep the compiler
// from moving accesses across this point and wait for the LDS queue.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive prefix sum over the 64 lanes, all in DPP: shifted adds within
// each 16-lane row (row_shr 1, 2, 4, 8), then row 0's total into row 1 and
// row 2's into row 3 (row_bcast15), then rows 0-1's total into rows 2-3
// (row_bcast31).  A lane without a source adds the zero passed as `old`.
__device__ __forceinline__ std::uint32_t wave_scan(std::uint32_t v, int) {
    int x = int(v);
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false);
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false);
    return std::uint32_t(x);
}

// 16-byte group as a native clang vector (an array of HIP's union-based
// uint4 would be kept in private scratch memory by the compiler).
typedef std::uint32_t quad_t __attribute__((ext_vector_type(4)));

template <bool FP32> using real_t = typename std::conditional<FP32, float, double>::type;

// LDS of one wave: a head word that decoding may read below the first
// stream, the current item's data plus slack that lanes decoding past the end
// of a stream may read (64 codes of 32 bits and the word after), then the
// item's inputs (x rows of its columns, or the workspace partial sums of
// kLeft), KB values per lane, which also take the item's rows at its end,
// then the scaled inputs of 16 components (zt), then the item's scale
// products and widths (lanes = components).
constexpr int kStageBytes = kItemQuads * 16 + 272;
constexpr int kLdsHead = 256;  // readable bytes before the stage (decode_lm)
template <int KB> constexpr int wave_lds_meta() {
    return kLdsHead + kStageBytes + kTile * KB * 8 + 16 * KB * 8;
}
// Narrow passes in FP32 mode also keep the x rows in FP32 (lanes = columns).
template <int KB> constexpr int wave_lds_xf() { return wave_lds_meta<KB>() + 64 * 12; }
template <int KB> constexpr int wave_lds_bytes() {
    return wave_lds_xf<KB>() + (KB < 8 ? 64 * KB * 4 : 0);
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

// Loads go through buffer resources (V#) sized to the bytes a request may
// touch: lanes outside an item (or past its columns / components) use an
// out-of-range offset and read zero, so no load is predicated and every
// load of an item is issued, whatever its shape (the compiler then counts
// the loads in flight exactly, which lets two items be in flight).
constexpr int kBufferFlags = 0x00020000;      // dword 3 of a raw 32-bit buffer resource
constexpr std::uint32_t kOutOfRange = 0x7FFFFFF0u;  // >= every num_records used here

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer(const void* base, std::uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(bytes), kBufferFlags);
}
__device__ __forceinline__ double load_f64(__amdgpu_buffer_rsrc_t r, std::uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0);
    return __longlong_as_double(static_cast<long long>(
        (static_cast<unsigned long long>(v[1]) << 32) | static_cast<unsigned long long>(v[0])));
}

// Row sets per tile in the workspace (four workgroups per pass), and x in
// clustered order after them: xc[i][p] = x[perm[i]][p], written once per
// apply (permute_tile) so that an item's x rows are one contiguous run.
constexpr int kRowSets = 2 * kPass1Groups;
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

// Matrix cores.  v_mfma_f32_16x16x4f32 and v_mfma_f64_16x16x4f64 form
// D += A B with A 16 x 4 and B 4 x 16: lane l = 16 q + r supplies A[r][q] and
// B[q][r], and slot t of its result holds D[4q + t][r] (FP32) or D[4t + q][r]
// (FP64).  In FP64 mode lane r supplies row a_row(r) of A instead, which puts
// D[4q + t][r] in slot t as well, so both modes share one result layout.
// Each result element is its own dot product: column r of D depends on
// column r of B only, so a right-hand side's arithmetic does not depend on
// the others in its pass.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

// Right-hand sides per pass from which the matrix cores take the codes (below,
// one code's products are too few to fill them and the vector path is faster).
constexpr int kMatrixMinKB = 8;

template <bool FP32> struct Mma;
template <> struct Mma<true> {
    using acc = f32x4;
    static __device__ __forceinline__ acc zero() { return acc{0.0f, 0.0f, 0.0f, 0.0f}; }
    static __device__ __forceinline__ acc mad(float a, float b, acc d) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, d, 0, 0, 0);
    }
};
template <> struct Mma<false> {
    using acc = f64x4;
    static __device__ __forceinline__ acc zero() { return acc{0.0, 0.0, 0.0, 0.0}; }
    static __device__ __forceinline__ acc mad(double a, double b, acc d) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, d, 0, 0, 0);
    }
};
template <bool FP32> __host__ __device__ __forceinline__ int a_row(int r) {
    return FP32 ? r : 4 * (r & 3) + (r >> 2);
}

// The code of lane `lane` for the component with packed fields m in a staged
// side block (word w of lane L at side[64 w + L]; one row of 64 readable
// words is kept before the stage and after the item): the 32 bits [E - 32,
// E) of the lane's stream from words (E >> 5) - 1 and E >> 5
// (v_alignbit_b32), shifted arithmetically right by 32 - b, which
// sign-extends the field (b = 32 included) and drops the bits below it.  No
// mask and no branch; every lane reads its own column of words, so the 64
// lanes of a read hit 64 consecutive words.
__device__ __forceinline__ int decode_lm(const std::uint32_t* side, std::uint32_t m, int lane) {
    const std::uint32_t E = m >> 6;
    const std::uint32_t* w = side + 64u * (E >> 5) + std::uint32_t(lane);
    return static_cast<int>(__builtin_amdgcn_alignbit(w[0], w[-64], E)) >> (32u - (m & 63u));
}

// Codes of columns c4 .. c4 + 3 (c4 a multiple of 4) for the component with
// packed fields m in a staged right side block, whose rows are swizzled
// (lane_slot): the four columns stay consecutive within each of the two
// rows read, so each row is one aligned read of four words per lane.
__device__ __forceinline__ void decode_right4(const std::uint32_t* side, std::uint32_t m, int c4,
                                              int (&code)[4]) {
    const int E = int(m >> 6), kk = E >> 5;
    const std::uint32_t sh = 32u - (m & 63u);
    const std::uint32_t* hi = side + 64 * kk + (c4 ^ (4 * (kk & 15)));
    const std::uint32_t* lo = side + 64 * (kk - 1) + (c4 ^ (4 * ((kk - 1) & 15)));
#pragma unroll
    for (int u = 0; u < 4; ++u)
        code[u] = static_cast<int>(__builtin_amdgcn_alignbit(hi[u], lo[u], std::uint32_t(E))) >> sh;
}

// A code as h 2^16 + l with l its sign-extended low 16 bits: both halves are
// exact in FP32, and a code of small magnitude stays in l (no cancellation
// between the halves).
__device__ __forceinline__ void split_code(int code, float& h, float& l) {
    // (code - l) / 2^16 without leaving 32 bits: the arithmetic high half,
    // plus one when the low half is negative
    h = float((code >> 16) + ((code >> 15) & 1));
    l = float(static_cast<int>(__builtin_amdgcn_sbfe(code, 0u, 16u)));
}

// ---------------------------------------------------------------------------
// Narrow passes (KB < kMatrixMinKB): vector ALUs.  Lanes are the item's rows
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


// ---------------------------------------------------------------------------
// Wide passes (KB >= kMatrixMinKB): matrix cores.

// Per-item rows on the matrix cores: slot t of d[rb] in lane 16q + r holds
// row 16 rb + 4q + t of the item's tile for right-hand side r mod KB.
template <bool FP32> struct ItemRows {
    typename Mma<FP32>::acc d[4];
    __device__ __forceinline__ ItemRows() {
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) d[rb] = Mma<FP32>::zero();
    }
};

// Left-factor rows of components [g0, g0 + 16) of an item (g0 a multiple of
// 16) on the matrix cores, one MFMA per 16 rows x 4 components: lane 16q + r
// decodes row 16 rb + a_row(r) of component g0 + 4s + q and supplies that
// component's scaled input for right-hand side r mod KB from the LDS table
// zt (zt[(c - g0) KB + p], zero past the item's components, whose dummy
// streams decode in bounds).  Rows past the tile's decode neighbouring bits,
// and only their own results, which are never used.  WIDE (FP32 mode, an
// item with a code wider than 24 bits): every code is split (split_code), h
// against 2^16 zs and l against zs, into the same sums.
template <int KB, bool FP32, bool WIDE>
__device__ __forceinline__ void left_group(const std::uint32_t* stage, const real_t<FP32>* zt,
                                           std::uint32_t lmeta_l, std::uint32_t g0, int lane,
                                           ItemRows<FP32>& rows) {
    using Mm = Mma<FP32>;
    using Real = real_t<FP32>;
    const int q = lane >> 4, r = lane & 15;
    const int row = a_row<FP32>(r);
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const std::uint32_t c = g0 + 4u * std::uint32_t(s) + std::uint32_t(q);  // <= 63
        const std::uint32_t meta = std::uint32_t(__shfl(int(lmeta_l), int(c), 64));
        const Real zs = zt[(4 * s + q) * KB + (r & (KB - 1))];
        if constexpr (!WIDE) {
#pragma unroll
            for (int rb = 0; rb < 4; ++rb)
                rows.d[rb] = Mm::mad(Real(decode_lm(stage, meta, 16 * rb + row)), zs, rows.d[rb]);
        } else {
            const float zh = float(zs) * 65536.0f;
#pragma unroll
            for (int rb = 0; rb < 4; ++rb) {
                float h, l;
                split_code(decode_lm(stage, meta, 16 * rb + row), h, l);
                rows.d[rb] = Mma<true>::mad(h, zh, rows.d[rb]);
                rows.d[rb] = Mma<true>::mad(l, float(zs), rows.d[rb]);
            }
        }
    }
}

// Diagonal-block columns on the FP64 matrix cores: lane 16q + r supplies
// value (row 16 rb + a_row(r), column 4s + q) of the staged column-major
// R x cnt block (zero for columns past cnt, whose x is zero too) and x of
// that column for right-hand side r mod KB.
template <int KB>
__device__ __forceinline__ void dense_rows(const double* dv, const double* xs, std::uint32_t R,
                                           std::uint32_t cnt, int lane, ItemRows<false>& rows) {
    const int q = lane >> 4, r = lane & 15;
    const std::uint32_t row = std::uint32_t(a_row<false>(r));
#pragma unroll
    for (int s = 0; s < kDenseColumns / 4; ++s) {
        const std::uint32_t c = 4u * std::uint32_t(s) + std::uint32_t(q);
        const double xv = xs[c * KB + (r & (KB - 1))];
#pragma unroll
        for (int rb = 0; rb < 4; ++rb) {
            const double a = c < cnt ? dv[c * R + 16u * std::uint32_t(rb) + row] : 0.0;
            rows.d[rb] = Mma<false>::mad(a, xv, rows.d[rb]);
        }
    }
}

// Adds an item's rows to the wave's row sums (lanes = rows) through the LDS
// tile T[row][KB] (the item's input area, no longer read by then), in the
// precision the rows were summed in.
template <int KB, bool FP32>
__device__ __forceinline__ void add_item_rows(const ItemRows<FP32>& rows, void* tile, int lane,
                                              double (&acc)[KB]) {
    using Real = real_t<FP32>;
    Real* T = static_cast<Real*>(tile);
    const int q = lane >> 4, r = lane & 15;
    if (r < KB) {
#pragma unroll
        for (int rb = 0; rb < 4; ++rb)
#pragma unroll
            for (int t = 0; t < 4; ++t) T[(16 * rb + 4 * q + t) * KB + r] = rows.d[rb][t];
    }
    wave_lds_sync();
#pragma unroll
    for (int p = 0; p < KB; ++p) acc[p] += double(T[lane * KB + p]);
}

// Right-factor sums of components [g0, g0 + 16) of a kRight or kFused item
// on the matrix cores, one MFMA per 4 x 16 codes: z[t] in lane 16q + r is
//   sum over the item's 64 columns c of code(c, g0 + 4q + t) x[c][r mod KB]
// in FP64.  MFMA m (0..15) puts columns 16q + m in its k slots: each lane
// decodes its own component's code of column 16q + m (stepping the end bit
// by the width) and reads that column's x from LDS.  Columns past the item's
// hold zero x; components past its count decode the dummy stream and their
// sums are never used.  WIDE (FP32 mode, an item with a code wider than 24
// bits): every code is split into exact halves (split_code) summed apart and
// joined in FP64.  Both depend only on the item, so a component's rounding
// is the same for every right-hand-side count.
template <int KB, bool FP32, bool WIDE>
__device__ __forceinline__ void right_group(const std::uint32_t* stage, const double* xs,
                                            std::uint32_t rmeta_l, std::uint32_t g0, int lane,
                                            double (&z)[4]) {
    using Mm = Mma<FP32>;
    using Real = real_t<FP32>;
    const int q = lane >> 4, r = lane & 15;
    const std::uint32_t j = g0 + std::uint32_t(a_row<FP32>(r));  // <= 63
    const std::uint32_t meta = std::uint32_t(__shfl(int(rmeta_l), int(j), 64));
    const double* xcol = xs + 16 * q * KB + (r & (KB - 1));
    if constexpr (!WIDE) {
        typename Mm::acc d = Mm::zero();
#pragma unroll 2
        for (int t = 0; t < 4; ++t) {
            int code[4];
            decode_right4(stage, meta, 16 * q + 4 * t, code);
#pragma unroll
            for (int u = 0; u < 4; ++u) d = Mm::mad(Real(code[u]), Real(xcol[(4 * t + u) * KB]), d);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) z[t] = double(d[t]);
    } else {
        f32x4 dh = Mma<true>::zero(), dl = Mma<true>::zero();
#pragma unroll 2
        for (int t = 0; t < 4; ++t) {
            int code[4];
            decode_right4(stage, meta, 16 * q + 4 * t, code);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                float h, l;
                split_code(code[u], h, l);
                const float xv = float(xcol[(4 * t + u) * KB]);
                dh = Mma<true>::mad(h, xv, dh);
                dl = Mma<true>::mad(l, xv, dl);
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) z[t] = double(dh[t]) * 65536.0 + double(dl[t]);
    }
}

// kFused components [g0, g0 + 16): right-factor sums, their scaled inputs
// through the LDS table zt, and the left streams of the same components.
template <int KB, bool FP32, bool WIDE>
__device__ __forceinline__ void fused_group(const Item& f, const std::uint32_t* stage,
                                            const std::uint32_t* lstage, const double* xs,
                                            real_t<FP32>* zt, std::uint32_t rmeta_l,
                                            double scale_l, std::uint32_t g0, int lane,
                                            ItemRows<FP32>& rows) {
    using Real = real_t<FP32>;
    const int q = lane >> 4, r = lane & 15;
    double z[4];
    right_group<KB, FP32, WIDE>(stage, xs, rmeta_l, g0, lane, z);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const std::uint32_t c = g0 + 4u * std::uint32_t(q) + std::uint32_t(t);  // <= 63
        const double sc = __shfl(scale_l, int(c), 64);
        if (r < KB) zt[(4 * q + t) * KB + r] = c < f.count ? Real(sc * z[t]) : Real(0);
    }
    wave_lds_sync();
    left_group<KB, FP32, WIDE>(lstage, zt, rmeta_l, g0, lane, rows);
    wave_lds_sync();  // the next group rewrites zt
}

// kRight components [g0, g0 + 16): right-factor sums into the workspace rows
// of the block's column tile.
template <int KB, bool FP32, bool WIDE>
__device__ __forceinline__ void right_rows(const MemberView& M, const Item& f,
                                           const std::uint32_t* stage, const double* xs,
                                           std::uint32_t rmeta_l, std::uint32_t g0, int lane,
                                           int p0, int kk) {
    const int q = lane >> 4, r = lane & 15;
    double z[4];
    right_group<KB, FP32, WIDE>(stage, xs, rmeta_l, g0, lane, z);
    if (r < kk) {
        double* out = M.zp + (std::uint64_t(f.zrow) + g0 + 4u * std::uint32_t(q)) *
                                 std::uint64_t(M.k) + p0 + r;
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (g0 + 4u * std::uint32_t(q) + std::uint32_t(t) < f.count)
                out[std::uint64_t(t) * std::uint64_t(M.k)] = z[t];
    }
}

// Component widths and scale products of the staged item (lanes =
// components), read back from LDS after the staging.
struct Staged {
    std::uint32_t bits_l;
    double scale_l;
};

// Moves a fetched item into the wave's LDS: its data, its x rows (lanes =
// columns; zero past the item's columns) or kLeft partial sums (lanes =
// pairs; zero past them), and its widths and scale products.  Every
// register of the fetch buffer is dead afterwards, so the next fetch into it
// needs no copy of a load still in flight.
template <int KB>
__device__ __forceinline__ void stage_item(const Item& f, const Fetch<KB>& F, int lane,
                                           std::uint8_t* wave_lds) {
    quad_t* stage_q = reinterpret_cast<quad_t*>(wave_lds + kLdsHead);
    double* xs = reinterpret_cast<double*>(wave_lds + kLdsHead + kStageBytes);
#pragma unroll
    for (int t = 0; t < kItemQuads / 64; ++t)
        if (std::uint32_t(lane + 64 * t) < f.quads) stage_q[lane + 64 * t] = F.q[t];
#pragma unroll
    for (int p = 0; p < KB; ++p) xs[lane * KB + p] = F.in[p];
    reinterpret_cast<double*>(wave_lds + wave_lds_meta<KB>())[lane] = F.scale;
    reinterpret_cast<std::uint32_t*>(wave_lds + wave_lds_meta<KB>() + 512)[lane] = F.bits;
}
template <int KB>
__device__ __forceinline__ Staged staged_meta(const std::uint8_t* wave_lds, int lane) {
    return Staged{reinterpret_cast<const std::uint32_t*>(wave_lds + wave_lds_meta<KB>() + 512)[lane],
                  reinterpret_cast<const double*>(wave_lds + wave_lds_meta<KB>())[lane]};
}

// A kFused, kRight (pass 1) or kLeft (pass 2) item, 16 components at a
// time.  Row contributions are formed on the matrix cores (in FP32 for the
// codes in FP32 mode) and added to acc (lanes = rows) in FP64.
template <int KB, bool FP32, int PASS, bool WIDE>
__device__ __forceinline__ void lowrank_item(const MemberView& M, const Item& f,
                                             const Staged& st, int lane, int p0, int kk,
                                             std::uint32_t R, std::uint8_t* wave_lds,
                                             double (&acc)[KB]) {
    using Real = real_t<FP32>;
    const std::uint32_t* stage = reinterpret_cast<const std::uint32_t*>(wave_lds + kLdsHead);
    double* xs = reinterpret_cast<double*>(wave_lds + kLdsHead + kStageBytes);
    Real* zt = reinterpret_cast<Real*>(wave_lds + kLdsHead + kStageBytes + kTile * KB * 8);
    const bool comp = std::uint32_t(lane) < f.count;
    const std::uint32_t bits_l = st.bits_l;
    const double scale_l = st.scale_l;
    ItemRows<FP32> rows;
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
        std::uint32_t total;
        const std::uint32_t lmeta_l = side_meta(comp, bits_l, lane, total);
        for (std::uint32_t g0 = 0; g0 < f.count; g0 += 16) {
            if (std::uint32_t(lane) - g0 < 16u) {
#pragma unroll
                for (int p = 0; p < KB; ++p) zt[(std::uint32_t(lane) - g0) * KB + p] = Real(scale_l * z[p]);
            }
            wave_lds_sync();
            left_group<KB, FP32, WIDE>(stage, zt, lmeta_l, g0, lane, rows);
            wave_lds_sync();  // the next group rewrites zt
        }
    } else {
        std::uint32_t total;
        const std::uint32_t rmeta_l = side_meta(comp, bits_l, lane, total);
        if (f.kind == kRight) {
            for (std::uint32_t g0 = 0; g0 < f.count; g0 += 16)
                right_rows<KB, FP32, WIDE>(M, f, stage, xs, rmeta_l, g0, lane, p0, kk);
            return;
        }
        const std::uint32_t* lstage = stage + side_words(total);
        for (std::uint32_t g0 = 0; g0 < f.count; g0 += 16)
            fused_group<KB, FP32, WIDE>(f, stage, lstage, xs, zt, rmeta_l, scale_l, g0, lane, rows);
    }
    add_item_rows<KB, FP32>(rows, xs, lane, acc);
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
        // x rows in the arithmetic precision (FP32 mode: converted once)
        const Real* xr;
        if constexpr (FP32) {
            float* xf = reinterpret_cast<float*>(wave_lds + wave_lds_xf<KB>());
#pragma unroll
            for (int p = 0; p < KB; ++p) xf[lane * KB + p] = float(xs[lane * KB + p]);
            wave_lds_sync();
            xr = xf;
        } else {
            xr = xs;
        }
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

// Decodes and applies one staged item.  Pass 1 holds kDense, kFused and
// kRight items, pass 2 kLeft items only.  Narrow passes run on the vector
// ALUs, wide ones on the matrix cores; in FP32 mode on the matrix cores an
// item with a code wider than 24 bits takes the split path throughout.
template <int KB, bool FP32, int PASS>
__device__ __forceinline__ void compute_item(const MemberView& M, const Item& f,
                                             const Staged& st, int lane, int p0, int kk,
                                             std::uint32_t R, std::uint8_t* wave_lds,
                                             double (&acc)[KB]) {
    if constexpr (KB < kMatrixMinKB) {
        valu_item<KB, FP32, PASS>(M, f, st, lane, p0, kk, R, wave_lds, acc);
    } else if (PASS == 1 && f.kind == kDense) {
        double* xs = reinterpret_cast<double*>(wave_lds + kLdsHead + kStageBytes);
        ItemRows<false> rows;
        dense_rows<KB>(reinterpret_cast<const double*>(wave_lds + kLdsHead), xs, R,
                       min(f.count, std::uint32_t(kDenseColumns)), lane, rows);
        add_item_rows<KB, false>(rows, xs, lane, acc);
    } else if (FP32 && __ballot(std::uint32_t(lane) < f.count && st.bits_l > 24u) != 0) {
        lowrank_item<KB, FP32, PASS, FP32>(M, f, st, lane, p0, kk, R, wave_lds, acc);
    } else {
        lowrank_item<KB, FP32, PASS, false>(M, f, st, lane, p0, kk, R, wave_lds, acc);
    }
}

// One wave's items: [b1, e1) followed by [b2, e2).  Row contributions
// (kDense, kFused, kLeft) are added to acc (lanes = the R rows of the tile);
// kRight items write their partial sums to the workspace.  Two items are in
// flight while one is decoded: the fetch buffers alternate.
template <int KB, bool FP32, int PASS>
__device__ __forceinline__ void run_items(const MemberView& M, std::uint32_t b1, std::uint32_t e1,
                                          std::uint32_t b2, std::uint32_t e2, int p0, int kk,
                                          std::uint32_t R, std::uint8_t* __restrict__ wave_lds,
                                          double (&acc)[KB]) {
    const int lane = threadIdx.x & 63;
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
        Item fa = slot(0), fb = slot(1), fc = slot(2), fd = slot(3);
        Fetch<KB> FA, FB;
        fetch<KB>(M, fa, lane, p0, kk, FA);
        fetch<KB>(M, fb, lane, p0, kk, FB);

        // Steady state: buffer A holds item s, buffer B item s+1.  Step s
        // waits for item s (fetched two steps earlier) and fetches item s+2
        // into the freed buffer, so two items' data stay in flight during a
        // decode.
        for (std::uint32_t s = 0; s < n; s += 2) {
            {
                stage_item<KB>(fa, FA, lane, wave_lds);
                fetch<KB>(M, fc, lane, p0, kk, FA);
                wave_lds_sync();
                compute_item<KB, FP32, PASS>(M, fa, staged_meta<KB>(wave_lds, lane), lane, p0, kk,
                                             R, wave_lds, acc);
                wave_lds_sync();  // the next item overwrites the stage
                fa = fc;
                fc = slot(s + 4);
            }
            // Narrow passes take no exit here: an odd count runs one empty
            // item, whose loads all read zero, so that both steps always run
            // and the waits for loads in flight stay exact across the loop.
            // Wide passes exit (one path less keeps their registers in bounds).
            if constexpr (KB >= kMatrixMinKB) {
                if (s + 1 >= n) break;
            }
            {
                stage_item<KB>(fb, FB, lane, wave_lds);
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

// The row sums of a tile: each pass's 16 item runs (each summed in item
// order from zero) go to four workgroups of four waves; workgroup q of pass
// P sums its waves as ((r0 + r1) + (r2 + r3)) into its own set of rows in the
// workspace (set 4(P-1) + q), and the combine launch forms
//   y = ((g0 + g1) + (g2 + g3)) + ((h0 + h1) + (h2 + h3))
// over the pass-1 sets g and the pass-2 sets h: one fixed tree over the 32
// runs, the same for every right-hand-side count and launch shape.
constexpr int kGroupWaves = kWaves / kPass1Groups;

template <int KB, bool FP32, int PASS>
__device__ __forceinline__ void group_rows(const MemberView& M, std::uint32_t group, int p0,
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
    if (PASS == 1)
        run_items<KB, FP32, 1>(M, tile->rsplit[r], tile->rsplit[r + 1], tile->split1[r],
                               tile->split1[r + 1], p0, kk, R, lds + wave * wave_lds_bytes<KB>(),
                               total);
    else
        run_items<KB, FP32, 2>(M, 0u, 0u, tile->split2[r], tile->split2[r + 1], p0, kk, R,
                               lds + wave * wave_lds_bytes<KB>(), total);
    double zero[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) zero[p] = 0.0;
    const std::uint32_t set = (PASS == 1 ? 0u : std::uint32_t(kPass1Groups)) + q;
    double* rows = M.zp + (M.z_rows + set * M.order + tile->row0 + lane) * std::uint64_t(M.k) + p0;
    combine_rows<KB, W>(total, zero, lds, wave, lane, [&](int pc, double a0, double a1) {
        if (lane < int(R)) {
            if (pc < kk) rows[pc] = a0;
            if (pc + 1 < kk) rows[pc + 1] = a1;
        }
    });
}

// y rows of tile t: the tree over the eight row sets, one thread per
// (row, right-hand side) so that reads and writes run along the columns.
__device__ __forceinline__ void combine_tile(const MemberView& M, std::uint32_t t) {
    const TileDev* __restrict__ tile = M.tiles + t;
    const std::uint32_t R = min(tile->rows, std::uint32_t(kTile));
    const std::uint32_t k = std::uint32_t(M.k);
    const std::uint64_t stride = std::uint64_t(M.order) * k;
    const double* rows = M.zp + (M.z_rows + tile->row0) * std::uint64_t(k);
    for (std::uint32_t e = threadIdx.x; e < R * k; e += blockDim.x) {
        const std::uint32_t i = e / k, p = e - i * k;
        double v[kRowSets];
#pragma unroll
        for (int s = 0; s < kRowSets; ++s) v[s] = rows[std::uint64_t(s) * stride + e];
        M.y[std::int64_t(M.perm[tile->row0 + i]) * M.ldy + p] =
            ((v[0] + v[1]) + (v[2] + v[3])) + ((v[4] + v[5]) + (v[6] + v[7]));
    }
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
    group_rows<KB, FP32, 1>(M, blockIdx.x, blockIdx.y * KB, lds);
}

template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kGroupWaves, pass1_occupancy(KB)) void apply_pass2(
    MemberView M) {
    __shared__ __align__(16) std::uint8_t lds[kGroupWaves * wave_lds_bytes<KB>()];
    group_rows<KB, FP32, 2>(M, blockIdx.x, blockIdx.y * KB, lds);
}

__global__ __launch_bounds__(256) void apply_combine(MemberView M) { combine_tile(M, blockIdx.x); }

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

template <int KB, bool FP32, int PASS>
__global__ __launch_bounds__(64 * kGroupWaves, batch1_occupancy(KB)) void batch_pass(
    const MemberView* __restrict__ members, const std::uint32_t* __restrict__ first,
    std::uint32_t count, const double* x, double* y, double* zp) {
    __shared__ __align__(16) std::uint8_t lds[kGroupWaves * wave_lds_bytes<KB>()];
    const std::uint32_t g = blockIdx.x / kPass1Groups;
    const std::uint32_t m = locate_member(first, count, g);
    const MemberView M = bind_member(members, m, x, y, zp);
    const int p0 = blockIdx.y * KB;
    if (p0 >= M.k) return;  // the whole workgroup leaves together
    group_rows<KB, FP32, PASS>(M, blockIdx.x - first[m] * kPass1Groups, p0, lds);
}

__global__ __launch_bounds__(256) void batch_permute(const MemberView* __restrict__ members,
                                                    const std::uint32_t* __restrict__ first,
                                                    std::uint32_t count, const double* x,
                                                    double* y, double* zp) {
    const std::uint32_t m = locate_member(first, count, blockIdx.x);
    const MemberView M = bind_member(members, m, x, y, zp);
    permute_tile(M, blockIdx.x - first[m]);
}

__global__ __launch_bounds__(256) void batch_combine(const MemberView* __restrict__ members,
                                                    const std::uint32_t* __restrict__ first,
                                                    std::uint32_t count, const double* x,
                                                    double* y, double* zp) {
    const std::uint32_t m = locate_member(first, count, blockIdx.x);
    const MemberView M = bind_member(members, m, x, y, zp);
    combine_tile(M, blockIdx.x - first[m]);
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
    if (!p.armed || 3 * (p.used + 1) > p.events.size()) return nullptr;
    return reinterpret_cast<hipEvent_t*>(p.events.data() + 3 * p.used++);
}

// Waves per workgroup: fewer waves with more registers each as the
// right-hand sides per pass grow (the wave lists are cut for 16 waves; a
// wave of a smaller workgroup takes 16/W consecutive lists).
// Right-hand sides per pass for an apply of k: the tuned choice when one
// was set for this k (qh_store_set_rhs_block), else k rounded up to a power
// of two, at most 16.
int default_rhs_block(std::int64_t k) { return k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : k <= 8 ? 8 : 16; }
int rhs_block(const DeviceStore& s, std::int64_t k) {
    const auto it = s.rhs_block.find(k);
    return it == s.rhs_block.end() ? default_rhs_block(k) : it->second;
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
    hipLaunchKernelGGL((apply_pass2<KB, FP32>), dim3(s.n_tiles * kPass1Groups, passes),
                       dim3(64 * kGroupWaves), 0, stream, M);
    hip_check(hipGetLastError(), "apply pass 2 launch");
    hipLaunchKernelGGL(apply_combine, dim3(s.n_tiles), dim3(256), 0, stream, M);
    hip_check(hipGetLastError(), "apply combine launch");
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
    hipLaunchKernelGGL((batch_pass<KB, FP32, 1>), dim3(c.total_tiles * kPass1Groups, passes),
                       dim3(64 * kGroupWaves), 0, stream, c.members, c.first_tile, count, x, y, zp);
    hip_check(hipGetLastError(), "batched pass 1 launch");
    hipLaunchKernelGGL((batch_pass<KB, FP32, 2>), dim3(c.total_tiles * kPass1Groups, passes),
                       dim3(64 * kGroupWaves), 0, stream, c.members, c.first_tile, count, x, y, zp);
    hip_check(hipGetLastError(), "batched pass 2 launch");
    hipLaunchKernelGGL(batch_combine, dim3(c.total_tiles), dim3(256), 0, stream, c.members,
                       c.first_tile, count, x, y, zp);
    hip_check(hipGetLastError(), "batched combine launch");
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
    switch (rhs_block(s, k)) {
        case 1: return launch_apply<1, FP32>(s, x, ldx, k, y, ldy, zp, st);
        case 2: return launch_apply<2, FP32>(s, x, ldx, k, y, ldy, zp, st);
        case 4: return launch_apply<4, FP32>(s, x, ldx, k, y, ldy, zp, st);
        case 8: return launch_apply<8, FP32>(s, x, ldx, k, y, ldy, zp, st);
        default: return launch_apply<16, FP32>(s, x, ldx, k, y, ldy, zp, st);
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

// Cuts `cost` into `parts` contiguous runs of similar total; out[i] is the
// first index of run i (out has parts + 1 entries, offset by `base`).
void split_runs(const std::vector<std::uint64_t>& cost, std::uint32_t parts, std::uint32_t base,
                std::uint32_t* out) {
    std::uint64_t total = 0;
    for (auto c : cost) total += c;
    std::uint32_t e = 0;
    std::uint64_t run = 0;
    out[0] = base;
    for (std::uint32_t w = 1; w < parts; ++w) {
        const std::uint64_t goal = total * w / parts;
        while (e < cost.size() && run + cost[e] / 2 < goal) run += cost[e++];
        out[w] = base + e;
    }
    out[parts] = base + std::uint32_t(cost.size());
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
            split_runs(cost, kWaves, std::uint32_t(items.size()), split);
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
    // The right-factor items fill the pass-1 runs up to an even total: run r
    // takes items (in order) while the cost of runs 0..r stays under r + 1
    // means, so that rounding does not pile up in the last run.
    std::uint64_t all = 0;
    for (auto c : run_cost) all += c;
    all += right_load;
    const std::uint64_t nruns = std::max<std::uint64_t>(1, run_cost.size());
    std::uint32_t next = 0;
    std::uint64_t done = 0;  // cost of the runs filled so far
    for (std::size_t r = 0; r < run_cost.size(); ++r) {
        TileDev& T = plan.tiles[r / kWaves];
        const int w = int(r % kWaves);
        T.rsplit[w] = next;
        done += run_cost[r];
        const std::uint64_t goal = all * (r + 1) / nruns;
        const bool last = r + 1 == run_cost.size();
        while (next < right_items.size() &&
               (last || done + cost_of(right_items[next]) / 2 <= goal)) {
            done += cost_of(right_items[next]);
            ++next;
        }
        T.rsplit[w + 1] = next;
    }
    plan.z_rows = z_rows;

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

void reconstruct(const DeviceStore& s, double* out, std::int64_t ld, void* stream) {
    require(out != nullptr, "null output pointer");
    require(ld >= std::int64_t(s.order), "leading dimension is smaller than the matrix order");
    DeviceScope scope(s.device);
    {
        std::lock_guard<std::mutex> guard(s.rec.lock);
        if (!s.rec.uploaded) {
            std::uint64_t bytes = 0;
            s.rec.tasks = upload(s.rec.host.tasks, bytes);
            s.rec.chunks = upload(s.rec.host.chunks, bytes);
            s.rec.lscale = upload(s.rec.host.lscale, bytes);
            s.rec.rscale = upload(s.rec.host.rscale, bytes);
            s.rec.uploaded = true;
        }
    }
    if (s.rec.host.tasks.empty()) return;
    hipLaunchKernelGGL(reconstruct_tiles, dim3(unsigned(s.rec.host.tasks.size())), dim3(256), 0,
                       static_cast<hipStream_t>(stream), s.rec.tasks, s.rec.chunks, s.words,
                       s.widths, s.rec.lscale, s.rec.rscale, s.perm, out, ld);
    hip_check(hipGetLastError(), "reconstruct launch");
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

void destroy_device_store(DeviceStore* s) {
    if (!s) return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(s->device);
    for (void* e : s->profile.events) (void)hipEventDestroy(static_cast<hipEvent_t>(e));
    for (void* p : {static_cast<void*>(s->perm), static_cast<void*>(s->words),
                    static_cast<void*>(s->widths), static_cast<void*>(s->cscale),
                    static_cast<void*>(s->items), static_cast<void*>(s->tiles),
                    static_cast<void*>(s->rec.tasks), static_cast<void*>(s->rec.chunks),
                    static_cast<void*>(s->rec.lscale), static_cast<void*>(s->rec.rscale)})
        if (p) (void)hipFree(p);
    if (prev >= 0) (void)hipSetDevice(prev);
    delete s;
}

std::uint64_t apply_workspace_bytes(const DeviceStore& s, std::int64_t k) {
    // partial sums, the row sets, then x in clustered order
    return (s.z_rows + (kRowSets + 1) * s.order) * std::uint64_t(k) * sizeof(double);
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
    // the kernels address x and the workspace with 32-bit byte offsets
    require(((s.order - 1) * std::uint64_t(ldx) + std::uint64_t(k)) * sizeof(double) < kOutOfRange &&
                need < kOutOfRange,
            "right-hand sides too large for one apply (split the columns)");
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
    require(kb == 0 || kb == 1 || kb == 2 || kb == 4 || kb == 8 || kb == 16,
            "right-hand sides per pass must be 1, 2, 4, 8 or 16 (0 = default)");
    if (kb == 0) s.rhs_block.erase(k);
    else s.rhs_block[k] = kb;
}

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
            zoff += apply_workspace_bytes(*stores[i], k[i]) / sizeof(double);
        }
        b->total = off;
        b->z_total = zoff;
        DeviceScope scope(b->device);
        for (int kb : {1, 2, 4, 8, 16}) {
            BatchClass c;
            c.kb = kb;
            std::vector<MemberView> views;
            std::vector<std::uint32_t> ft{0};
            for (std::uint64_t i = 0; i < count; ++i) {
                if (rhs_block(*b->stores[i], b->k[i]) != kb) continue;
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
        for (void* p : {static_cast<void*>(c.members), static_cast<void*>(c.first_tile)})
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
    require(b.total * sizeof(double) < kOutOfRange && need < kOutOfRange,
            "batch right-hand sides too large for one launch");
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
