// Device helpers shared by the apply kernels (device_store.hip: narrow
// passes, reconstruction, decode; wide_passes.hip: matrix-core passes).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "device_store.hpp"

namespace qhamd {
inline void hip_check(hipError_t e, const char* what) {
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

inline void require(bool ok, const char* message) {
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


// Workgroups of a 1-D launch are handed to the eight XCDs in turn (workgroup
// L runs on XCD L mod 8), and each XCD has its own L2.  A launch of `groups`
// work groups with `slices` slices each (right-hand-side slices, row groups)
// takes group g on XCD g mod 8 and runs its slices there one after another,
// so the slices, which read the same codes, share one L2, while the groups
// stay spread over the XCDs in turn (batches mix groups of very different
// cost, so a contiguous share per XCD would not balance).  The grid has
// xcd_grid(groups, slices) workgroups; workgroup L takes group *g and slice
// *s, with *g >= groups for the padding.
constexpr std::uint32_t kXcds = 8;
__host__ __device__ __forceinline__ void xcd_order(std::uint32_t L, std::uint32_t slices,
                                                   std::uint32_t& g, std::uint32_t& s) {
    const std::uint32_t j = L / kXcds;
    g = kXcds * (j / slices) + L % kXcds;
    s = j % slices;
}
inline std::uint32_t xcd_grid(std::uint64_t groups, std::uint64_t slices) {
    return std::uint32_t((groups + kXcds - 1) / kXcds * kXcds * slices);
}

// Issue priority by progress (four levels): a wave that has done less of
// its run than the waves it shares a SIMD with is issued first, so that the
// waves of a SIMD finish together (by default the oldest wave is issued
// first and the youngest finishes last, running alone).  done and total are
// wave-uniform.
__device__ __forceinline__ void progress_priority(std::uint32_t done, std::uint32_t total) {
    if (4 * done < total) __builtin_amdgcn_s_setprio(3);
    else if (4 * done < 2 * total) __builtin_amdgcn_s_setprio(2);
    else if (4 * done < 3 * total) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
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

// LDS of a staged item: a head row of 64 words that a decode may read below
// the first stream, the item (<= kItemQuads 16-byte groups), and slack past it.
constexpr int kStageBytes = kItemQuads * 16 + 272;
constexpr int kLdsHead = 256;  // readable bytes before the stage (decode_lm)

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


}  // namespace qhamd
