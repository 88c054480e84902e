// MI355X (gfx950) apply of a qhmat H-matrix, y = H x.
//
// The packed store is uploaded once.  An apply is two launches on the caller's
// stream, both deterministic (fixed reduction order, one writer per output):
//
//  1. right-factor pass (one wave per work item = <=256 columns of one
//     low-rank block): lanes own columns, decode their b-bit codes straight
//     from the resident words, multiply by x gathered through the permutation
//     and reduce across the wave per component:
//         zp[item][j][p] = sum_{c in item} codeR(c, j) * x[perm[col0 + c], p]
//  2. row-tile pass (one workgroup of four waves per tile of <=64 rows of one
//     row slab): lanes own rows; the waves split the tile's blocks and
//     accumulate
//         y[row, p] += sum_j codeL(row, j) * (sL_j sR_j sum_items zp[.][j][p])
//     for low-rank blocks and D[row, :] x for the diagonal block, then the
//     four partial rows are added in wave order and written once to
//     y[perm[row], p].
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

// ---------------------------------------------------------------------------
// Device helpers

// Sign-extended b-bit code number `idx` of the component stream at `cw`.
// The word pair is always read: a code that does not straddle a word boundary
// takes all its bits from the low word, and the array is padded so the
// second read stays in bounds.  v_alignbit_b32 + v_bfe_i32.
__device__ __forceinline__ int decode_code(const std::uint32_t* __restrict__ cw, std::uint32_t idx,
                                           std::uint32_t bits) {
    const std::uint32_t pos = idx * bits;
    const std::uint32_t w = pos >> 5;
    const std::uint32_t lo = cw[w];
    const std::uint32_t hi = cw[w + 1];
    const std::uint32_t raw = __builtin_amdgcn_alignbit(hi, lo, pos & 31u);
    return bits == 32 ? static_cast<int>(raw) : __builtin_amdgcn_sbfe(raw, 0u, bits);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ std::uint64_t component_words(std::uint32_t rows, std::uint32_t bits) {
    return (std::uint64_t(rows) * bits + 31u) >> 5;
}

// ---------------------------------------------------------------------------
// Pass 1: right factor.  KB right-hand sides per launch row (grid.y).
template <int KB, bool FP32>
__global__ __launch_bounds__(64) void right_factor_pass(
    const ItemDev* __restrict__ items, const LowRankDev* __restrict__ blocks,
    const std::int32_t* __restrict__ widths, const std::uint32_t* __restrict__ words,
    const std::int32_t* __restrict__ perm, const double* __restrict__ x, std::int64_t ldx, int k,
    double* __restrict__ zp) {
    const ItemDev it = items[blockIdx.x];
    const LowRankDev B = blocks[it.block];
    const int lane = threadIdx.x;
    const int p0 = blockIdx.y * KB;
    const int kk = min(KB, k - p0);
    constexpr int kSteps = kItemCols / 64;
    using Real = typename std::conditional<FP32, float, double>::type;

    // x rows of this lane's columns, gathered through the permutation once
    // (kept in registers for small KB, re-read through the cache otherwise).
    constexpr bool kKeep = KB <= 4;
    std::int64_t src[kSteps];
    Real xv[kKeep ? kSteps : 1][kKeep ? KB : 1];
#pragma unroll
    for (int t = 0; t < kSteps; ++t) {
        const std::uint32_t c = it.c0 + t * 64 + lane;
        src[t] = c < it.c1 ? std::int64_t(perm[B.col0 + c]) * ldx + p0 : 0;
        if constexpr (kKeep) {
#pragma unroll
            for (int p = 0; p < KB; ++p)
                xv[t][p] = (c < it.c1 && p < kk) ? Real(x[src[t] + p]) : Real(0);
        }
    }
    auto xat = [&](int t, int p) -> Real {
        if constexpr (kKeep) return xv[t][p];
        else return p < kk ? Real(x[src[t] + p]) : Real(0);
    };

    std::uint64_t woff = B.rbase;
    for (std::uint32_t j = 0; j < B.rank; ++j) {
        const std::uint32_t bits = widths[B.comp_base + j];
        const std::uint32_t* cw = words + woff;
        double total[KB];
        if (!FP32 || bits <= 24) {
            Real part[KB];
#pragma unroll
            for (int p = 0; p < KB; ++p) part[p] = Real(0);
#pragma unroll
            for (int t = 0; t < kSteps; ++t) {
                const std::uint32_t c = it.c0 + t * 64 + lane;
                if (c < it.c1) {
                    const Real code = Real(decode_code(cw, c, bits));
#pragma unroll
                    for (int p = 0; p < KB; ++p) part[p] = fma(code, xat(t, p), part[p]);
                }
            }
#pragma unroll
            for (int p = 0; p < KB; ++p) total[p] = double(wave_sum(part[p]));
        } else {
            // 25..32-bit codes: signed high half and unsigned low 16 bits,
            // both exact in FP32, summed apart and joined in FP64.
            float hi[KB], lo[KB];
#pragma unroll
            for (int p = 0; p < KB; ++p) hi[p] = lo[p] = 0.0f;
#pragma unroll
            for (int t = 0; t < kSteps; ++t) {
                const std::uint32_t c = it.c0 + t * 64 + lane;
                if (c < it.c1) {
                    const int code = decode_code(cw, c, bits);
                    const float h = float(code >> 16);
                    const float l = float(code & 0xFFFF);
#pragma unroll
                    for (int p = 0; p < KB; ++p) {
                        const float xf = float(xat(t, p));
                        hi[p] = fmaf(h, xf, hi[p]);
                        lo[p] = fmaf(l, xf, lo[p]);
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < KB; ++p)
                total[p] = double(wave_sum(hi[p])) * 65536.0 + double(wave_sum(lo[p]));
        }
        if (lane == 0) {
            double* out = zp + (std::uint64_t(it.zbase) + j) * std::uint64_t(k) + p0;
#pragma unroll
            for (int p = 0; p < KB; ++p)
                if (p < kk) out[p] = total[p];
        }
        woff += component_words(B.n, bits);
    }
}

// ---------------------------------------------------------------------------
// Pass 2: row tiles.
template <int KB, bool FP32>
__global__ __launch_bounds__(64 * kWaves) void row_tile_pass(
    const TileDev* __restrict__ tiles, const EntryDev* __restrict__ entries,
    const LowRankDev* __restrict__ blocks, const DenseDev* __restrict__ denses,
    const ItemDev* __restrict__ items, const std::int32_t* __restrict__ widths,
    const double* __restrict__ lscale, const double* __restrict__ rscale,
    const std::uint32_t* __restrict__ words, const double* __restrict__ dvals,
    const std::int32_t* __restrict__ perm, const double* __restrict__ zp,
    const double* __restrict__ x, std::int64_t ldx, int k, double* __restrict__ y,
    std::int64_t ldy) {
    __shared__ double partial[kWaves - 1][KB][64];
    const TileDev tile = tiles[blockIdx.x];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int p0 = blockIdx.y * KB;
    const int kk = min(KB, k - p0);
    const bool active = lane < int(tile.rows);

    double acc[KB];
#pragma unroll
    for (int p = 0; p < KB; ++p) acc[p] = 0.0;

    const std::uint32_t e_end = tile.split[wave + 1];
    for (std::uint32_t e = tile.split[wave]; e < e_end; ++e) {
        const EntryDev ent = entries[e];
        const std::uint32_t loc = ent.row + lane;
        if (ent.what & kDenseBit) {
            const DenseDev D = denses[ent.what & ~kDenseBit];
            const double* col = dvals + D.off + loc;
            for (std::uint32_t c = 0; c < D.n; ++c) {
                const double d = active ? col[std::uint64_t(c) * D.m] : 0.0;
                const double* xr = x + std::int64_t(perm[D.col0 + c]) * ldx + p0;
#pragma unroll
                for (int p = 0; p < KB; ++p)
                    if (p < kk) acc[p] = fma(d, xr[p], acc[p]);
            }
            continue;
        }
        const LowRankDev B = blocks[ent.what];
        std::uint64_t woff = B.lbase;
        float block_acc[KB];
#pragma unroll
        for (int p = 0; p < KB; ++p) block_acc[p] = 0.0f;
        for (std::uint32_t j = 0; j < B.rank; ++j) {
            const std::uint32_t cj = B.comp_base + j;
            const std::uint32_t bits = widths[cj];
            const double s = lscale[cj] * rscale[cj];
            double z[KB];
#pragma unroll
            for (int p = 0; p < KB; ++p) z[p] = 0.0;
            for (std::uint32_t i = B.item_begin; i < B.item_end; ++i) {
                const double* zr = zp + (std::uint64_t(items[i].zbase) + j) * std::uint64_t(k) + p0;
#pragma unroll
                for (int p = 0; p < KB; ++p)
                    if (p < kk) z[p] += zr[p];
            }
            const int code = active ? decode_code(words + woff, loc, bits) : 0;
            if (!FP32) {
                const double cd = double(code);
#pragma unroll
                for (int p = 0; p < KB; ++p) acc[p] = fma(cd, s * z[p], acc[p]);
            } else if (bits <= 24) {
                const float cf = float(code);
#pragma unroll
                for (int p = 0; p < KB; ++p) block_acc[p] = fmaf(cf, float(s * z[p]), block_acc[p]);
            } else {
                const float h = float(code >> 16), l = float(code & 0xFFFF);
#pragma unroll
                for (int p = 0; p < KB; ++p) {
                    const float zf = float(s * z[p]);
                    block_acc[p] = fmaf(h, zf * 65536.0f, block_acc[p]);
                    block_acc[p] = fmaf(l, zf, block_acc[p]);
                }
            }
            woff += component_words(B.m, bits);
        }
        if (FP32) {
#pragma unroll
            for (int p = 0; p < KB; ++p) acc[p] += double(block_acc[p]);
        }
    }

    if (wave > 0) {
#pragma unroll
        for (int p = 0; p < KB; ++p) partial[wave - 1][p][lane] = acc[p];
    }
    __syncthreads();
    if (wave == 0 && active) {
#pragma unroll
        for (int w = 0; w < kWaves - 1; ++w)
#pragma unroll
            for (int p = 0; p < KB; ++p) acc[p] += partial[w][p][lane];
        double* yr = y + std::int64_t(perm[tile.row0 + lane]) * ldy + p0;
#pragma unroll
        for (int p = 0; p < KB; ++p)
            if (p < kk) yr[p] = acc[p];
    }
}

// Test hook: codes of one factor side, component-major.
__global__ void decode_side(const std::uint32_t* __restrict__ words, std::uint64_t base,
                            const std::int32_t* __restrict__ widths, std::uint32_t rows,
                            std::int32_t* __restrict__ out) {
    const std::uint32_t j = blockIdx.x;
    std::uint64_t off = base;
    for (std::uint32_t t = 0; t < j; ++t) off += component_words(rows, widths[t]);
    const std::uint32_t bits = widths[j];
    for (std::uint32_t i = threadIdx.x; i < rows; i += blockDim.x)
        out[std::uint64_t(j) * rows + i] = decode_code(words + off, i, bits);
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
    hipEvent_t* ev = profile_slot(s);
    if (ev) hip_check(hipEventRecord(ev[0], stream), "hipEventRecord");
    if (s.n_items) {
        hipLaunchKernelGGL((right_factor_pass<KB, FP32>), dim3(s.n_items, passes), dim3(64), 0,
                           stream, s.items, s.lowrank, s.widths, s.words, s.perm, x, ldx, k, zp);
        hip_check(hipGetLastError(), "right-factor pass launch");
    }
    if (ev) hip_check(hipEventRecord(ev[1], stream), "hipEventRecord");
    hipLaunchKernelGGL((row_tile_pass<KB, FP32>), dim3(s.n_tiles, passes), dim3(64 * kWaves), 0,
                       stream, s.tiles, s.entries, s.lowrank, s.denses, s.items, s.widths,
                       s.lscale, s.rscale, s.words, s.dense, s.perm, zp, x, ldx, k, y, ldy);
    hip_check(hipGetLastError(), "row-tile pass launch");
    if (ev) hip_check(hipEventRecord(ev[2], stream), "hipEventRecord");
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

void require(bool ok, const char* message) {
    if (!ok) throw std::invalid_argument(message);
}

}  // namespace

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

DeviceStore* create_device_store(const qh_store_desc& d, int device) {
    require(d.order > 0 && d.perm && d.nblocks > 0 && d.rect && d.is_dense && d.rank,
            "incomplete store description");
    require(d.order < (std::uint64_t{1} << 26),
            "matrix order exceeds the device store limit (2^26)");
    const int count = device_count();
    if (count == 0) throw std::runtime_error("no CUDA device is available (no HIP device found)");
    require(device >= 0 && device < count, "invalid device ordinal");
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

    auto* s = new DeviceStore;
    s->device = device;
    s->order = n;
    std::vector<std::uint32_t> words;
    std::vector<std::int32_t> widths;
    std::vector<double> lsc, rsc, dvals;
    std::vector<LowRankDev> lowrank;
    std::vector<DenseDev> denses;
    std::vector<ItemDev> items;
    s->block_to_lowrank.assign(d.nblocks, -1);
    s->block_rows.resize(d.nblocks);
    s->block_cols.resize(d.nblocks);
    std::uint64_t codes = 0, lr_flops = 0, dn_flops = 0, z_rows = 0;
    std::uint64_t n_lowrank_all = 0, left_words = 0, right_words = 0;
    // Entry candidates in insertion order: (row0, row1, what).
    struct Cover {
        std::uint64_t row0, row1;
        std::uint32_t what;
        std::uint64_t cost;
    };
    std::vector<Cover> covers;
    covers.reserve(d.nblocks);
    for (std::uint64_t b = 0; b < d.nblocks; ++b) {
        const std::uint64_t r0 = d.rect[4 * b], r1 = d.rect[4 * b + 1];
        const std::uint64_t c0 = d.rect[4 * b + 2], c1 = d.rect[4 * b + 3];
        require(r0 < r1 && r1 <= n && c0 < c1 && c1 <= n, "invalid block rectangle");
        const std::uint64_t m = r1 - r0, nc = c1 - c0;
        s->block_rows[b] = m;
        s->block_cols[b] = nc;
        if (d.is_dense[b]) {
            require(d.dense && d.dense[b], "dense block without values");
            DenseDev D{std::uint32_t(r0), std::uint32_t(c0), std::uint32_t(m), std::uint32_t(nc),
                       dvals.size()};
            const double* v = d.dense[b];
            const std::size_t at = dvals.size();
            dvals.resize(at + m * nc);
            for (std::uint64_t i = 0; i < m; ++i)  // row-major in, column-major out
                for (std::uint64_t c = 0; c < nc; ++c) dvals[at + c * m + i] = v[i * nc + c];
            covers.push_back({r0, r1, kDenseBit | std::uint32_t(denses.size()), nc});
            denses.push_back(D);
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
        LowRankDev L{};
        L.row0 = std::uint32_t(r0);
        L.col0 = std::uint32_t(c0);
        L.m = std::uint32_t(m);
        L.n = std::uint32_t(nc);
        L.rank = std::uint32_t(r);
        L.comp_base = std::uint32_t(widths.size());
        std::uint64_t lw = 0, rw = 0;
        for (std::uint64_t j = 0; j < r; ++j) {
            const int bits = d.widths[b][j];
            require(bits >= 2 && bits <= 32, "bit width must be in [2, 32]");
            widths.push_back(bits);
            lsc.push_back(d.lscale[b][j]);
            rsc.push_back(d.rscale[b][j]);
            lw += (m * bits + 31) / 32;
            rw += (nc * bits + 31) / 32;
        }
        left_words += lw;
        right_words += rw;
        L.lbase = words.size();
        words.insert(words.end(), d.lwords[b], d.lwords[b] + lw);
        L.rbase = words.size();
        words.insert(words.end(), d.rwords[b], d.rwords[b] + rw);
        L.item_begin = std::uint32_t(items.size());
        for (std::uint64_t c = 0; c < nc; c += kItemCols) {
            items.push_back({std::uint32_t(lowrank.size()), std::uint32_t(c),
                             std::uint32_t(std::min<std::uint64_t>(nc, c + kItemCols)),
                             std::uint32_t(z_rows)});
            z_rows += r;
        }
        L.item_end = std::uint32_t(items.size());
        s->block_to_lowrank[b] = std::int64_t(lowrank.size());
        covers.push_back({r0, r1, std::uint32_t(lowrank.size()), 4 + 3 * r});
        lowrank.push_back(L);
        codes += (m + nc) * r;
        lr_flops += 2 * r * (m + nc);
    }
    require(z_rows < (std::uint64_t{1} << 31), "store too large for one device plan");
    words.resize(words.size() + 4, 0u);  // pair reads past the last code

    // Row tiles: slabs between distinct row boundaries, cut into <= 64 rows.
    std::vector<std::uint64_t> cuts{0, n};
    for (std::uint64_t b = 0; b < d.nblocks; ++b) {
        cuts.push_back(d.rect[4 * b]);
        cuts.push_back(d.rect[4 * b + 1]);
    }
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    std::vector<TileDev> tiles;
    for (std::size_t sl = 0; sl + 1 < cuts.size(); ++sl)
        for (std::uint64_t r = cuts[sl]; r < cuts[sl + 1]; r += kTileRows) {
            TileDev t{};
            t.row0 = std::uint32_t(r);
            t.rows = std::uint32_t(std::min<std::uint64_t>(kTileRows, cuts[sl + 1] - r));
            tiles.push_back(t);
        }
    std::vector<std::vector<EntryDev>> per_tile(tiles.size());
    std::vector<std::vector<std::uint64_t>> per_cost(tiles.size());
    for (const Cover& c : covers) {
        auto first = std::lower_bound(tiles.begin(), tiles.end(), c.row0,
                                      [](const TileDev& t, std::uint64_t r) { return t.row0 < r; });
        for (auto t = first; t != tiles.end() && t->row0 < c.row1; ++t) {
            const std::size_t ti = std::size_t(t - tiles.begin());
            per_tile[ti].push_back({c.what, std::uint32_t(t->row0 - c.row0)});
            per_cost[ti].push_back(c.cost);
        }
    }
    std::vector<EntryDev> entries;
    for (std::size_t ti = 0; ti < tiles.size(); ++ti) {
        const auto& list = per_tile[ti];
        const auto& cost = per_cost[ti];
        const std::uint32_t base = std::uint32_t(entries.size());
        entries.insert(entries.end(), list.begin(), list.end());
        // Contiguous split of the entries into kWaves runs of similar cost.
        std::uint64_t total = 0;
        for (auto c : cost) total += c;
        std::uint32_t e = 0;
        std::uint64_t run = 0;
        tiles[ti].split[0] = base;
        for (int w = 1; w < kWaves; ++w) {
            const std::uint64_t goal = total * std::uint64_t(w) / kWaves;
            while (e < list.size() && run + cost[e] / 2 < goal) run += cost[e++];
            tiles[ti].split[w] = base + e;
        }
        tiles[ti].split[kWaves] = base + std::uint32_t(list.size());
    }

    try {
        DeviceScope scope(device);
        std::uint64_t bytes = 0;
        s->perm = upload(perm, bytes);
        s->words = upload(words, bytes);
        s->widths = upload(widths, bytes);
        s->lscale = upload(lsc, bytes);
        s->rscale = upload(rsc, bytes);
        s->dense = upload(dvals, bytes);
        s->lowrank = upload(lowrank, bytes);
        s->denses = upload(denses, bytes);
        s->items = upload(items, bytes);
        s->tiles = upload(tiles, bytes);
        s->entries = upload(entries, bytes);
        s->device_bytes = bytes;
    } catch (...) {
        destroy_device_store(s);
        throw;
    }
    s->n_words = words.size() - 4;
    s->n_comp = widths.size();
    s->n_dense_values = dvals.size();
    s->n_lowrank = std::uint32_t(lowrank.size());
    s->n_dense = std::uint32_t(denses.size());
    s->n_items = std::uint32_t(items.size());
    s->n_tiles = std::uint32_t(tiles.size());
    s->n_entries = std::uint32_t(entries.size());
    s->z_rows = z_rows;

    qh_store_stats& st = s->stats;
    st.order = n;
    st.nblocks = d.nblocks;
    st.ndense = denses.size();
    st.nlowrank = n_lowrank_all;
    st.ncomponents = widths.size();
    st.code_words = s->n_words;
    st.left_code_words = left_words;
    st.right_code_words = right_words;
    st.codes = codes;
    st.dense_values = dvals.size();
    st.device_bytes = s->device_bytes;
    st.lowrank_flops_per_rhs = lr_flops;
    st.dense_flops_per_rhs = dn_flops;
    st.tiles = tiles.size();
    st.items = items.size();
    st.entries = entries.size();
    st.device = device;
    return s;
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
                    static_cast<void*>(s->lowrank), static_cast<void*>(s->denses),
                    static_cast<void*>(s->items), static_cast<void*>(s->tiles),
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
    LowRankDev L;
    DeviceScope scope(s.device);
    hip_check(hipMemcpy(&L, s.lowrank + li, sizeof(L), hipMemcpyDeviceToHost), "hipMemcpy");
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(decode_side, dim3(L.rank), dim3(256), 0, st, s.words,
                       side == 0 ? L.lbase : L.rbase, s.widths + L.comp_base,
                       side == 0 ? L.m : L.n, out);
    hip_check(hipGetLastError(), "decode launch");
}

}  // namespace qhamd
