// Device-resident packed store and the apply plan (internal to the library).
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/qhmat_hip.h"

namespace qhamd {

// Row tiles are owned by one workgroup of kWaves waves; a tile has at most
// kTileRows rows (one per lane) and never crosses a row-slab boundary, so
// every block covering one row of a tile covers all of them.
constexpr int kTileRows = 64;
constexpr int kWaves = 4;
// Columns of a low-rank block reduced by one wave in the right-factor pass.
constexpr int kItemCols = 256;

// One off-diagonal block with rank > 0.  lbase/rbase: first word of its left
// and right factor in the word array; components follow the packed format
// (component j at the previous one's end, ceil(rows*b/32) words each).
struct LowRankDev {
    std::uint32_t row0, col0, m, n;
    std::uint32_t rank, comp_base;
    std::uint32_t item_begin, item_end;
    std::uint64_t lbase, rbase;
};

// A diagonal block: m x n FP64 stored column-major at values[off].
struct DenseDev {
    std::uint32_t row0, col0, m, n;
    std::uint64_t off;
};

// Right-factor work item: columns [c0, c1) of block `block`; its partial
// dot products go to zp[(zbase + j) * k + p].
struct ItemDev {
    std::uint32_t block, c0, c1, zbase;
};

// Row tile: rows [row0, row0+rows); entries [split[w], split[w+1]) are summed
// by wave w, waves are combined in order 0..kWaves-1.
struct TileDev {
    std::uint32_t row0, rows;
    std::uint32_t split[kWaves + 1];
    std::uint32_t pad;
};

// Entry of a tile: bit 31 set = dense block, low 31 bits = index into the
// dense or low-rank array; `row` = first tile row relative to the block.
struct EntryDev {
    std::uint32_t what, row;
};
constexpr std::uint32_t kDenseBit = 0x80000000u;

struct DeviceStore {
    int device = 0;
    std::uint64_t order = 0;
    // device arrays
    std::int32_t* perm = nullptr;
    std::uint32_t* words = nullptr;
    std::int32_t* widths = nullptr;
    double* lscale = nullptr;
    double* rscale = nullptr;
    double* dense = nullptr;
    LowRankDev* lowrank = nullptr;
    DenseDev* denses = nullptr;
    ItemDev* items = nullptr;
    TileDev* tiles = nullptr;
    EntryDev* entries = nullptr;
    // sizes
    std::uint64_t n_words = 0, n_comp = 0, n_dense_values = 0;
    std::uint32_t n_lowrank = 0, n_dense = 0, n_items = 0, n_tiles = 0, n_entries = 0;
    std::uint64_t z_rows = 0;  // sum of rank over items: workspace = z_rows * k doubles
    // host mirrors needed by entry points
    std::vector<std::int64_t> block_to_lowrank;  // insertion index -> lowrank index or -1
    std::vector<std::uint64_t> block_rows, block_cols;
    qh_store_stats stats{};
    std::uint64_t device_bytes = 0;
    // Optional per-pass event timing (qh_profile_begin / qh_profile_end).
    struct Profile {
        std::vector<void*> events;  // hipEvent_t, 3 per apply
        std::uint64_t used = 0;     // applies recorded
        bool armed = false;
    };
    mutable Profile profile;
};

void profile_begin(DeviceStore& s, std::uint64_t max_applies);
qh_profile profile_end(DeviceStore& s);

// Throw std::invalid_argument / std::runtime_error like the host store.
DeviceStore* create_device_store(const qh_store_desc& desc, int device);
void destroy_device_store(DeviceStore* s);
std::uint64_t apply_workspace_bytes(const DeviceStore& s, std::int64_t k);
void apply(const DeviceStore& s, const double* x, std::int64_t ldx, std::int64_t k, double* y,
           std::int64_t ldy, int precision, void* stream, void* workspace,
           std::uint64_t workspace_bytes);
void decode_codes(const DeviceStore& s, std::uint64_t block, int side, std::int32_t* out,
                  void* stream);
int device_count();

}  // namespace qhamd
