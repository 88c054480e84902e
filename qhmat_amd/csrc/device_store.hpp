// Device-resident packed store and the apply plan (internal to the library).
//
// Layout in HBM ("tile-major"): the matrix rows are cut into row tiles of at
// most kTile rows that never cross a block boundary, and the columns into
// column tiles the same way.  Each low-rank block's left factor is stored as
// one chunk per row tile it covers and its right factor as one chunk per
// column tile: a chunk holds, for every component j, the codes of that
// tile's rows (or columns) as a word-aligned LSB-first stream of
// ceil(R*b_j/32) words -- the reference's packed format (core.cpp:235-267)
// cut at tile boundaries -- and is padded to a multiple of four words so
// that a wave can stream it with 16-byte loads.  When a tile starts on a
// multiple of 32 codes the chunk's words are the reference's words verbatim.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/qhmat_hip.h"

namespace qhamd {

constexpr int kTile = 64;          // rows per row tile, columns per column tile
constexpr int kRowWaves = 16;      // waves per row-tile workgroup
constexpr int kStageWords = 1024;  // LDS words a wave stages at a time
constexpr int kPassComponents = 16;  // components per right-factor pass (4 lanes each)
constexpr int kUnitWaves = 4;        // right-factor units (one per wave) per workgroup

// Low-rank block with rank > 0.  Right-factor partial sums of column tile s,
// component j live in workspace row zbase + s * rank + j.
struct BlockDev {
    std::uint32_t rank, comp_base;
    std::uint32_t zbase, nseg;
};

// Right-factor unit: block `block` restricted to columns [col0, col0+cols).
struct UnitDev {
    std::uint32_t block, col0, cols, zrow;
    std::uint32_t words, pad;  // words of the chunk (all components, unpadded)
    std::uint64_t chunk;       // first word, multiple of 4
};

// Row-tile entry: kDenseBit | dense index, or low-rank block index.
struct EntryDev {
    std::uint32_t what;
    std::uint32_t row;    // dense: first tile row relative to the block
    std::uint32_t words;  // low-rank: words of the chunk (all components); dense: first column
    std::uint32_t cols;   // dense: columns of this entry (<= kDenseCols)
    std::uint64_t chunk;  // low-rank: first word of the block's chunk for this tile
};
constexpr int kDenseCols = 16;  // columns of a diagonal block per row-tile entry
constexpr std::uint32_t kDenseBit = 0x80000000u;

// Row tile: rows [row0, row0+rows); entries [split[w], split[w+1]) are summed
// by wave w and the waves are combined in order 0..kRowWaves-1.
struct TileDev {
    std::uint32_t row0, rows;
    std::uint32_t split[kRowWaves + 1];
    std::uint32_t pad;
};

// Diagonal block: m x n FP64 stored column-major at values[off].
struct DenseDev {
    std::uint32_t row0, col0, m, n;
    std::uint64_t off;
};

// One store's device arrays plus one call's operands: everything a work item
// of either pass reads.  Passed by value for a single apply; a batch keeps a
// table of them in device memory with x, y and workspace given as offsets
// from the call's base pointers.
struct MemberView {
    const std::int32_t* perm;
    const std::uint32_t* words;
    const std::int32_t* widths;
    const double* lscale;
    const double* rscale;
    const double* dense;
    const BlockDev* blocks;
    const DenseDev* denses;
    const UnitDev* units;
    const TileDev* tiles;
    const EntryDev* entries;
    const double* x;
    double* y;
    double* zp;
    std::int64_t ldx, ldy;
    std::int32_t k;
    std::uint32_t units_count, tiles_count;
    std::uint32_t pad;
};

// Host record of one chunk (test hook and statistics).
struct ChunkRef {
    std::uint32_t first, count;  // rows/columns of the block it holds
    std::uint64_t chunk;
};

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
    BlockDev* blocks = nullptr;
    DenseDev* denses = nullptr;
    UnitDev* units = nullptr;
    TileDev* tiles = nullptr;
    EntryDev* entries = nullptr;
    // sizes
    std::uint64_t n_words = 0, n_comp = 0, n_dense_values = 0;
    std::uint32_t n_lowrank = 0, n_dense = 0, n_units = 0, n_tiles = 0, n_entries = 0;
    std::uint64_t z_rows = 0;  // workspace = z_rows * k doubles
    // host mirrors needed by entry points
    std::vector<std::int64_t> block_to_lowrank;  // insertion index -> low-rank index or -1
    std::vector<std::vector<ChunkRef>> left_chunks, right_chunks;  // per low-rank block
    std::vector<std::uint32_t> comp_base;                          // per low-rank block
    std::vector<std::uint32_t> ranks;                              // per low-rank block
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

// Everything the device store holds, built on the host before the upload.
struct Plan {
    std::uint64_t order = 0;
    std::vector<std::int32_t> perm;
    std::vector<std::uint32_t> words;  // chunks + 8 words of padding
    std::vector<std::int32_t> widths;
    std::vector<double> lscale, rscale, dense;
    std::vector<BlockDev> blocks;
    std::vector<DenseDev> denses;
    std::vector<UnitDev> units;
    std::vector<TileDev> tiles;
    std::vector<EntryDev> entries;
    std::vector<std::int64_t> block_to_lowrank;
    std::vector<std::vector<ChunkRef>> left_chunks, right_chunks;
    std::uint64_t z_rows = 0, stored_words = 0;
    qh_store_stats stats{};
};

// Throw std::invalid_argument / std::runtime_error like the host store.
Plan build_plan(const qh_store_desc& desc);
// Runs the two passes' index arithmetic on the host against the plan's
// arrays, with every array access bounds-checked (std::out_of_range), and
// returns y = H x in FP64.  A check of the layout without a device.
void emulate_plan(const Plan& plan, const double* x, std::int64_t k, double* y);
DeviceStore* create_device_store(const qh_store_desc& desc, int device);

// One launch pair of a batch: the members whose right-hand-side count rounds
// up to the same power of two kb (<= 16; larger counts take several passes).
struct BatchClass {
    int kb = 1;
    std::uint64_t member_count = 0;
    std::int64_t max_k = 1;
    std::uint32_t total_units = 0, total_tiles = 0;
    MemberView* members = nullptr;        // device table; x, y, zp hold offsets
    std::uint32_t* first_unit = nullptr;  // device, member_count + 1 prefix counts
    std::uint32_t* first_tile = nullptr;
};

// A fixed set of resident stores applied together.  Member i reads its
// order_i x k_i right-hand sides at x + offset[i] (row-major, leading
// dimension k_i), writes y at the same offset, and uses workspace doubles
// from zoffset[i].
struct Batch {
    int device = 0;
    std::vector<const DeviceStore*> stores;
    std::vector<std::int64_t> k;
    std::vector<std::uint64_t> offset, zoffset;
    std::uint64_t total = 0, z_total = 0;
    std::vector<BatchClass> classes;
};

Batch* create_batch(const DeviceStore* const* stores, const std::int64_t* k, std::uint64_t count);
void destroy_batch(Batch* b);
void batch_apply(const Batch& b, const double* x, double* y, int precision, void* stream,
                 void* workspace, std::uint64_t workspace_bytes);
// Captures one batch apply on a private stream into an instantiated graph
// (hipGraphExec_t returned as void*); the workspace must be given.
void* batch_graph(const Batch& b, const double* x, double* y, int precision, void* workspace,
                  std::uint64_t workspace_bytes);
void graph_launch(void* exec, void* stream);
void graph_destroy(void* exec);
void destroy_device_store(DeviceStore* s);
std::uint64_t apply_workspace_bytes(const DeviceStore& s, std::int64_t k);
void apply(const DeviceStore& s, const double* x, std::int64_t ldx, std::int64_t k, double* y,
           std::int64_t ldy, int precision, void* stream, void* workspace,
           std::uint64_t workspace_bytes);
void decode_codes(const DeviceStore& s, std::uint64_t block, int side, std::int32_t* out,
                  void* stream);
int device_count();
void profile_begin(DeviceStore& s, std::uint64_t max_applies);
qh_profile profile_end(DeviceStore& s);

// Appends codes [first, first+count) of one component stream of width `bits`
// to `out` as a fresh word-aligned stream (host side of the upload).
void append_codes(const std::uint32_t* stream, std::uint64_t stream_words, std::uint64_t first,
                  std::uint64_t count, int bits, std::vector<std::uint32_t>& out);

}  // namespace qhamd
