// Device-resident packed store and the apply plan (internal to the library).
//
// Layout in HBM.  The matrix rows are cut into row tiles of at most kTile
// rows that never cross a block boundary, and the columns into column tiles
// the same way.  All work of an apply is a list of *items*; each item's data
// (codes or diagonal values) is one contiguous, 16-byte aligned run of at
// most kItemQuads 16-byte groups in the `words` array, so that a wave can
// request an item with a fixed number of 16-byte loads per lane before it
// needs any of it.  Item kinds:
//
//  kDense  columns [col, col+count) of a diagonal block restricted to one
//          row tile: R x count FP64 values, column-major (R = tile rows).
//  kFused  a low-rank block that lies in one row tile and one column tile,
//          components [comp0, comp0+count): first the right-factor codes of
//          those components (one word-aligned stream of ceil(cols*b/32)
//          words each), then their left-factor codes (ceil(R*b/32) words
//          each).  The apply forms z = V^T x and y += U z inside the item.
//  kRight  one column tile of a larger low-rank block, components
//          [comp0, comp0+count): right-factor streams only.  Writes the
//          partial sums of its column tile to the workspace.
//  kLeft   one row tile of a larger low-rank block, components
//          [comp0, comp0+count): left-factor streams only.  Reads the
//          partial sums of the block's column tiles from the workspace.
//
// Codes form side blocks of 64 lanes (lane L = the tile's row or column L):
// lane L's codes of the item's components follow one another in its own
// LSB-first bit stream (the reference's two's-complement b-bit codes,
// core.cpp:235-267), word w of lane L at word 64 w + slot(w, L) of the block
// (right-factor blocks swizzled by row, see lane_slot).  A kFused item holds
// its right block, then its left block.
//
// An apply is four launches:
//  permute  x in clustered order into the workspace;
//  pass 1   four workgroups per row tile, each running four of the tile's 16
//           item runs: the tile's kDense and kFused items, whose row sums go
//           to the workspace, and a share of the kRight items;
//  pass 2   the same shape over the tile's kLeft items;
//  combine  one workgroup per row tile sums the eight row sets into y
//           through the permutation.
#pragma once

#include <atomic>
#include <cstdint>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <vector>

#include "../../include/qhmat_hip.h"

namespace qhamd {

constexpr int kTile = 64;          // rows per row tile, columns per column tile
constexpr int kWaves = 16;         // item runs per tile and pass
constexpr int kPass1Groups = 4;    // pass-1 workgroups per tile (4 runs each)
constexpr int kItemQuads = 256;    // 16-byte groups of data per item (4 KiB)
constexpr int kItemComponents = 64;  // components per item
constexpr int kDenseColumns = 8;     // columns per kDense item (64 x 8 FP64 = 4 KiB)

enum ItemKind : std::uint8_t { kDense = 0, kFused = 1, kRight = 2, kLeft = 3 };

// One item (32 bytes).  Fields not listed for a kind are zero.
struct ItemDev {
    std::uint32_t quad;     // first 16-byte group of the data in `words`
    std::uint16_t quads;    // 16-byte groups of data (<= kItemQuads)
    std::uint8_t kind;      // ItemKind
    std::uint8_t count;     // kDense: columns; otherwise components (<= 64)
    std::uint32_t comp0;    // first component (index into widths / cscale)
    std::uint32_t col;      // kDense / kFused / kRight: first column (clustered index)
    std::uint32_t cols;     // kFused / kRight: columns of the block in this tile
    std::uint32_t zrow;     // kRight: workspace row of comp0; kLeft: same, column tile 0
    std::uint32_t zstride;  // kLeft: workspace rows between column tiles (the block rank)
    std::uint32_t nseg;     // kLeft: column tiles of the block
};
static_assert(sizeof(ItemDev) == 32, "ItemDev is 32 bytes");

// Where a narrow-pass wave starts: its run [b1, e1) followed by [b2, e2)
// (pass 1: a share of the kRight items, then the tile's own items; pass 2:
// an empty first range, then the tile's kLeft items) and the records of the
// run's first two items (zero records past its end), so that a wave's first
// fetches need one memory round trip instead of three (bounds, records,
// data).  Wave r of tile t in pass P (1 or 2) reads heads[((P - 1) tiles +
// t) kWaves + r].
struct WaveHead {
    std::uint32_t b1, e1, b2, e2;
    ItemDev first[2];
};
static_assert(sizeof(WaveHead) == 80, "WaveHead is 80 bytes");

// Row tile: rows [row0, row0+rows).  Run w (of kWaves) of the tile's pass-1
// workgroups streams the kRight items [rsplit[w], rsplit[w+1]) (a share of
// all right factors, independent of the tile) and the tile's items
// [split1[w], split1[w+1]); run w of pass 2 the items [split2[w],
// split2[w+1]).  Each pass runs a tile's 16 runs as four workgroups of four
// waves.
struct TileDev {
    std::uint32_t row0, rows;
    std::uint32_t rsplit[kWaves + 1];
    std::uint32_t split1[kWaves + 1];
    std::uint32_t split2[kWaves + 1];
};

// ---- wide passes: 16, 32 or 64 right-hand sides per pass on the matrix cores
//
// The same side blocks as above, in a second item table (`witems`) ordered
// for two passes and a combine:
//  pass R  workgroups of kWideWaves waves own one column tile: x of its
//          columns is loaded into LDS once, and each wave streams a run of
//          the tile's right-factor side blocks (every kRight item, and the
//          right half of every kFused item), decoding each code once for all
//          the pass's right-hand sides and writing the scaled partial sums
//          z = cscale * V^T x of the tile to the workspace: per component for
//          a kRight item, in virtual-row order (the order pass L reads them
//          in, see wide_passes.hip) for a kFused item's right half;
//  zsum    one workgroup per (block, component range) of the kLeft items
//          sums the partial sums of the block's column tiles into rows in
//          virtual-row order;
//  pass L  workgroups of kWideWaves waves own a run of one row tile's items
//          (its diagonal kDense items, kLeft items and the left halves of
//          its kFused items); wave w forms rows 16 w .. 16 w + 15, summing
//          each block's product in FP32 (FP32 mode) and the blocks in FP64;
//          a tile with one run writes y, others write row sets;
//  combine sums a tile's row sets in a fixed order into y.
constexpr int kWideWaves = 4;
constexpr std::uint32_t kNoSet = 0xFFFFFFFFu;

// A run of wide items [begin, end): pass R, one per wave (all waves of a
// workgroup share `tile`, a column tile); pass L, one per workgroup (`tile`
// a row tile, `set` its row set or kNoSet when the run writes y).
struct WideRun {
    std::uint32_t tile, begin, end, set;
};
// Row tile `tile` whose pass-L runs wrote row sets [set0, set0 + count).
struct WideSum {
    std::uint32_t tile, set0, count, pad;
};
// Column tile: clustered columns [first, first + count).
struct Span {
    std::uint32_t first, count;
};
// The partial sums of a block's components [comp0, comp0 + count) over its
// nseg column tiles (rows zrow + s zstride + j), summed in order between the
// two passes into rows vzrow + v in the components' virtual-row order (the
// order pass L reads them in).
struct ZSum {
    std::uint32_t zrow, zstride, nseg, count, vzrow, comp0, pad0, pad1;
};
static_assert(sizeof(ZSum) == 32, "ZSum is 32 bytes");

// Everything the wide passes read, plus one call's operands (x, y and the
// workspace as offsets from the call's bases in a batch table).
struct WideView {
    const std::int32_t* perm;
    const std::uint32_t* words;
    const std::int32_t* widths;
    const double* cscale;
    const ItemDev* items;
    const WideRun* rruns;
    const WideRun* lruns;
    const WideSum* sums;
    const ZSum* zsums;
    const Span* ctiles;
    const TileDev* tiles;
    const double* x;
    double* y;
    void* zw;  // workspace: z_rows x k partial sums (FP32 in FP32 mode, FP64
               // otherwise, row stride k), then the row sets: sets x 64 x k FP64
    std::int64_t ldx, ldy;
    std::int32_t k;
    std::uint32_t pad;
    std::uint64_t z_rows;
};

// One store's device arrays plus one call's operands: everything a work item
// of either pass reads.  Passed by value for a single apply; a batch keeps a
// table of them in device memory with x, y and workspace given as offsets
// from the call's base pointers.
struct MemberView {
    const std::int32_t* perm;
    const std::uint32_t* words;
    const std::int32_t* widths;
    const double* cscale;        // lscale * rscale per component
    const ItemDev* items;
    const TileDev* tiles;
    const WaveHead* heads;
    const double* x;
    double* y;
    double* zp;     // workspace: z_rows x k partial sums, then the row sets (kPass1Groups
                    // sets of order x k row sums, one per pass-1 workgroup of a tile),
                    // then x in clustered order
    std::int64_t ldx, ldy;
    std::int32_t k;
    std::uint32_t tiles_count;  // row tiles
    std::uint32_t order;
    std::uint32_t pad;
    std::uint64_t z_rows;       // first workspace row of the row sums
};

// Host record of where one factor side of a block's codes lives: components
// [j0, j0+jn) of rows/columns [first, first+count) start at word `word`.
struct ChunkRef {
    std::uint32_t first, count, j0, jn;
    std::uint64_t word;
};

// One (block, row tile, column tile) of a reconstruction (qh_reconstruct).
struct RecTask {
    std::uint32_t row0, rows, col0, cols;  // clustered rows and columns of the tile
    std::uint32_t kind;                    // 0: zero (rank 0), 1: dense, 2: low rank
    std::uint32_t rank, comp_base;         // low rank: components
    std::uint32_t lchunk, nl, rchunk, nr;  // low rank: runs of the two sides' chunks
    std::uint32_t quad;                    // dense: first item of the block's row tile
    std::uint32_t bcol0;                   // dense: first column of the block
    std::uint32_t pad[3];
};
static_assert(sizeof(RecTask) == 64, "RecTask is 64 bytes");

// Tables of a reconstruction, built with the plan and uploaded on first use.
struct RecTables {
    std::vector<RecTask> tasks;
    std::vector<ChunkRef> chunks;
    std::vector<double> lscale, rscale;  // per component, as the factors were packed
};

struct DeviceStore {
    int device = 0;
    // batches whose member tables hold this store's device pointers; the
    // store refuses destruction while any exist
    mutable std::atomic<int> pins{0};
    std::uint64_t order = 0;
    // device arrays
    std::int32_t* perm = nullptr;
    std::uint32_t* words = nullptr;
    std::int32_t* widths = nullptr;
    double* cscale = nullptr;
    ItemDev* items = nullptr;
    TileDev* tiles = nullptr;
    WaveHead* heads = nullptr;
    // wide passes
    ItemDev* witems = nullptr;
    WideRun* wruns_r = nullptr;
    WideRun* wruns_l = nullptr;
    WideSum* wsums = nullptr;
    ZSum* wzsums = nullptr;
    Span* ctiles = nullptr;
    std::uint32_t n_wr_groups = 0, n_wl_groups = 0, n_wsums = 0, n_wzsums = 0, wsets = 0;
    std::uint64_t wz_rows = 0;
    // sizes
    std::uint64_t n_words = 0, n_comp = 0;
    std::uint32_t n_tiles = 0, n_items = 0;
    std::uint64_t z_rows = 0;  // workspace = (z_rows + (2 kPass1Groups + 1) order) k doubles
    // host mirrors needed by entry points
    std::vector<std::int64_t> block_to_lowrank;  // insertion index -> low-rank index or -1
    std::vector<std::vector<ChunkRef>> left_chunks, right_chunks;  // per low-rank block
    std::vector<std::uint32_t> comp_base;                          // per low-rank block
    std::vector<std::uint32_t> ranks;                              // per low-rank block
    qh_store_stats stats{};
    std::uint64_t device_bytes = 0;
    // Right-hand sides per pass chosen for a right-hand-side count
    // (qh_store_set_rhs_block, e.g. by the autotuner); others use the default.
    std::map<std::int64_t, int> rhs_block;
    mutable std::shared_mutex rhs_lock;  // applies read rhs_block, set_rhs_block writes it
    // Reconstruction tables: host copies, device copies made by the first
    // qh_reconstruct.
    struct Rec {
        RecTables host;
        RecTask* tasks = nullptr;
        ChunkRef* chunks = nullptr;
        double* lscale = nullptr;
        double* rscale = nullptr;
        bool uploaded = false;
        std::mutex lock;
    };
    mutable Rec rec;
    // Optional per-pass event timing (qh_profile_begin / qh_profile_end).
    struct Profile {
        std::vector<void*> events;  // hipEvent_t, 3 per apply
        std::uint64_t used = 0;     // applies recorded
        bool armed = false;
        std::mutex lock;
    };
    mutable Profile profile;
};

// Everything the device store holds, built on the host before the upload.
struct Plan {
    std::uint64_t order = 0;
    std::vector<std::int32_t> perm;
    std::vector<std::uint32_t> words;  // item data, 16-byte aligned runs, + padding
    std::vector<std::int32_t> widths;
    std::vector<double> cscale;
    std::vector<ItemDev> items;
    std::vector<TileDev> tiles;
    std::vector<WaveHead> heads;  // 2 x tiles x kWaves
    std::vector<std::int64_t> block_to_lowrank;
    std::vector<std::vector<ChunkRef>> left_chunks, right_chunks;
    std::vector<std::uint32_t> comp_base, ranks;
    std::uint64_t z_rows = 0, stored_words = 0;
    qh_store_stats stats{};
    RecTables rec;
    // wide passes
    std::vector<ItemDev> witems;
    std::vector<WideRun> wruns_r, wruns_l;
    std::vector<WideSum> wsums;
    std::vector<ZSum> wzsums;
    std::vector<Span> ctiles;
    std::uint64_t wz_rows = 0;  // partial-sum rows of the wide passes
    std::uint32_t wsets = 0;    // row sets of the wide passes
};

// Throw std::invalid_argument / std::runtime_error like the host store.
Plan build_plan(const qh_store_desc& desc);
// Runs both passes' index arithmetic on the host against the plan's arrays,
// with every array access bounds-checked (std::out_of_range), and returns
// y = H x in FP64.  A check of the layout without a device.
void emulate_plan(const Plan& plan, const double* x, std::int64_t k, double* y);
// The same for the wide passes' tables, in one precision's virtual-row order.
void emulate_wide(const Plan& plan, const double* x, std::int64_t k, double* y, bool fp32);
DeviceStore* create_device_store(const qh_store_desc& desc, int device);

// One launch chain of a batch on the narrow passes: the members applied with
// kb (1, 2 or 4) right-hand sides per pass.
struct BatchClass {
    int kb = 1;
    std::uint64_t member_count = 0;
    std::int64_t max_k = 1;
    std::uint32_t total_tiles = 0;
    MemberView* members = nullptr;        // device table; x, y, zp hold offsets
    std::uint32_t* first_tile = nullptr;  // device, member_count + 1 prefix counts
};

// Workgroups of one apply's launches.
struct WideCounts {
    std::uint32_t right_groups = 0;  // pass R (kWideWaves runs each)
    std::uint32_t zsums = 0;         // partial-sum reductions between the passes
    std::uint32_t left_groups = 0;   // pass L (one run each)
    std::uint32_t sums = 0;          // row tiles with several pass-L runs
};

// One class of a batch: the members applied with kc right-hand sides per
// pass, their tables in device memory (x, y and the workspace as offsets in
// doubles, doubles and bytes from the call's bases) and the prefix counts
// of each launch's workgroups.
struct WideBatchClass {
    int kc = 16;
    std::uint64_t member_count = 0;
    std::int64_t max_k = 1;
    std::uint32_t right_groups = 0, left_groups = 0, sums = 0, zsums = 0;
    WideView* members = nullptr;
    std::uint32_t* first_zsum = nullptr;
    std::uint32_t* first_right = nullptr;
    std::uint32_t* first_left = nullptr;
    std::uint32_t* first_sum = nullptr;
};

// A fixed set of resident stores applied together.  Member i reads its
// order_i x k_i right-hand sides at x + offset[i] (row-major, leading
// dimension k_i), writes y at the same offset, and uses workspace doubles
// from zoffset[i].
struct Batch {
    int device = 0;
    // graphs captured from this batch (they replay its member tables); the
    // batch refuses destruction while any exist
    std::atomic<int> pins{0};
    std::vector<const DeviceStore*> stores;
    std::vector<std::int64_t> k;
    std::vector<std::uint64_t> offset, zoffset;
    std::uint64_t total = 0, z_total = 0;
    std::vector<BatchClass> classes;
    std::vector<WideBatchClass> wclasses;
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
// Bytes of the static device scratch that qh_apply_noalloc uses as its
// workspace (one array per device, held until the process ends).
constexpr std::uint64_t kScratchBytes = std::uint64_t{64} << 20;
// qh_apply with that scratch as the workspace: no allocation of any kind.
// Rejects (std::invalid_argument) an apply that needs more workspace.
void apply_noalloc(const DeviceStore& s, const double* x, std::int64_t ldx, std::int64_t k,
                   double* y, std::int64_t ldy, int precision, void* stream);
void decode_codes(const DeviceStore& s, std::uint64_t block, int side, std::int32_t* out,
                  void* stream);
// Dense n x n FP64 matrix in source order into device memory (row-major,
// leading dimension ld), stream-ordered.
void reconstruct(const DeviceStore& s, double* out, std::int64_t ld, void* stream);

// A list of factor sides to decode into FP32 (code x its side's scale) in
// one launch, for batched-GEMM applications of the store: side 0 of a block
// as rows x rank row-major, side 1 as rank x cols row-major (V^T), each at
// its slot's offset (in floats) of the output.
struct FactorTask {
    std::uint64_t word;    // first word of the chunk's side block
    std::uint64_t offset;  // output offset of the slot (floats)
    std::uint32_t first, count, j0, jn;  // rows [first, first+count), components [j0, j0+jn)
    std::uint32_t comp_base, side, rows, rank;
};
static_assert(sizeof(FactorTask) == 48, "FactorTask is 48 bytes");
struct FactorPlan {
    const DeviceStore* store = nullptr;
    FactorTask* tasks = nullptr;  // device
    std::uint32_t count = 0;
};
FactorPlan* create_factor_plan(const DeviceStore& s, const std::uint64_t* blocks, const int* sides,
                               const std::uint64_t* offsets, std::uint64_t count,
                               std::uint64_t out_floats);
void factor_plan_decode(const FactorPlan& p, float* out, void* stream);
void destroy_factor_plan(FactorPlan* p);
int device_count();
void profile_begin(DeviceStore& s, std::uint64_t max_applies);
void set_rhs_block(DeviceStore& s, std::int64_t k, int kb);
qh_profile profile_end(DeviceStore& s);

}  // namespace qhamd
