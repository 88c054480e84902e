/*
 * qhmat_hip.h -- C ABI of the MI355X (gfx950) H-matrix apply library
 * (qhmat_amd/libqhmat_hip.so).
 *
 * Two groups of entry points:
 *
 *  1. The device path -- the drop-in for the reference's GPU backend.  The
 *     reference binding NativeHMatrix.matmul(x, use_cuda=True)
 *     (src/qhmat/bindings.cpp:152-179) calls HMatrix::matmul_cuda
 *     (src/qhmat/cuda.cu:106-164), which flattens and uploads the whole store
 *     on every call.  Here the store is uploaded once (qh_store_create) and
 *     every apply (qh_apply) works on device pointers on a caller stream.
 *
 *  2. The host store -- what NativeHMatrix holds (src/qhmat/core.hpp:52-102,
 *     bindings.cpp:124-198): block insertion, packing, finalize, the host
 *     apply, reconstruction, accounting and QHMAT001 files.  It is the
 *     source the device store is uploaded from.
 *
 * All functions return a qh_status; on failure qh_last_error() returns the
 * message for the calling thread.  Status codes map onto the exceptions the
 * reference raises: QH_EINVAL -> std::invalid_argument (Python ValueError),
 * QH_ERUNTIME / QH_ENODEV -> std::runtime_error (RuntimeError),
 * QH_EOVERFLOW -> std::overflow_error (OverflowError), QH_ENOMEM -> MemoryError.
 *
 * No symbol here takes a torch type: pointers, sizes and a hipStream_t passed
 * as void*.
 */
#ifndef QHMAT_HIP_H
#define QHMAT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    QH_OK = 0,
    QH_EINVAL = 1,
    QH_ERUNTIME = 2,
    QH_EOVERFLOW = 3,
    QH_ENOMEM = 4,
    QH_ENODEV = 5
} qh_status;

/* Arithmetic of the off-diagonal code x vector products. */
typedef enum {
    /* FP64 throughout: the parity anchor against the FP64 CPU apply. */
    QH_PRECISION_FP64 = 0,
    /* Codes of <= 24 bits as FP32 (exact); wider codes split into a high
     * half and a sign-extended low 16-bit half, each exact in FP32 and
     * accumulated apart.  FP32 accumulation inside a block; scales, the
     * diagonal blocks and the cross-block sum into y in FP64. */
    QH_PRECISION_FP32SPLIT = 1
} qh_precision;

typedef struct qh_host qh_host;   /* host packed store (NativeHMatrix)        */
typedef struct qh_store qh_store; /* device-resident store, uploaded once     */

/* Flat, read-only view of a finalized host store in the reference's packed
 * format.  This is the shape any caller holding a reference HMatrix can
 * produce from its private members (core.hpp:77-95): per block rectangle,
 * dense values or widths + scales + the left and right word streams, where
 * component j of a factor with `rows` rows occupies ceil(rows*b_j/32) words
 * starting right after component j-1 (io.cpp:233-243). */
typedef struct {
    uint64_t order;
    const uint64_t *perm;        /* [order] clustered position -> source index */
    uint64_t nblocks;
    const uint64_t *rect;        /* [4*nblocks] row0,row1,col0,col1            */
    const uint8_t *is_dense;     /* [nblocks]                                  */
    const uint64_t *rank;        /* [nblocks] retained components (0 if dense) */
    const double *const *dense;  /* [nblocks] row-major m*m values, dense only */
    const int32_t *const *widths;      /* [nblocks] -> rank widths            */
    const double *const *lscale;       /* [nblocks] -> rank left scales       */
    const double *const *rscale;       /* [nblocks] -> rank right scales      */
    const uint32_t *const *lwords;     /* [nblocks] -> left factor words      */
    const uint32_t *const *rwords;     /* [nblocks] -> right factor words     */
} qh_store_desc;

/* Resident sizes and per-apply algorithmic counts of a device store. */
typedef struct {
    uint64_t order, nblocks, ndense, nlowrank, ncomponents;
    uint64_t code_words;          /* packed code stream, 32-bit words         */
    uint64_t left_code_words, right_code_words;
    uint64_t codes;               /* sum over low-rank blocks of (m+n)*r      */
    uint64_t dense_values;        /* FP64 values in diagonal blocks           */
    uint64_t device_bytes;        /* everything resident on the device        */
    uint64_t lowrank_flops_per_rhs; /* sum 2 r (m+n)                          */
    uint64_t dense_flops_per_rhs;   /* sum 2 m n over dense blocks            */
    uint64_t tiles;               /* row tiles (at most 64 rows each)         */
    uint64_t items;               /* work items of both passes                */
    uint64_t right_waves;         /* pass-1 waves streaming right factors     */
    uint64_t pass_code_words[2];  /* of code_words: read by pass 1 / pass 2    */
    uint64_t fused_components;    /* components of blocks in one row and one
                                     column tile (both factors in one item)   */
    uint64_t pass_fused_components[2]; /* of fused_components: pass 1 / 2     */
    uint64_t pass_dense_values[2];     /* of dense_values: pass 1 / pass 2    */
    int32_t device;
} qh_store_stats;

const char *qh_last_error(void);
/* Version of this library ("major.minor.patch"). */
const char *qh_version(void);

/* ---- device path ------------------------------------------------------- */

/* Number of visible HIP devices (0 when none). */
qh_status qh_device_count(int *count);

/* Uploads a finalized store to `device` once.  Replaces the per-call flatten +
 * upload at cuda.cu:119-150. */
qh_status qh_store_create(const qh_store_desc *desc, int device, qh_store **out);
/* Refused (QH_EINVAL, the store kept) while a batch, a factor plan or a
 * server holds it. */
qh_status qh_store_destroy(qh_store *store);
qh_status qh_store_get_stats(const qh_store *store, qh_store_stats *out);

/* Launch-shape choice (no counterpart in the reference): applies of `k`
 * right-hand sides run in passes of `kb` right-hand sides, kb in {1, 2, 4}
 * (vector ALUs) or {16, 32, 64} (matrix cores); 0 restores the default (k
 * rounded up to one of them, 64 for k > 32).  Results are bitwise the same for
 * every kb of one of the two groups and agree to the precision's tolerance
 * across them; the autotuner (qhmat_amd/tune.py) times the choices and keeps
 * the fastest.  Batches created afterwards group members by it. */
qh_status qh_store_set_rhs_block(qh_store *store, int64_t k, int kb);

/* Workspace bytes qh_apply needs for `k` right-hand sides. */
qh_status qh_apply_workspace_bytes(const qh_store *store, int64_t k, uint64_t *bytes);

/* y = H x on the device (cuda.cu:106-164 semantics, minus the copies).
 * x: order x k FP64, row-major, leading dimension ldx >= k, source order.
 * y: order x k FP64, leading dimension ldy >= k; every entry is written.
 * stream: a hipStream_t (NULL = null stream); the call is stream-ordered and
 * does not synchronise.  workspace: device memory of at least
 * qh_apply_workspace_bytes, or NULL to let the library take it from the
 * stream-ordered allocator.  x must be finite (callers check; the numpy path
 * of the Python wrapper raises the reference's "nonfinite right-hand side").
 * Concurrent calls on one store are safe when each has its own workspace. */
qh_status qh_apply(const qh_store *store, const double *x, int64_t ldx, int64_t k, double *y,
                   int64_t ldy, int precision, void *stream, void *workspace,
                   uint64_t workspace_bytes);

/* qh_apply without any allocation (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The
 * workspace is a static device array of qh_noalloc_scratch_bytes() (64 MiB)
 * in each device's memory, created when the library's code object is loaded
 * and held until the process ends; its address is looked up once per device
 * on the first call there.  An apply whose qh_apply_workspace_bytes exceeds
 * it fails with QH_EINVAL.  Every call on one device shares that device's
 * array, so calls that overlap on one device (two streams, two threads)
 * race: serialise them.  Same kernels and results as qh_apply. */
qh_status qh_apply_noalloc(const qh_store *store, const double *x, int64_t ldx, int64_t k,
                           double *y, int64_t ldy, int precision, void *stream);
qh_status qh_noalloc_scratch_bytes(uint64_t *bytes);

/* ---- batches of resident stores (no counterpart in the reference) -------
 * Many stores of uneven order, rank and right-hand-side count applied
 * together: one launch pair per right-hand-side class (k rounded up to
 * 1, 2, 4, 8 or 16) covers every member of the class.  Member i reads its
 * order_i x k_i row-major block at x + offsets[i] and writes y at the same
 * offset; the stores must live on one device.  A member's result is bitwise
 * the one qh_apply gives for it alone. */
typedef struct qh_batch qh_batch;
qh_status qh_batch_create(const qh_store *const *stores, const int64_t *k, uint64_t count,
                          qh_batch **out);
/* Refused (QH_EINVAL, the batch kept) while a graph captured from it exists;
 * a batch keeps its member stores from being destroyed until it is. */
qh_status qh_batch_destroy(qh_batch *batch);
/* offsets: count entries (doubles); total: doubles in x (and y). */
qh_status qh_batch_layout(const qh_batch *batch, uint64_t *offsets, uint64_t *total,
                          uint64_t *workspace_bytes);
qh_status qh_batch_apply(const qh_batch *batch, const double *x, double *y, int precision,
                         void *stream, void *workspace, uint64_t workspace_bytes);

/* The batch apply captured once into a HIP graph (fixed x, y, workspace and
 * precision) and replayed with one qh_graph_launch. */
typedef struct qh_graph qh_graph;
qh_status qh_batch_graph_create(const qh_batch *batch, const double *x, double *y,
                                int precision, void *workspace, uint64_t workspace_bytes,
                                qh_graph **out);
qh_status qh_graph_launch(const qh_graph *graph, void *stream);
qh_status qh_graph_destroy(qh_graph *graph);

/* Per-pass timing of qh_apply on the stream the passes are launched on.
 * qh_profile_begin arms hipEvents around the launches of the next
 * `max_applies` calls on this store (the events are created here, not in the
 * timed calls); qh_profile_end waits for the recorded events, returns the
 * totals and disarms.  Vector-ALU path: pass_ms[0] = pass 1 (right factors,
 * diagonal and fused blocks), pass_ms[1] = pass 2 (left factors and the write
 * of y); the short launch that puts x in clustered order runs before both.
 * Matrix-core path: pass_ms[0] = pass R and the partial-sum reduction,
 * pass_ms[1] = pass L and the row-set sum. */
typedef struct {
    uint64_t applies;
    double pass_ms[2];
} qh_profile;
qh_status qh_profile_begin(qh_store *store, uint64_t max_applies);
qh_status qh_profile_end(qh_store *store, qh_profile *out);

/* ---- serving loop (no counterpart in the reference) ----------------------
 * Resident stores served from native threads: requests (a store, x of
 * order x k FP64 row-major in device memory, complete when submitted, and y
 * of the same shape, written when the request completes) are queued; a
 * dispatcher thread takes every queued request of the store at the head of
 * the queue, up to max_rhs columns, as one pass on a free slot (a stream with
 * its own pass buffers and workspace) -- several requests are gathered into
 * one apply of all their columns and scattered back -- and a completion
 * thread marks them done as their passes finish.  A request's result is
 * bitwise qh_apply's for it alone when both run on the same arithmetic path
 * (vector ALUs up to 4 columns per pass, matrix cores above), and within the
 * precision's tolerance otherwise.  Latency = submission to observed
 * completion. */
typedef struct qh_server qh_server;
typedef struct {
    int32_t device;
    int32_t precision;     /* qh_precision of every pass                       */
    int32_t slots;         /* passes in flight at once (streams), 1..16        */
    int64_t max_rhs;       /* columns per pass and per request, 1..4096        */
    uint64_t hbm_budget;   /* admitted stores plus pass buffers, bytes         */
} qh_server_config;
typedef struct {
    uint64_t completed, completed_rhs, passes, resident_bytes, buffer_bytes;
    double p50_ms, p99_ms, mean_ms; /* since the last reset (0 when none)    */
} qh_server_stats;
typedef struct {
    double offered_rate;        /* requests per second asked for             */
    double offered_rate_drawn;  /* sessions drawn x session length / duration */
    uint64_t sessions, requests, passes, errors;
    double seconds, throughput_rps, p50_ms, p99_ms, mean_ms;
} qh_load_result;
enum { QH_TICKET_PENDING = 0, QH_TICKET_DONE = 1, QH_TICKET_FAILED = 2 };

qh_status qh_server_create(const qh_server_config *config, qh_server **out);
/* Stops the threads after the passes in flight; queued requests are dropped. */
qh_status qh_server_destroy(qh_server *server);
/* Keeps `store` resident for the server (QH_EINVAL past the HBM budget) and
 * returns its id; the store refuses destruction until evicted or the server
 * is destroyed.  Waits for the passes in flight. */
qh_status qh_server_admit(qh_server *server, const qh_store *store, int32_t *id);
/* Queued requests of the store fail; passes already launched finish. */
qh_status qh_server_evict(qh_server *server, int32_t id);
qh_status qh_server_submit(qh_server *server, int32_t id, const double *x, int64_t k, double *y,
                           uint64_t *ticket);
/* Waits up to timeout_s seconds (< 0: no limit); *state is a QH_TICKET_*.
 * A finished ticket is forgotten once reported. */
qh_status qh_server_wait(qh_server *server, uint64_t ticket, double timeout_s, int32_t *state);
qh_status qh_server_get_stats(qh_server *server, qh_server_stats *out);
qh_status qh_server_reset_stats(qh_server *server);
/* Closed-loop load (a GMRES client per session): sessions start at Poisson
 * times, rate / session per second, each on a store of `ids` and a width of
 * `ks` drawn uniformly, and issue `session` requests one after another, the
 * next from the completion thread as soon as the previous one is done.  One
 * request per (store, width) runs first and is dropped from the statistics.
 * The operands are generated on the device; sessions on one (store, width)
 * share them (the results are not read).  Returns when every session ends. */
qh_status qh_server_closed_loop(qh_server *server, const int32_t *ids, uint64_t nids,
                                const int64_t *ks, uint64_t nks, double rate, double duration,
                                int32_t session, uint64_t seed, qh_load_result *out);

/* Layout check without a device: builds the device plan of `desc` on the
 * host and runs both passes' index arithmetic against its arrays with every
 * access bounds-checked, writing y = H x (FP64, order x k row-major).  A
 * test hook for the tile-major layout; no apply path calls it. */
qh_status qh_plan_emulate(const qh_store_desc *desc, const double *x, int64_t k, double *y);
/* The same check for the tables of the wide (16/32/64 right-hand sides per
 * pass, matrix-core) passes, whose partial-sum rows follow the precision's
 * order of virtual rows (`precision` a qh_precision). */
qh_status qh_plan_emulate_wide(const qh_store_desc *desc, const double *x, int64_t k, double *y,
                               int precision);

/* Dense reconstruction on the device: writes the order x order FP64 matrix
 * in source index order into device memory `out` (row-major, leading
 * dimension ld >= order), every entry formed as the host reconstruction
 * forms it, so the result is bitwise the host's.  Stream-ordered; the first
 * call on a store uploads its reconstruction tables (per-tile tasks, chunk
 * offsets, per-component scales).
 * Replaces HMatrix::reconstruct (src/qhmat/core.cpp:392-410), bound as
 * NativeHMatrix.reconstruct (src/qhmat/bindings.cpp:180-190) and
 * HMatrix.to_dense (python/qhmat/matrix.py:309-311), for large-order checks
 * without a host pass over n^2 entries. */
qh_status qh_reconstruct(const qh_store *store, double *out, int64_t ld, void *stream);

/* Test hook: the integer codes of one side (0 = left, 1 = right) of block
 * `block` (insertion index), decoded on the device from the resident words
 * into out[j*rows + i] (int32, device memory, rank*rows entries).  They must
 * equal the host unpacker bit for bit (core.cpp:269-280). */
qh_status qh_decode_codes(const qh_store *store, uint64_t block, int side, int32_t *out,
                          void *stream);

/* Decoded FP32 factors for batched-GEMM applications of a resident store
 * (the application harness's methods that decode on every application:
 * experiments/cuda/application.cu:189-222, `adaptive_unpacked_batched`, whose
 * unpack_jobs kernel this replaces; no counterpart in the library's API).
 * A plan lists factor sides: block `blocks[i]` (insertion index, a low-rank
 * block of rank > 0), side `sides[i]` (0: the left factor as rows x rank
 * row-major; 1: the right factor transposed, rank x cols row-major), written
 * at `offsets[i]` floats into an output of `out_floats` floats.  Each value
 * is float(code x that side's component scale).  qh_factor_plan_decode
 * decodes every listed side from the resident words in one stream-ordered
 * launch; the plan holds a task table in device memory until destroyed,
 * and the store refuses destruction (QH_EINVAL) while a plan on it exists. */
typedef struct qh_factor_plan qh_factor_plan;
qh_status qh_factor_plan_create(const qh_store *store, const uint64_t *blocks, const int32_t *sides,
                                const uint64_t *offsets, uint64_t count, uint64_t out_floats,
                                qh_factor_plan **plan);
qh_status qh_factor_plan_decode(const qh_factor_plan *plan, float *out, void *stream);
qh_status qh_factor_plan_destroy(qh_factor_plan *plan);

/* ---- host store (NativeHMatrix, bindings.cpp:124-198) ------------------- */

qh_status qh_host_create(uint64_t order, const uint64_t *perm, qh_host **out);
qh_status qh_host_destroy(qh_host *h);
/* core.cpp:224-233; values row-major (row1-row0) x (col1-col0). */
qh_status qh_host_add_dense(qh_host *h, uint64_t row0, uint64_t row1, uint64_t col0,
                            uint64_t col1, int admissible, const double *values,
                            uint64_t count);
/* core.cpp:284-330; u: m x rank and v: n x rank column-major, sigma: rank. */
qh_status qh_host_add_low_rank(qh_host *h, uint64_t row0, uint64_t row1, uint64_t col0,
                               uint64_t col1, int admissible, const double *u,
                               const double *sigma, uint64_t rank, const double *v,
                               const int32_t *widths, uint64_t width_count,
                               uint64_t menu_size, double truncation_error);
qh_status qh_host_finalize(qh_host *h);
/* core.cpp:360-390 semantics on the host (device="cpu"): x, y order x rhs. */
qh_status qh_host_matmul(const qh_host *h, const double *x, uint64_t rhs, double *y,
                         int threads);
/* core.cpp:392-410: order*order values, source order. */
qh_status qh_host_reconstruct(const qh_host *h, double *out);

typedef struct {
    uint64_t order, nblocks;
    int32_t finalized;
    uint64_t logical_bits, allocated_bytes, resident_bytes;
    double reconstruction_bound;
} qh_host_info;
qh_status qh_host_get_info(const qh_host *h, qh_host_info *out);

typedef struct {
    uint64_t row0, row1, col0, col1;
    int32_t admissible, dense;
    uint64_t rank;
    double truncation_error, quantization_bound;
    uint64_t logical_bits, allocated_bytes;
} qh_block_info;
/* core.cpp:412-436 for block `index` (insertion order). */
qh_status qh_host_block_info(const qh_host *h, uint64_t index, qh_block_info *out);
/* Clustered -> source permutation, order entries. */
qh_status qh_host_permutation(const qh_host *h, uint64_t *out);
/* Words / scales / widths of one block side (0 left, 1 right): sizes first
 * (pass NULL buffers), then the copy. */
qh_status qh_host_block_words(const qh_host *h, uint64_t index, int side, uint32_t *words,
                              uint64_t *word_count, double *scales, int32_t *widths);
/* io.cpp:127-257 format.  save refuses to replace an existing file. */
qh_status qh_host_save(const qh_host *h, const char *path);
qh_status qh_host_load(const char *path, qh_host **out);
/* Fills `desc` with pointers into `h` (valid while h lives, unmodified).
 * The pointer tables are owned by h. */
qh_status qh_host_describe(qh_host *h, qh_store_desc *desc);

/* ---- construction helpers (core.cpp:56-203) ----------------------------- */

qh_status qh_coefficient(int bits, double *out);
/* core.cpp:61-115.  bits_out: rank entries. */
qh_status qh_allocate(const double *sigma, uint64_t rank, const int32_t *menu,
                      uint64_t menu_count, double budget, uint64_t rows, uint64_t cols,
                      int32_t *bits_out, double *bound, uint64_t *payload_bits,
                      int32_t *feasible);

typedef struct qh_partition qh_partition;
/* core.cpp:117-203; points/lower/upper: count x dim row-major. */
qh_status qh_partition_create(const double *points, const double *lower, const double *upper,
                              uint64_t count, uint64_t dim, uint64_t leaf_size, double eta,
                              qh_partition **out);
qh_status qh_partition_sizes(const qh_partition *p, uint64_t *count, uint64_t *nblocks);
/* perm: count entries; rects: 5 per block (row0,row1,col0,col1,admissible). */
qh_status qh_partition_copy(const qh_partition *p, uint64_t *perm, uint64_t *rects);
qh_status qh_partition_destroy(qh_partition *p);

#ifdef __cplusplus
}
#endif
#endif /* QHMAT_HIP_H */
