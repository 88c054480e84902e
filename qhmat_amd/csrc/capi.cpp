// extern "C" surface of libqhmat_hip.so (declarations and reference
// citations in include/qhmat_hip.h).  Exceptions never cross the ABI: they
// become a status code plus a per-thread message.
#include <cstring>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qhmat_hip.h"
#include "device_store.hpp"
#include "host_store.hpp"
#include "server.hpp"

struct qh_host {
    qhamd::HostStore store;
    // Pointer tables handed out by qh_host_describe.
    std::vector<std::uint64_t> rect, rank;
    std::vector<std::uint8_t> is_dense;
    std::vector<const double*> dense, lscale, rscale;
    std::vector<const std::int32_t*> widths;
    std::vector<std::vector<std::int32_t>> widths32;
    std::vector<const std::uint32_t*> lwords, rwords;
    explicit qh_host(qhamd::HostStore s) : store(std::move(s)) {}
};

struct qh_store {
    qhamd::DeviceStore* dev = nullptr;
};

struct qh_partition {
    qhamd::PartitionOut out;
};

struct qh_batch {
    qhamd::Batch* batch = nullptr;
};

struct qh_server {
    qhamd::Server* server = nullptr;
};

struct qh_graph {
    void* exec = nullptr;  // hipGraphExec_t
    qhamd::Batch* batch = nullptr;  // pinned while the graph exists
};

namespace {
thread_local std::string g_error;

template <class F> qh_status guarded(F&& f) {
    try {
        f();
        return QH_OK;
    } catch (const std::invalid_argument& e) {
        g_error = e.what();
        return QH_EINVAL;
    } catch (const std::overflow_error& e) {
        g_error = e.what();
        return QH_EOVERFLOW;
    } catch (const std::bad_alloc&) {
        g_error = "out of memory";
        return QH_ENOMEM;
    } catch (const std::exception& e) {
        g_error = e.what();
        const bool nodev = std::strstr(e.what(), "no CUDA device") != nullptr;
        return nodev ? QH_ENODEV : QH_ERUNTIME;
    } catch (...) {
        g_error = "unknown error";
        return QH_ERUNTIME;
    }
}

void need(const void* p, const char* what) {
    if (!p) throw std::invalid_argument(std::string("null ") + what);
}

qhamd::Rect rect_of(std::uint64_t r0, std::uint64_t r1, std::uint64_t c0, std::uint64_t c1,
                    int admissible) {
    return qhamd::Rect{r0, r1, c0, c1, admissible != 0};
}
}  // namespace

extern "C" {

const char* qh_last_error(void) { return g_error.c_str(); }
const char* qh_version(void) { return "0.1.0"; }

qh_status qh_device_count(int* count) {
    return guarded([&] {
        need(count, "count");
        *count = qhamd::device_count();
    });
}

qh_status qh_store_create(const qh_store_desc* desc, int device, qh_store** out) {
    return guarded([&] {
        need(desc, "description");
        need(out, "output handle");
        *out = nullptr;
        auto* s = new qh_store;
        try {
            s->dev = qhamd::create_device_store(*desc, device);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

qh_status qh_store_destroy(qh_store* store) {
    return guarded([&] {
        if (!store) return;
        qhamd::destroy_device_store(store->dev);
        delete store;
    });
}

qh_status qh_store_get_stats(const qh_store* store, qh_store_stats* out) {
    return guarded([&] {
        need(store, "store");
        need(out, "stats");
        *out = store->dev->stats;
    });
}

qh_status qh_apply_workspace_bytes(const qh_store* store, int64_t k, uint64_t* bytes) {
    return guarded([&] {
        need(store, "store");
        need(bytes, "bytes");
        if (k <= 0) throw std::invalid_argument("right-hand-side shape mismatch");
        *bytes = qhamd::apply_workspace_bytes(*store->dev, k);
    });
}

qh_status qh_apply(const qh_store* store, const double* x, int64_t ldx, int64_t k, double* y,
                   int64_t ldy, int precision, void* stream, void* workspace,
                   uint64_t workspace_bytes) {
    return guarded([&] {
        need(store, "store");
        qhamd::apply(*store->dev, x, ldx, k, y, ldy, precision, stream, workspace,
                     workspace_bytes);
    });
}

qh_status qh_apply_noalloc(const qh_store* store, const double* x, int64_t ldx, int64_t k,
                           double* y, int64_t ldy, int precision, void* stream) {
    return guarded([&] {
        need(store, "store");
        qhamd::apply_noalloc(*store->dev, x, ldx, k, y, ldy, precision, stream);
    });
}

qh_status qh_noalloc_scratch_bytes(uint64_t* bytes) {
    return guarded([&] {
        need(bytes, "bytes");
        *bytes = qhamd::kScratchBytes;
    });
}

qh_status qh_batch_create(const qh_store* const* stores, const int64_t* k, uint64_t count,
                          qh_batch** out) {
    return guarded([&] {
        need(stores, "stores");
        need(k, "right-hand-side counts");
        need(out, "output handle");
        *out = nullptr;
        std::vector<const qhamd::DeviceStore*> dev(count);
        for (uint64_t i = 0; i < count; ++i) {
            need(stores[i], "store");
            dev[i] = stores[i]->dev;
        }
        auto* b = new qh_batch;
        try {
            b->batch = qhamd::create_batch(dev.data(), k, count);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

qh_status qh_batch_destroy(qh_batch* batch) {
    return guarded([&] {
        if (!batch) return;
        qhamd::destroy_batch(batch->batch);
        delete batch;
    });
}

qh_status qh_batch_layout(const qh_batch* batch, uint64_t* offsets, uint64_t* total,
                          uint64_t* workspace_bytes) {
    return guarded([&] {
        need(batch, "batch");
        const auto& b = *batch->batch;
        if (offsets)
            for (std::size_t i = 0; i < b.offset.size(); ++i) offsets[i] = b.offset[i];
        if (total) *total = b.total;
        if (workspace_bytes) *workspace_bytes = b.z_total * sizeof(double);
    });
}

qh_status qh_batch_apply(const qh_batch* batch, const double* x, double* y, int precision,
                         void* stream, void* workspace, uint64_t workspace_bytes) {
    return guarded([&] {
        need(batch, "batch");
        qhamd::batch_apply(*batch->batch, x, y, precision, stream, workspace, workspace_bytes);
    });
}

qh_status qh_batch_graph_create(const qh_batch* batch, const double* x, double* y,
                                int precision, void* workspace, uint64_t workspace_bytes,
                                qh_graph** out) {
    return guarded([&] {
        need(batch, "batch");
        need(out, "output handle");
        *out = nullptr;
        void* exec = qhamd::batch_graph(*batch->batch, x, y, precision, workspace,
                                        workspace_bytes);
        batch->batch->pins.fetch_add(1);
        *out = new qh_graph{exec, batch->batch};
    });
}

qh_status qh_graph_launch(const qh_graph* graph, void* stream) {
    return guarded([&] {
        need(graph, "graph");
        qhamd::graph_launch(graph->exec, stream);
    });
}

qh_status qh_graph_destroy(qh_graph* graph) {
    return guarded([&] {
        if (!graph) return;
        qhamd::graph_destroy(graph->exec);
        if (graph->batch) graph->batch->pins.fetch_sub(1);
        delete graph;
    });
}

qh_status qh_plan_emulate(const qh_store_desc* desc, const double* x, int64_t k, double* y) {
    return guarded([&] {
        need(desc, "description");
        need(x, "right-hand side");
        need(y, "result");
        if (k <= 0) throw std::invalid_argument("right-hand-side shape mismatch");
        const qhamd::Plan plan = qhamd::build_plan(*desc);
        qhamd::emulate_plan(plan, x, k, y);
    });
}

qh_status qh_plan_emulate_wide(const qh_store_desc* desc, const double* x, int64_t k,
                               double* y, int precision) {
    return guarded([&] {
        need(desc, "description");
        need(x, "right-hand side");
        need(y, "result");
        if (k <= 0) throw std::invalid_argument("right-hand-side shape mismatch");
        const qhamd::Plan plan = qhamd::build_plan(*desc);
        if (precision != QH_PRECISION_FP64 && precision != QH_PRECISION_FP32SPLIT)
            throw std::invalid_argument("unknown precision mode");
        qhamd::emulate_wide(plan, x, k, y, precision == QH_PRECISION_FP32SPLIT);
    });
}

qh_status qh_store_set_rhs_block(qh_store* store, int64_t k, int kb) {
    return guarded([&] {
        need(store, "store");
        qhamd::set_rhs_block(*store->dev, k, kb);
    });
}

qh_status qh_profile_begin(qh_store* store, uint64_t max_applies) {
    return guarded([&] {
        need(store, "store");
        qhamd::profile_begin(*store->dev, max_applies);
    });
}

qh_status qh_profile_end(qh_store* store, qh_profile* out) {
    return guarded([&] {
        need(store, "store");
        need(out, "profile");
        *out = qhamd::profile_end(*store->dev);
    });
}

qh_status qh_reconstruct(const qh_store* store, double* out, int64_t ld, void* stream) {
    return guarded([&] {
        need(store, "store");
        qhamd::reconstruct(*store->dev, out, ld, stream);
    });
}

qh_status qh_decode_codes(const qh_store* store, uint64_t block, int side, int32_t* out,
                          void* stream) {
    return guarded([&] {
        need(store, "store");
        qhamd::decode_codes(*store->dev, block, side, out, stream);
    });
}

struct qh_factor_plan {
    qhamd::FactorPlan* plan = nullptr;
};

qh_status qh_factor_plan_create(const qh_store* store, const uint64_t* blocks, const int32_t* sides,
                                const uint64_t* offsets, uint64_t count, uint64_t out_floats,
                                qh_factor_plan** plan) {
    return guarded([&] {
        need(store, "store");
        need(plan, "output handle");
        *plan = nullptr;
        std::vector<int> sd(sides, sides + (sides ? count : 0));
        auto* h = new qh_factor_plan;
        try {
            h->plan = qhamd::create_factor_plan(*store->dev, blocks, sd.data(), offsets, count,
                                                out_floats);
        } catch (...) {
            delete h;
            throw;
        }
        *plan = h;
    });
}

qh_status qh_factor_plan_decode(const qh_factor_plan* plan, float* out, void* stream) {
    return guarded([&] {
        need(plan, "plan");
        qhamd::factor_plan_decode(*plan->plan, out, stream);
    });
}

qh_status qh_factor_plan_destroy(qh_factor_plan* plan) {
    if (plan) qhamd::destroy_factor_plan(plan->plan);
    delete plan;
    return QH_OK;
}

qh_status qh_host_create(uint64_t order, const uint64_t* perm, qh_host** out) {
    return guarded([&] {
        need(out, "output handle");
        *out = nullptr;
        if (order > 0) need(perm, "permutation");
        std::vector<std::uint64_t> p(perm, perm + order);
        *out = new qh_host(qhamd::HostStore(order, std::move(p)));
    });
}

qh_status qh_host_destroy(qh_host* h) {
    delete h;
    return QH_OK;
}

qh_status qh_host_add_dense(qh_host* h, uint64_t row0, uint64_t row1, uint64_t col0,
                            uint64_t col1, int admissible, const double* values,
                            uint64_t count) {
    return guarded([&] {
        need(h, "store");
        if (count) need(values, "values");
        h->store.add_dense(rect_of(row0, row1, col0, col1, admissible), values, count);
    });
}

qh_status qh_host_add_low_rank(qh_host* h, uint64_t row0, uint64_t row1, uint64_t col0,
                               uint64_t col1, int admissible, const double* u,
                               const double* sigma, uint64_t rank, const double* v,
                               const int32_t* widths, uint64_t width_count,
                               uint64_t menu_size, double truncation_error) {
    return guarded([&] {
        need(h, "store");
        if (rank) {
            need(u, "left factor");
            need(v, "right factor");
            need(sigma, "singular values");
        }
        if (width_count) need(widths, "widths");
        std::vector<int> w(widths, widths + width_count);
        h->store.add_low_rank(rect_of(row0, row1, col0, col1, admissible), u, sigma, rank, v,
                              w.data(), width_count, menu_size, truncation_error);
    });
}

qh_status qh_host_finalize(qh_host* h) {
    return guarded([&] {
        need(h, "store");
        h->store.finalize();
    });
}

qh_status qh_host_matmul(const qh_host* h, const double* x, uint64_t rhs, double* y,
                         int threads) {
    return guarded([&] {
        need(h, "store");
        need(x, "right-hand side");
        need(y, "result");
        h->store.matmul(x, rhs, y, threads);
    });
}

qh_status qh_host_reconstruct(const qh_host* h, double* out) {
    return guarded([&] {
        need(h, "store");
        need(out, "output");
        h->store.reconstruct(out);
    });
}

qh_status qh_host_get_info(const qh_host* h, qh_host_info* out) {
    return guarded([&] {
        need(h, "store");
        need(out, "info");
        out->order = h->store.order();
        out->nblocks = h->store.blocks().size();
        out->finalized = h->store.finalized() ? 1 : 0;
        out->logical_bits = h->store.logical_bits();
        out->allocated_bytes = h->store.allocated_bytes();
        out->resident_bytes = h->store.resident_bytes();
        out->reconstruction_bound = h->store.reconstruction_bound();
    });
}

qh_status qh_host_block_info(const qh_host* h, uint64_t index, qh_block_info* out) {
    return guarded([&] {
        need(h, "store");
        need(out, "info");
        const auto& blocks = h->store.blocks();
        if (index >= blocks.size()) throw std::invalid_argument("block index out of range");
        const auto s = h->store.summary(index);
        out->row0 = s.rect.row0;
        out->row1 = s.rect.row1;
        out->col0 = s.rect.col0;
        out->col1 = s.rect.col1;
        out->admissible = s.rect.admissible ? 1 : 0;
        out->dense = s.dense ? 1 : 0;
        out->rank = s.rank;
        out->truncation_error = s.truncation_error;
        out->quantization_bound = s.quantization_bound;
        out->logical_bits = s.logical_bits;
        out->allocated_bytes = s.allocated_bytes;
    });
}

qh_status qh_host_permutation(const qh_host* h, uint64_t* out) {
    return guarded([&] {
        need(h, "store");
        need(out, "output");
        const auto& p = h->store.permutation();
        std::memcpy(out, p.data(), p.size() * sizeof(std::uint64_t));
    });
}

qh_status qh_host_block_words(const qh_host* h, uint64_t index, int side, uint32_t* words,
                              uint64_t* word_count, double* scales, int32_t* widths) {
    return guarded([&] {
        need(h, "store");
        const auto& blocks = h->store.blocks();
        if (index >= blocks.size()) throw std::invalid_argument("block index out of range");
        if (side != 0 && side != 1) throw std::invalid_argument("side must be 0 or 1");
        const auto& b = blocks[index];
        const auto& f = side == 0 ? b.left : b.right;
        if (word_count) *word_count = f.words.size();
        if (words) std::memcpy(words, f.words.data(), f.words.size() * sizeof(std::uint32_t));
        if (scales) std::memcpy(scales, f.scales.data(), f.scales.size() * sizeof(double));
        if (widths)
            for (std::size_t j = 0; j < b.widths.size(); ++j) widths[j] = b.widths[j];
    });
}

qh_status qh_host_save(const qh_host* h, const char* path) {
    return guarded([&] {
        need(h, "store");
        need(path, "path");
        h->store.save(path);
    });
}

qh_status qh_host_load(const char* path, qh_host** out) {
    return guarded([&] {
        need(path, "path");
        need(out, "output handle");
        *out = nullptr;
        *out = new qh_host(qhamd::HostStore::load(path));
    });
}

qh_status qh_host_describe(qh_host* h, qh_store_desc* desc) {
    return guarded([&] {
        need(h, "store");
        need(desc, "description");
        if (!h->store.finalized())
            throw std::invalid_argument("matrix must be finalized before use");
        const auto& blocks = h->store.blocks();
        const std::size_t nb = blocks.size();
        h->rect.assign(4 * nb, 0);
        h->rank.assign(nb, 0);
        h->is_dense.assign(nb, 0);
        h->dense.assign(nb, nullptr);
        h->lscale.assign(nb, nullptr);
        h->rscale.assign(nb, nullptr);
        h->widths.assign(nb, nullptr);
        h->widths32.assign(nb, {});
        h->lwords.assign(nb, nullptr);
        h->rwords.assign(nb, nullptr);
        for (std::size_t b = 0; b < nb; ++b) {
            const auto& blk = blocks[b];
            h->rect[4 * b] = blk.rect.row0;
            h->rect[4 * b + 1] = blk.rect.row1;
            h->rect[4 * b + 2] = blk.rect.col0;
            h->rect[4 * b + 3] = blk.rect.col1;
            h->is_dense[b] = blk.is_dense() ? 1 : 0;
            if (blk.is_dense()) {
                h->dense[b] = blk.dense.data();
                continue;
            }
            h->rank[b] = blk.widths.size();
            h->widths32[b].assign(blk.widths.begin(), blk.widths.end());
            h->widths[b] = h->widths32[b].data();
            h->lscale[b] = blk.left.scales.data();
            h->rscale[b] = blk.right.scales.data();
            h->lwords[b] = blk.left.words.data();
            h->rwords[b] = blk.right.words.data();
        }
        desc->order = h->store.order();
        desc->perm = h->store.permutation().data();
        desc->nblocks = nb;
        desc->rect = h->rect.data();
        desc->is_dense = h->is_dense.data();
        desc->rank = h->rank.data();
        desc->dense = h->dense.data();
        desc->widths = h->widths.data();
        desc->lscale = h->lscale.data();
        desc->rscale = h->rscale.data();
        desc->lwords = h->lwords.data();
        desc->rwords = h->rwords.data();
    });
}

qh_status qh_coefficient(int bits, double* out) {
    return guarded([&] {
        need(out, "output");
        *out = qhamd::coefficient(bits);
    });
}

qh_status qh_allocate(const double* sigma, uint64_t rank, const int32_t* menu,
                      uint64_t menu_count, double budget, uint64_t rows, uint64_t cols,
                      int32_t* bits_out, double* bound, uint64_t* payload_bits,
                      int32_t* feasible) {
    return guarded([&] {
        if (rank) need(sigma, "singular values");
        if (menu_count) need(menu, "menu");
        std::vector<double> s(sigma, sigma + rank);
        std::vector<int> m(menu, menu + menu_count);
        const auto a = qhamd::allocate_widths(s, m, budget, rows, cols);
        if (bits_out)
            for (std::size_t j = 0; j < a.bits.size(); ++j) bits_out[j] = a.bits[j];
        if (bound) *bound = a.bound;
        if (payload_bits) *payload_bits = a.payload_bits;
        if (feasible) *feasible = a.feasible ? 1 : 0;
    });
}

qh_status qh_partition_create(const double* points, const double* lower, const double* upper,
                              uint64_t count, uint64_t dim, uint64_t leaf_size, double eta,
                              qh_partition** out) {
    return guarded([&] {
        need(out, "output handle");
        *out = nullptr;
        const std::uint64_t total = count * dim;
        if (total) {
            need(points, "points");
            need(lower, "support_lower");
            need(upper, "support_upper");
        }
        auto* p = new qh_partition;
        try {
            p->out = qhamd::spatial_partition(std::vector<double>(points, points + total),
                                              std::vector<double>(lower, lower + total),
                                              std::vector<double>(upper, upper + total), count,
                                              dim, leaf_size, eta);
        } catch (...) {
            delete p;
            throw;
        }
        *out = p;
    });
}

qh_status qh_partition_sizes(const qh_partition* p, uint64_t* count, uint64_t* nblocks) {
    return guarded([&] {
        need(p, "partition");
        if (count) *count = p->out.permutation.size();
        if (nblocks) *nblocks = p->out.blocks.size();
    });
}

qh_status qh_partition_copy(const qh_partition* p, uint64_t* perm, uint64_t* rects) {
    return guarded([&] {
        need(p, "partition");
        if (perm)
            std::memcpy(perm, p->out.permutation.data(),
                        p->out.permutation.size() * sizeof(std::uint64_t));
        if (rects)
            for (std::size_t b = 0; b < p->out.blocks.size(); ++b) {
                const auto& r = p->out.blocks[b];
                rects[5 * b] = r.row0;
                rects[5 * b + 1] = r.row1;
                rects[5 * b + 2] = r.col0;
                rects[5 * b + 3] = r.col1;
                rects[5 * b + 4] = r.admissible ? 1 : 0;
            }
    });
}

qh_status qh_partition_destroy(qh_partition* p) {
    delete p;
    return QH_OK;
}

// ---- serving loop ---------------------------------------------------------

qh_status qh_server_create(const qh_server_config* config, qh_server** out) {
    return guarded([&] {
        need(config, "config");
        need(out, "output handle");
        *out = nullptr;
        qhamd::ServerConfig c;
        c.device = config->device;
        c.precision = config->precision;
        c.slots = config->slots;
        c.max_rhs = config->max_rhs;
        c.hbm_budget = config->hbm_budget;
        auto* s = new qh_server;
        try {
            s->server = new qhamd::Server(c);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

qh_status qh_server_destroy(qh_server* server) {
    return guarded([&] {
        if (!server) return;
        delete server->server;
        delete server;
    });
}

qh_status qh_server_admit(qh_server* server, const qh_store* store, int32_t* id) {
    return guarded([&] {
        need(server, "server");
        need(store, "store");
        need(id, "id");
        *id = server->server->admit(store->dev);
    });
}

qh_status qh_server_evict(qh_server* server, int32_t id) {
    return guarded([&] {
        need(server, "server");
        server->server->evict(id);
    });
}

qh_status qh_server_submit(qh_server* server, int32_t id, const double* x, int64_t k, double* y,
                           uint64_t* ticket) {
    return guarded([&] {
        need(server, "server");
        need(ticket, "ticket");
        *ticket = server->server->submit(id, x, k, y);
    });
}

qh_status qh_server_wait(qh_server* server, uint64_t ticket, double timeout_s, int32_t* state) {
    return guarded([&] {
        need(server, "server");
        need(state, "state");
        *state = server->server->wait(ticket, timeout_s);
    });
}

qh_status qh_server_get_stats(qh_server* server, qh_server_stats* out) {
    return guarded([&] {
        need(server, "server");
        need(out, "stats");
        const qhamd::ServerStats st = server->server->stats();
        out->completed = st.completed;
        out->completed_rhs = st.completed_rhs;
        out->passes = st.passes;
        out->resident_bytes = st.resident_bytes;
        out->buffer_bytes = st.buffer_bytes;
        out->p50_ms = st.p50_ms;
        out->p99_ms = st.p99_ms;
        out->mean_ms = st.mean_ms;
    });
}

qh_status qh_server_reset_stats(qh_server* server) {
    return guarded([&] {
        need(server, "server");
        server->server->reset_stats();
    });
}

qh_status qh_server_closed_loop(qh_server* server, const int32_t* ids, uint64_t nids,
                                const int64_t* ks, uint64_t nks, double rate, double duration,
                                int32_t session, uint64_t seed, qh_load_result* out) {
    return guarded([&] {
        need(server, "server");
        need(ids, "store ids");
        need(ks, "widths");
        need(out, "result");
        const std::vector<int> id(ids, ids + nids);
        const std::vector<std::int64_t> k(ks, ks + nks);
        const qhamd::LoadResult r = server->server->closed_loop(id, k, rate, duration, session, seed);
        out->offered_rate = r.offered_rate;
        out->offered_rate_drawn = r.offered_rate_drawn;
        out->sessions = r.sessions;
        out->requests = r.requests;
        out->passes = r.passes;
        out->errors = r.errors;
        out->seconds = r.seconds;
        out->throughput_rps = r.throughput_rps;
        out->p50_ms = r.p50_ms;
        out->p99_ms = r.p99_ms;
        out->mean_ms = r.mean_ms;
    });
}

}  // extern "C"
