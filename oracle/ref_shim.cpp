// C entry points over the reference's own compiled core (core.cpp, io.cpp),
// built from the reference sources where they lie by oracle/Makefile into
// oracle/_ref/.  TEST INFRASTRUCTURE ONLY: it is the checker for the oracle
// restatement and the "reference" CPU baseline; the product never loads it.
#include "qhmat/core.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace {
thread_local std::string last_error;

template <class F> int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        last_error = e.what();
        return -1;
    }
}

struct PartitionResult {
    std::vector<std::size_t> permutation;
    std::vector<std::size_t> rects;  // row0,row1,col0,col1,admissible per block
};
}  // namespace

extern "C" {

const char* ref_last_error() { return last_error.c_str(); }

void* ref_create(std::size_t order, const std::size_t* perm) {
    qhmat::HMatrix* out = nullptr;
    guard([&] { out = new qhmat::HMatrix(order, std::vector<std::size_t>(perm, perm + order)); });
    return out;
}

void ref_destroy(void* h) { delete static_cast<qhmat::HMatrix*>(h); }

int ref_add_dense(void* h, std::size_t r0, std::size_t r1, std::size_t c0, std::size_t c1,
                  const double* values) {
    return guard([&] {
        std::vector<double> v(values, values + (r1 - r0) * (c1 - c0));
        static_cast<qhmat::HMatrix*>(h)->add_dense(qhmat::Rectangle(r0, r1, c0, c1, false), v);
    });
}

// u: m x rank column-major, v: n x rank column-major (the layout bindings.cpp:138-139 builds).
int ref_add_low_rank(void* h, std::size_t r0, std::size_t r1, std::size_t c0, std::size_t c1,
                     int admissible, const double* u, const double* sigma, std::size_t rank,
                     const double* v, const int* widths, std::size_t menu_size,
                     double truncation_error) {
    return guard([&] {
        const std::size_t m = r1 - r0, n = c1 - c0;
        std::vector<double> uu(u, u + m * rank), vv(v, v + n * rank), s(sigma, sigma + rank);
        std::vector<int> w(widths, widths + rank);
        static_cast<qhmat::HMatrix*>(h)->add_low_rank(
            qhmat::Rectangle(r0, r1, c0, c1, admissible != 0), uu, s, vv, w, menu_size,
            truncation_error);
    });
}

int ref_finalize(void* h) { return guard([&] { static_cast<qhmat::HMatrix*>(h)->finalize(); }); }

int ref_matmul(void* h, const double* x, std::size_t rhs, double* y) {
    return guard([&] {
        auto* m = static_cast<qhmat::HMatrix*>(h);
        std::vector<double> xx(x, x + m->order() * rhs);
        auto out = m->matmul(xx, rhs);
        std::memcpy(y, out.data(), out.size() * sizeof(double));
    });
}

int ref_reconstruct(void* h, double* out) {
    return guard([&] {
        auto d = static_cast<qhmat::HMatrix*>(h)->reconstruct();
        std::memcpy(out, d.data(), d.size() * sizeof(double));
    });
}

double ref_reconstruction_bound(void* h) {
    return static_cast<qhmat::HMatrix*>(h)->reconstruction_bound();
}
std::uint64_t ref_logical_bits(void* h) { return static_cast<qhmat::HMatrix*>(h)->logical_bits(); }
std::uint64_t ref_allocated_bytes(void* h) {
    return static_cast<qhmat::HMatrix*>(h)->allocated_bytes();
}
std::size_t ref_order(void* h) { return static_cast<qhmat::HMatrix*>(h)->order(); }

int ref_save(void* h, const char* path) {
    return guard([&] { static_cast<qhmat::HMatrix*>(h)->save(path); });
}

void* ref_load(const char* path) {
    qhmat::HMatrix* out = nullptr;
    guard([&] { out = new qhmat::HMatrix(qhmat::HMatrix::load(path)); });
    return out;
}

double ref_coefficient(int bits) {
    double c = 0;
    if (guard([&] { c = qhmat::coefficient(bits); })) return -1.0;
    return c;
}

int ref_allocate(const double* sigma, std::size_t rank, const int* menu, std::size_t menu_size,
                 double budget, std::size_t rows, std::size_t cols, int* bits_out,
                 double* bound, std::uint64_t* payload_bits, int* feasible) {
    return guard([&] {
        auto a = qhmat::allocate(std::vector<double>(sigma, sigma + rank),
                                 std::vector<int>(menu, menu + menu_size), budget, rows, cols);
        for (std::size_t j = 0; j < a.bits.size(); ++j) bits_out[j] = a.bits[j];
        *bound = a.bound;
        *payload_bits = a.payload_bits;
        *feasible = a.feasible ? 1 : 0;
    });
}

void* ref_partition(const double* points, const double* lower, const double* upper,
                    std::size_t count, std::size_t dim, std::size_t leaf, double eta) {
    PartitionResult* out = nullptr;
    guard([&] {
        const std::size_t total = count * dim;
        auto p = qhmat::spatial_partition(std::vector<double>(points, points + total),
                                          std::vector<double>(lower, lower + total),
                                          std::vector<double>(upper, upper + total), count, dim,
                                          leaf, eta);
        out = new PartitionResult;
        out->permutation = p.permutation;
        for (const auto& r : p.blocks) {
            out->rects.insert(out->rects.end(), {r.row0, r.row1, r.col0, r.col1,
                                                 std::size_t(r.admissible ? 1 : 0)});
        }
    });
    return out;
}

std::size_t ref_partition_blocks(void* p) {
    return static_cast<PartitionResult*>(p)->rects.size() / 5;
}
void ref_partition_copy(void* p, std::size_t* perm, std::size_t* rects) {
    auto* r = static_cast<PartitionResult*>(p);
    std::memcpy(perm, r->permutation.data(), r->permutation.size() * sizeof(std::size_t));
    std::memcpy(rects, r->rects.data(), r->rects.size() * sizeof(std::size_t));
}
void ref_partition_free(void* p) { delete static_cast<PartitionResult*>(p); }

}  // extern "C"
