// Host-side packed store of a qhmat H-matrix: what NativeHMatrix holds
// (reference src/qhmat/core.hpp:52-102), built here so that the device store
// can be uploaded from it once.  The words, offsets and scales follow the
// reference's packed format exactly (core.cpp:235-282): per component a
// word-aligned, LSB-first stream of b-bit two's-complement codes.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace qhamd {

struct Rect {
    std::uint64_t row0 = 0, row1 = 0, col0 = 0, col1 = 0;
    bool admissible = false;
};

struct PackedSide {
    std::vector<std::uint32_t> words;
    std::vector<std::uint64_t> offsets;  // first word of each component
    std::vector<double> scales;          // max|column| / q, or 0
};

struct HostBlock {
    Rect rect;
    std::vector<double> dense;  // row-major, diagonal blocks only
    std::vector<int> widths;    // one per retained component
    PackedSide left, right;     // m and n rows per component
    std::uint64_t menu_size = 0;
    double truncation_error = 0.0;
    double quantization_bound = 0.0;
    bool is_dense() const { return !dense.empty(); }
    std::uint64_t rows() const { return rect.row1 - rect.row0; }
    std::uint64_t cols() const { return rect.col1 - rect.col0; }
};

struct BlockSummary {
    Rect rect;
    bool dense;
    std::uint64_t rank;
    double truncation_error, quantization_bound;
    std::uint64_t logical_bits, allocated_bytes;
};

// q = 2^(b-1) - 1 (core.cpp:32-35); throws invalid_argument outside [2, 32].
std::uint64_t code_limit(int bits);
// c(b) = 1/q + 1/(4 q^2) (core.cpp:56-59).
double coefficient(int bits);
// Words of one component of `rows` codes at `bits` bits.
std::uint64_t words_for(std::uint64_t rows, int bits);

struct Allocation {
    std::vector<int> bits;
    double bound = 0.0;
    std::uint64_t payload_bits = 0;
    bool feasible = false;
};
// Reserve-greedy width allocation (core.cpp:61-115).
Allocation allocate_widths(const std::vector<double>& sigma, const std::vector<int>& menu,
                           double budget, std::uint64_t rows, std::uint64_t cols);

struct PartitionOut {
    std::vector<std::uint64_t> permutation;
    std::vector<Rect> blocks;
};
// Median-split cluster tree + box admissibility (core.cpp:117-203).
PartitionOut spatial_partition(const std::vector<double>& points,
                               const std::vector<double>& lower,
                               const std::vector<double>& upper, std::uint64_t count,
                               std::uint64_t dim, std::uint64_t leaf_size, double eta);

class HostStore {
public:
    HostStore(std::uint64_t order, std::vector<std::uint64_t> permutation);

    void add_dense(const Rect& r, const double* values, std::uint64_t count);
    // u: m x rank and v: n x rank, column-major (component j at j*m / j*n).
    void add_low_rank(const Rect& r, const double* u, const double* sigma, std::uint64_t rank,
                      const double* v, const int* widths, std::uint64_t width_count,
                      std::uint64_t menu_size, double truncation_error);
    void finalize();

    // y = H x on the host; x and y are order x rhs row-major in source order.
    void matmul(const double* x, std::uint64_t rhs, double* y, int threads) const;
    // Dense FP64 matrix in source order (order*order values).
    void reconstruct(double* out) const;

    std::vector<BlockSummary> summaries() const;
    BlockSummary summary(std::size_t index) const;
    std::uint64_t logical_bits() const;
    std::uint64_t allocated_bytes() const;
    std::uint64_t resident_bytes() const;
    double reconstruction_bound() const;

    void save(const std::string& path) const;
    static HostStore load(const std::string& path);

    std::uint64_t order() const { return order_; }
    bool finalized() const { return finalized_; }
    const std::vector<std::uint64_t>& permutation() const { return perm_; }
    const std::vector<HostBlock>& blocks() const { return blocks_; }

private:
    void check_rect(const Rect& r, bool diagonal) const;

    std::uint64_t order_;
    std::vector<std::uint64_t> perm_;
    std::vector<HostBlock> blocks_;
    bool finalized_ = false;
};

// Packs one balanced factor (rows x widths.size(), column-major) into `side`.
void pack_side(const std::vector<double>& factor, std::uint64_t rows,
               const std::vector<int>& widths, PackedSide& side);

}  // namespace qhamd
