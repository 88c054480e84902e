// Host packed store, construction helpers, QHMAT001 persistence and the
// explicit host backend of the apply (device="cpu").  Semantics and error
// messages follow the reference (src/qhmat/core.cpp, io.cpp) so that callers
// see the same exceptions; the code is organised around the device upload.
#include "host_store.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace qhamd {
namespace {

void require(bool ok, const char* message) {
    if (!ok) throw std::invalid_argument(message);
}

std::uint64_t add_checked(std::uint64_t a, std::uint64_t b) {
    if (b > std::numeric_limits<std::uint64_t>::max() - a)
        throw std::overflow_error("storage count overflow");
    return a + b;
}

std::uint64_t mul_checked(std::uint64_t a, std::uint64_t b) {
    if (b != 0 && a > std::numeric_limits<std::uint64_t>::max() / b)
        throw std::overflow_error("storage count overflow");
    return a * b;
}

// Bits needed to label one entry of a precision menu (core.cpp:37-42).
std::uint64_t menu_label_bits(std::uint64_t menu_size) {
    require(menu_size >= 1 && menu_size <= 31, "precision menu size must be in [1, 31]");
    std::uint64_t bits = 0;
    while ((std::uint64_t{1} << bits) < menu_size) ++bits;
    return bits;
}

// Sign-extended b-bit code at index i of one component stream.
std::int64_t read_code(const std::uint32_t* w, std::uint64_t i, int bits) {
    const std::uint64_t bit = i * std::uint64_t(bits);
    std::uint64_t pair = w[bit >> 5];
    if ((bit & 31) + std::uint64_t(bits) > 32) pair |= std::uint64_t(w[(bit >> 5) + 1]) << 32;
    const std::uint64_t raw = (pair >> (bit & 31)) & ((std::uint64_t{1} << bits) - 1);
    const std::uint64_t sign = std::uint64_t{1} << (bits - 1);
    return static_cast<std::int64_t>(raw ^ sign) - static_cast<std::int64_t>(sign);
}

}  // namespace

std::uint64_t code_limit(int bits) {
    require(bits >= 2 && bits <= 32, "bit width must be in [2, 32]");
    return (std::uint64_t{1} << (bits - 1)) - 1;
}

double coefficient(int bits) {
    const double q = static_cast<double>(code_limit(bits));
    return 1.0 / q + 1.0 / (4.0 * q * q);
}

std::uint64_t words_for(std::uint64_t rows, int bits) {
    return add_checked(mul_checked(rows, std::uint64_t(bits)), 31) / 32;
}

// ---------------------------------------------------------------------------
// Width allocation.  The greedy walks the menu from the narrowest width and,
// for each level, lowers the weakest still-unassigned components while the
// total surrogate (cost so far + this component at the level + everything
// stronger at the widest width) stays within what is left of the budget.
// Sums are carried in long double in the same association as the reference
// so that feasibility decisions at the budget edge agree.
Allocation allocate_widths(const std::vector<double>& sigma, const std::vector<int>& menu_in,
                           double budget, std::uint64_t rows, std::uint64_t cols) {
    require(rows > 0 && cols > 0, "matrix dimensions must be positive");
    require(std::isfinite(budget) && budget >= 0.0, "budget must be finite and nonnegative");
    require(sigma.size() <= std::min(rows, cols), "rank exceeds block dimension");
    std::vector<int> menu = menu_in;
    require(!menu.empty(), "precision menu must not be empty");
    std::sort(menu.begin(), menu.end());
    for (std::size_t i = 1; i < menu.size(); ++i)
        require(menu[i] != menu[i - 1], "precision menu contains duplicates");
    for (int b : menu) code_limit(b);
    for (std::size_t j = 0; j < sigma.size(); ++j) {
        require(std::isfinite(sigma[j]) && sigma[j] >= 0.0,
                "singular values must be finite and nonnegative");
        require(j == 0 || sigma[j - 1] >= sigma[j], "singular values must be sorted descending");
    }
    const std::size_t r = sigma.size();
    std::vector<long double> head(r + 1, 0.0L);  // head[t] = sigma_0 + ... + sigma_{t-1}
    for (std::size_t j = 0; j < r; ++j) head[j + 1] = head[j] + sigma[j];

    Allocation out;
    out.bits.assign(r, menu.back());
    const long double widest = coefficient(menu.back());
    if (widest * head[r] <= static_cast<long double>(budget)) {
        long double left = budget;
        std::size_t open = r;  // components [0, open) still at the widest width
        for (std::size_t level = 0; level + 1 < menu.size() && open > 0; ++level) {
            const long double c = coefficient(menu[level]);
            long double spent = 0.0L;
            for (; open > 0; --open) {
                const long double trial = spent + c * sigma[open - 1] + widest * head[open - 1];
                if (trial > left) break;
                out.bits[open - 1] = menu[level];
                spent += c * sigma[open - 1];
            }
            left -= spent;
        }
    }
    long double bound = 0.0L;
    std::uint64_t width_sum = 0;
    for (std::size_t j = 0; j < r; ++j) {
        bound += static_cast<long double>(sigma[j]) * coefficient(out.bits[j]);
        width_sum = add_checked(width_sum, std::uint64_t(out.bits[j]));
    }
    out.bound = static_cast<double>(bound);
    out.payload_bits = mul_checked(add_checked(rows, cols), width_sum);
    out.feasible = bound <= static_cast<long double>(budget);
    return out;
}

// ---------------------------------------------------------------------------
// Cluster tree and block partition.
PartitionOut spatial_partition(const std::vector<double>& pts, const std::vector<double>& lo,
                               const std::vector<double>& hi, std::uint64_t count,
                               std::uint64_t dim, std::uint64_t leaf_size, double eta) {
    require(count > 0 && dim > 0 && dim <= 16, "invalid point-cloud dimensions");
    require(count <= std::numeric_limits<std::uint64_t>::max() / dim,
            "point-cloud dimensions overflow");
    require(pts.size() == count * dim && lo.size() == pts.size() && hi.size() == pts.size(),
            "point/support shapes differ");
    require(leaf_size > 0 && std::isfinite(eta) && eta > 0,
            "leaf size and admissibility must be positive");
    for (std::size_t i = 0; i < pts.size(); ++i)
        require(std::isfinite(pts[i]) && std::isfinite(lo[i]) && std::isfinite(hi[i]) &&
                    lo[i] <= pts[i] && pts[i] <= hi[i],
                "invalid point or support box");

    struct Node {
        std::uint64_t begin = 0, end = 0;
        std::vector<double> box_lo, box_hi;
        int child[2] = {-1, -1};
        bool leaf() const { return child[0] < 0; }
        double diameter() const {
            double s = 0.0;
            for (std::size_t d = 0; d < box_lo.size(); ++d) {
                const double g = box_hi[d] - box_lo[d];
                s += g * g;
            }
            return std::sqrt(s);
        }
    };
    PartitionOut out;
    out.permutation.reserve(count);
    std::vector<Node> nodes;

    // Builds the subtree over `idx` (source indices) and returns its node id.
    std::function<int(std::vector<std::uint64_t>&)> grow = [&](std::vector<std::uint64_t>& idx) {
        Node node;
        node.begin = out.permutation.size();
        const double inf = std::numeric_limits<double>::infinity();
        node.box_lo.assign(dim, inf);
        node.box_hi.assign(dim, -inf);
        std::vector<double> pmin(dim, inf), pmax(dim, -inf);
        for (std::uint64_t i : idx)
            for (std::uint64_t d = 0; d < dim; ++d) {
                const std::uint64_t at = i * dim + d;
                node.box_lo[d] = std::min(node.box_lo[d], lo[at]);
                node.box_hi[d] = std::max(node.box_hi[d], hi[at]);
                pmin[d] = std::min(pmin[d], pts[at]);
                pmax[d] = std::max(pmax[d], pts[at]);
            }
        const int id = static_cast<int>(nodes.size());
        nodes.push_back(std::move(node));
        if (idx.size() <= leaf_size) {
            out.permutation.insert(out.permutation.end(), idx.begin(), idx.end());
        } else {
            std::uint64_t axis = 0;  // widest point spread, first axis on ties
            for (std::uint64_t d = 1; d < dim; ++d)
                if (pmax[d] - pmin[d] > pmax[axis] - pmin[axis]) axis = d;
            std::stable_sort(idx.begin(), idx.end(), [&](std::uint64_t a, std::uint64_t b) {
                return pts[a * dim + axis] < pts[b * dim + axis];
            });
            const std::size_t half = idx.size() / 2;
            std::vector<std::uint64_t> first(idx.begin(), idx.begin() + half);
            std::vector<std::uint64_t> second(idx.begin() + half, idx.end());
            idx.clear();
            idx.shrink_to_fit();
            const int a = grow(first);
            const int b = grow(second);
            nodes[id].child[0] = a;
            nodes[id].child[1] = b;
        }
        nodes[id].end = out.permutation.size();
        return id;
    };
    std::vector<std::uint64_t> all(count);
    std::iota(all.begin(), all.end(), std::uint64_t{0});
    const int root = grow(all);

    std::function<void(int, int)> pair = [&](int ia, int ib) {
        const Node& a = nodes[ia];
        const Node& b = nodes[ib];
        double gap2 = 0.0;
        for (std::uint64_t d = 0; d < dim; ++d) {
            const double g = std::max({0.0, a.box_lo[d] - b.box_hi[d], b.box_lo[d] - a.box_hi[d]});
            gap2 += g * g;
        }
        const double gap = std::sqrt(gap2);
        const bool far = gap > 0.0 && gap >= eta * std::max(a.diameter(), b.diameter());
        if (far || (a.leaf() && b.leaf())) {
            out.blocks.push_back(Rect{a.begin, a.end, b.begin, b.end, far});
            return;
        }
        // A leaf stands for itself against the other side's children.
        const int ra[2] = {a.leaf() ? ia : a.child[0], a.leaf() ? -1 : a.child[1]};
        const int rb[2] = {b.leaf() ? ib : b.child[0], b.leaf() ? -1 : b.child[1]};
        for (int x : ra)
            for (int y : rb)
                if (x >= 0 && y >= 0) pair(x, y);
    };
    pair(root, root);
    return out;
}

// ---------------------------------------------------------------------------
// Packing: per component, scale = max|x| / q; code = nearbyint(x / scale)
// clamped to [-q, q]; codes are appended LSB-first through a 64-bit bit
// buffer so each component ends on a word boundary.
void pack_side(const std::vector<double>& factor, std::uint64_t rows,
               const std::vector<int>& widths, PackedSide& side) {
    side.words.clear();
    side.offsets.clear();
    side.scales.clear();
    std::uint64_t total = 0;
    for (int b : widths) total = add_checked(total, words_for(rows, b));
    side.words.reserve(total);
    for (std::size_t j = 0; j < widths.size(); ++j) {
        const int b = widths[j];
        const std::int64_t q = static_cast<std::int64_t>(code_limit(b));
        const double* col = factor.data() + j * rows;
        double peak = 0.0;
        for (std::uint64_t i = 0; i < rows; ++i) peak = std::max(peak, std::fabs(col[i]));
        const double scale = peak == 0 ? 0 : peak / static_cast<double>(q);
        side.scales.push_back(scale);
        side.offsets.push_back(side.words.size());
        const std::uint64_t mask = (std::uint64_t{1} << b) - 1;
        std::uint64_t buffer = 0;
        int filled = 0;
        for (std::uint64_t i = 0; i < rows; ++i) {
            std::int64_t code = 0;
            if (scale != 0) code = static_cast<std::int64_t>(std::nearbyint(col[i] / scale));
            code = std::min(q, std::max(-q, code));
            buffer |= (static_cast<std::uint64_t>(code) & mask) << filled;
            filled += b;
            if (filled >= 32) {
                side.words.push_back(static_cast<std::uint32_t>(buffer));
                buffer >>= 32;
                filled -= 32;
            }
        }
        if (filled > 0) side.words.push_back(static_cast<std::uint32_t>(buffer));
    }
}

// ---------------------------------------------------------------------------
HostStore::HostStore(std::uint64_t order, std::vector<std::uint64_t> permutation)
    : order_(order), perm_(std::move(permutation)) {
    require(order_ > 0 && perm_.size() == order_, "invalid matrix order/permutation");
    std::vector<char> seen(order_, 0);
    for (std::uint64_t s : perm_) {
        require(s < order_ && !seen[s], "permutation is not a bijection");
        seen[s] = 1;
    }
}

void HostStore::check_rect(const Rect& r, bool diagonal) const {
    require(!finalized_, "matrix has been finalized");
    require(r.row0 < r.row1 && r.row1 <= order_ && r.col0 < r.col1 && r.col1 <= order_,
            "invalid block rectangle");
    const bool same = r.row0 == r.col0 && r.row1 == r.col1;
    const bool apart = r.row1 <= r.col0 || r.col1 <= r.row0;
    require(diagonal ? same : apart,
            "only diagonal blocks may be dense; factorized blocks must be off-diagonal");
}

void HostStore::add_dense(const Rect& r, const double* values, std::uint64_t count) {
    check_rect(r, true);
    const std::uint64_t m = r.row1 - r.row0, n = r.col1 - r.col0;
    require(count == mul_checked(m, n), "dense block shape mismatch");
    for (std::uint64_t i = 0; i < count; ++i)
        require(std::isfinite(values[i]), "nonfinite dense block");
    HostBlock b;
    b.rect = r;
    b.dense.assign(values, values + count);
    blocks_.push_back(std::move(b));
}

void HostStore::add_low_rank(const Rect& r, const double* u, const double* sigma,
                             std::uint64_t rank, const double* v, const int* widths,
                             std::uint64_t width_count, std::uint64_t menu_size,
                             double truncation_error) {
    check_rect(r, false);
    const std::uint64_t m = r.row1 - r.row0, n = r.col1 - r.col0;
    require(rank <= std::min(m, n) && width_count == rank,
            "factor dimensions do not match block rank");
    require(std::isfinite(truncation_error) && truncation_error >= 0.0,
            "invalid truncation error");
    menu_label_bits(menu_size);
    HostBlock b;
    b.rect = r;
    b.widths.assign(widths, widths + rank);
    b.menu_size = menu_size;
    b.truncation_error = truncation_error;
    std::vector<double> left(m * rank), right(n * rank);
    for (std::uint64_t j = 0; j < rank; ++j) {
        require(std::isfinite(sigma[j]) && sigma[j] >= 0.0,
                "singular values must be finite and nonnegative");
        require(j == 0 || sigma[j - 1] >= sigma[j], "singular values must be sorted descending");
        code_limit(widths[j]);
        const double root = std::sqrt(sigma[j]);
        double un = 0.0, vn = 0.0;
        for (std::uint64_t i = 0; i < m; ++i) {
            const double e = u[j * m + i];
            require(std::isfinite(e), "nonfinite left factor");
            un += e * e;
            left[j * m + i] = root * e;
        }
        for (std::uint64_t i = 0; i < n; ++i) {
            const double e = v[j * n + i];
            require(std::isfinite(e), "nonfinite right factor");
            vn += e * e;
            right[j * n + i] = root * e;
        }
        require(un <= 1.0 + 1e-8 && vn <= 1.0 + 1e-8,
                "singular vectors must have norm at most one");
        b.quantization_bound += sigma[j] * coefficient(widths[j]);
    }
    pack_side(left, m, b.widths, b.left);
    pack_side(right, n, b.widths, b.right);
    blocks_.push_back(std::move(b));
}

void HostStore::finalize() {
    require(!finalized_, "matrix has already been finalized");
    require(!blocks_.empty(), "matrix has no blocks");
    // Row slabs: the intervals between consecutive distinct row boundaries.
    std::vector<std::uint64_t> cuts{0, order_};
    for (const auto& b : blocks_) {
        cuts.push_back(b.rect.row0);
        cuts.push_back(b.rect.row1);
    }
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    const std::size_t slabs = cuts.size() - 1;
    std::vector<std::vector<std::pair<std::uint64_t, std::uint64_t>>> spans(slabs);
    for (const auto& b : blocks_) {
        const std::size_t s0 = std::lower_bound(cuts.begin(), cuts.end(), b.rect.row0) - cuts.begin();
        for (std::size_t s = s0; s < slabs && cuts[s] < b.rect.row1; ++s)
            spans[s].emplace_back(b.rect.col0, b.rect.col1);
    }
    // Every slab's column intervals must tile [0, order) exactly.
    for (auto& list : spans) {
        std::sort(list.begin(), list.end());
        std::uint64_t reach = 0;
        for (const auto& iv : list) {
            require(iv.first == reach, "block partition has a gap or overlap");
            reach = iv.second;
        }
        require(reach == order_, "block partition does not cover the matrix");
    }
    finalized_ = true;
}

// Host backend.  Each low-rank block's right factor is reduced against x once
// (all right-hand sides together), then each row slab is owned by one thread,
// which adds its blocks in insertion order: no shared writes, fixed order.
// Every sum and product is the reference's (core.cpp:360-390: decoded value
// = code x scale, dot = sum over k in order, y += left value x dot per
// component in order, blocks in insertion order), so the result is bitwise
// the reference's for any thread count.
void HostStore::matmul(const double* x, std::uint64_t rhs, double* y, int threads) const {
    require(finalized_, "matrix must be finalized before use");
    require(rhs > 0 && rhs <= std::numeric_limits<std::uint64_t>::max() / order_,
            "right-hand-side shape mismatch");
    for (std::uint64_t i = 0; i < order_ * rhs; ++i)
        require(std::isfinite(x[i]), "nonfinite right-hand side");
    std::fill(y, y + order_ * rhs, 0.0);
    const unsigned workers =
        threads > 0 ? unsigned(threads) : std::max(1u, std::thread::hardware_concurrency());
    auto run = [&](std::size_t count, const std::function<void(std::size_t)>& body) {
        std::atomic<std::size_t> next{0};
        auto loop = [&] {
            for (std::size_t t; (t = next.fetch_add(1)) < count;) body(t);
        };
        std::vector<std::thread> pool;
        for (unsigned w = 1; w < std::min<std::size_t>(workers, count); ++w) pool.emplace_back(loop);
        loop();
        for (auto& t : pool) t.join();
    };
    // z[b][j*rhs + p] = sum_k (codeR(k, j) sR_j) x[perm[c0 + k], p], k in order
    std::vector<std::vector<double>> z(blocks_.size());
    run(blocks_.size(), [&](std::size_t bi) {
        const auto& b = blocks_[bi];
        if (b.is_dense()) return;
        const std::size_t r = b.widths.size();
        z[bi].assign(r * rhs, 0.0);
        for (std::size_t j = 0; j < r; ++j) {
            const std::uint32_t* w = b.right.words.data() + b.right.offsets[j];
            double* zj = z[bi].data() + j * rhs;
            for (std::uint64_t k = 0; k < b.cols(); ++k) {
                const double c = double(read_code(w, k, b.widths[j])) * b.right.scales[j];
                const double* xr = x + perm_[b.rect.col0 + k] * rhs;
                for (std::uint64_t p = 0; p < rhs; ++p) zj[p] += c * xr[p];
            }
        }
    });
    std::vector<std::uint64_t> cuts{0, order_};
    for (const auto& b : blocks_) {
        cuts.push_back(b.rect.row0);
        cuts.push_back(b.rect.row1);
    }
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    std::vector<std::vector<std::size_t>> owned(cuts.size() - 1);
    for (std::size_t bi = 0; bi < blocks_.size(); ++bi) {
        const auto& r = blocks_[bi].rect;
        std::size_t s = std::lower_bound(cuts.begin(), cuts.end(), r.row0) - cuts.begin();
        for (; s + 1 < cuts.size() && cuts[s] < r.row1; ++s) owned[s].push_back(bi);
    }
    run(owned.size(), [&](std::size_t s) {
        for (std::size_t bi : owned[s]) {
            const auto& b = blocks_[bi];
            for (std::uint64_t row = cuts[s]; row < cuts[s + 1]; ++row) {
                const std::uint64_t i = row - b.rect.row0;
                double* yr = y + perm_[row] * rhs;
                if (b.is_dense()) {
                    const double* d = b.dense.data() + i * b.cols();
                    for (std::uint64_t k = 0; k < b.cols(); ++k) {
                        const double* xr = x + perm_[b.rect.col0 + k] * rhs;
                        for (std::uint64_t p = 0; p < rhs; ++p) yr[p] += d[k] * xr[p];
                    }
                    continue;
                }
                for (std::size_t j = 0; j < b.widths.size(); ++j) {
                    const double l = double(read_code(b.left.words.data() + b.left.offsets[j], i,
                                                      b.widths[j])) * b.left.scales[j];
                    const double* zj = z[bi].data() + j * rhs;
                    for (std::uint64_t p = 0; p < rhs; ++p) yr[p] += l * zj[p];
                }
            }
        }
    });
}

void HostStore::reconstruct(double* out) const {
    require(finalized_, "matrix must be finalized before use");
    require(order_ <= std::numeric_limits<std::uint64_t>::max() / order_,
            "dense output size overflow");
    for (const auto& b : blocks_) {
        const std::uint64_t m = b.rows(), n = b.cols();
        // Decode both factors once, then form each entry.
        const std::size_t r = b.widths.size();
        std::vector<double> L(m * r), R(n * r);
        for (std::size_t j = 0; j < r; ++j) {
            for (std::uint64_t i = 0; i < m; ++i)
                L[i * r + j] = double(read_code(b.left.words.data() + b.left.offsets[j], i,
                                                b.widths[j])) * b.left.scales[j];
            for (std::uint64_t k = 0; k < n; ++k)
                R[k * r + j] = double(read_code(b.right.words.data() + b.right.offsets[j], k,
                                                b.widths[j])) * b.right.scales[j];
        }
        for (std::uint64_t i = 0; i < m; ++i)
            for (std::uint64_t k = 0; k < n; ++k) {
                double v = b.is_dense() ? b.dense[i * n + k] : 0.0;
                for (std::size_t j = 0; j < r; ++j) v += L[i * r + j] * R[k * r + j];
                out[perm_[b.rect.row0 + i] * order_ + perm_[b.rect.col0 + k]] = v;
            }
    }
}

std::vector<BlockSummary> HostStore::summaries() const {
    std::vector<BlockSummary> out;
    out.reserve(blocks_.size());
    for (std::size_t i = 0; i < blocks_.size(); ++i) out.push_back(summary(i));
    return out;
}

BlockSummary HostStore::summary(std::size_t index) const {
    const HostBlock& b = blocks_.at(index);
    {
        BlockSummary s{b.rect, b.is_dense(), b.widths.size(), b.truncation_error,
                       b.quantization_bound, 0, 0};
        if (b.is_dense()) {
            s.logical_bits = mul_checked(b.dense.size(), 64);
            s.allocated_bytes = mul_checked(b.dense.size(), sizeof(double));
        } else {
            // payload (m+n)*sum(b) + per component two FP64 scales and a menu label
            std::uint64_t width_sum = 0;
            for (int w : b.widths) width_sum = add_checked(width_sum, std::uint64_t(w));
            s.logical_bits = add_checked(
                mul_checked(add_checked(b.rows(), b.cols()), width_sum),
                mul_checked(b.widths.size(), 128 + menu_label_bits(b.menu_size)));
            std::uint64_t bytes =
                mul_checked(add_checked(b.left.words.size(), b.right.words.size()), 4);
            bytes = add_checked(bytes, mul_checked(add_checked(b.left.scales.size(),
                                                               b.right.scales.size()), 8));
            bytes = add_checked(bytes, mul_checked(add_checked(b.left.offsets.size(),
                                                               b.right.offsets.size()), 8));
            bytes = add_checked(bytes, mul_checked(b.widths.size(), sizeof(int)));
            s.allocated_bytes = bytes;
        }
        return s;
    }
}

std::uint64_t HostStore::logical_bits() const {
    std::uint64_t t = 0;
    for (const auto& s : summaries()) t = add_checked(t, s.logical_bits);
    return t;
}

std::uint64_t HostStore::allocated_bytes() const {
    std::uint64_t t = 0;
    for (const auto& s : summaries()) t = add_checked(t, s.allocated_bytes);
    return t;
}

std::uint64_t HostStore::resident_bytes() const {
    std::uint64_t t = sizeof(*this);
    t = add_checked(t, mul_checked(perm_.capacity(), sizeof(std::uint64_t)));
    t = add_checked(t, mul_checked(blocks_.capacity(), sizeof(HostBlock)));
    for (const auto& b : blocks_) {
        t = add_checked(t, mul_checked(b.dense.capacity(), sizeof(double)));
        t = add_checked(t, mul_checked(b.widths.capacity(), sizeof(int)));
        for (const PackedSide* s : {&b.left, &b.right}) {
            t = add_checked(t, mul_checked(s->words.capacity(), sizeof(std::uint32_t)));
            t = add_checked(t, mul_checked(s->offsets.capacity(), sizeof(std::uint64_t)));
            t = add_checked(t, mul_checked(s->scales.capacity(), sizeof(double)));
        }
    }
    return t;
}

double HostStore::reconstruction_bound() const {
    double worst = 0.0;
    for (const auto& b : blocks_)
        worst = std::max(worst, b.truncation_error + b.quantization_bound);
    return worst;
}

// ---------------------------------------------------------------------------
// QHMAT001 files (reference io.cpp:127-257): little-endian fields, an FNV-1a
// 64-bit checksum of every preceding byte as trailer, written to a temporary
// name and renamed so that a reader never sees a partial file.
namespace {
constexpr std::uint64_t kFnvBasis = 14695981039346656037ULL;
constexpr std::uint64_t kFnvPrime = 1099511628211ULL;

std::uint64_t fnv1a(const unsigned char* p, std::size_t n) {
    std::uint64_t h = kFnvBasis;
    for (std::size_t i = 0; i < n; ++i) h = (h ^ p[i]) * kFnvPrime;
    return h;
}

struct ByteSink {
    std::vector<unsigned char> bytes;
    void u8(unsigned v) { bytes.push_back(static_cast<unsigned char>(v)); }
    void u32(std::uint32_t v) {
        for (int i = 0; i < 4; ++i) u8((v >> (8 * i)) & 0xFF);
    }
    void u64(std::uint64_t v) {
        for (int i = 0; i < 8; ++i) u8((v >> (8 * i)) & 0xFF);
    }
    void f64(double v) {
        std::uint64_t b;
        std::memcpy(&b, &v, 8);
        u64(b);
    }
};

struct ByteSource {
    const unsigned char* p;
    std::size_t size;  // payload bytes (the 8-byte trailer excluded)
    std::size_t at = 0;
    void need(std::size_t n) const {
        if (n > size - at) throw std::runtime_error("truncated qhmat file");
    }
    unsigned u8() {
        need(1);
        return p[at++];
    }
    std::uint32_t u32() {
        need(4);
        std::uint32_t v = 0;
        for (int i = 0; i < 4; ++i) v |= std::uint32_t(p[at++]) << (8 * i);
        return v;
    }
    std::uint64_t u64() {
        need(8);
        std::uint64_t v = 0;
        for (int i = 0; i < 8; ++i) v |= std::uint64_t(p[at++]) << (8 * i);
        return v;
    }
    double f64() {
        const std::uint64_t b = u64();
        double v;
        std::memcpy(&v, &b, 8);
        return v;
    }
    void need_items(std::uint64_t count, std::uint64_t width) const {
        if (width && count > (size - at) / width) throw std::runtime_error("truncated qhmat payload");
    }
};
}  // namespace

void HostStore::save(const std::string& path) const {
    if (!finalized_) throw std::invalid_argument("matrix must be finalized before saving");
    const auto target = std::filesystem::u8path(path);
    if (std::filesystem::exists(target))
        throw std::invalid_argument("refusing to replace an existing qhmat file");
    ByteSink out;
    for (char c : std::string("QHMAT001")) out.u8(static_cast<unsigned char>(c));
    out.u64(1);
    out.u64(order_);
    out.u64(blocks_.size());
    for (std::uint64_t s : perm_) out.u64(s);
    for (const auto& b : blocks_) {
        out.u64(b.rect.row0);
        out.u64(b.rect.row1);
        out.u64(b.rect.col0);
        out.u64(b.rect.col1);
        out.u8(b.rect.admissible ? 1 : 0);
        out.u8(b.is_dense() ? 1 : 0);
        if (b.is_dense()) {
            for (double v : b.dense) out.f64(v);
            continue;
        }
        out.u64(b.widths.size());
        out.u64(b.menu_size);
        out.f64(b.truncation_error);
        out.f64(b.quantization_bound);
        for (int w : b.widths) out.u8(static_cast<unsigned>(w));
        for (double s : b.left.scales) out.f64(s);
        for (double s : b.right.scales) out.f64(s);
        for (std::uint32_t w : b.left.words) out.u32(w);
        for (std::uint32_t w : b.right.words) out.u32(w);
    }
    out.u64(fnv1a(out.bytes.data(), out.bytes.size()));
    static std::atomic<std::uint64_t> serial{0};
    auto temp = target;
    temp += ".partial-" + std::to_string(std::chrono::steady_clock::now().time_since_epoch().count()) +
            "-" + std::to_string(serial++);
    try {
        {
            std::ofstream f(temp, std::ios::binary);
            if (!f) throw std::runtime_error("cannot create qhmat file");
            f.write(reinterpret_cast<const char*>(out.bytes.data()),
                    static_cast<std::streamsize>(out.bytes.size()));
            f.flush();
            if (!f) throw std::runtime_error("qhmat write failed");
        }
        std::filesystem::rename(temp, target);
    } catch (...) {
        std::error_code ignored;
        std::filesystem::remove(temp, ignored);
        throw;
    }
}

HostStore HostStore::load(const std::string& path) {
    std::ifstream f(std::filesystem::u8path(path), std::ios::binary);
    if (!f) throw std::runtime_error("cannot open qhmat file");
    std::vector<unsigned char> data((std::istreambuf_iterator<char>(f)),
                                    std::istreambuf_iterator<char>());
    if (data.size() < 24) throw std::runtime_error("truncated qhmat file");
    ByteSource in{data.data(), data.size() - 8};
    std::string magic;
    for (int i = 0; i < 8; ++i) magic.push_back(static_cast<char>(in.u8()));
    if (magic != "QHMAT001") throw std::runtime_error("invalid qhmat file magic");
    if (in.u64() != 1) throw std::runtime_error("unsupported qhmat file version");
    const std::uint64_t order = in.u64();
    const std::uint64_t count = in.u64();
    if (!order || !count || order > (in.size - in.at) / 8)
        throw std::runtime_error("invalid qhmat file dimensions");
    std::vector<std::uint64_t> perm(order);
    for (auto& s : perm) s = in.u64();
    HostStore store(order, std::move(perm));
    if (count > (in.size - in.at) / 34) throw std::runtime_error("invalid qhmat block count");
    store.blocks_.reserve(count);
    for (std::uint64_t k = 0; k < count; ++k) {
        Rect r;
        r.row0 = in.u64();
        r.row1 = in.u64();
        r.col0 = in.u64();
        r.col1 = in.u64();
        const unsigned adm = in.u8(), dense = in.u8();
        if (adm > 1 || dense > 1) throw std::runtime_error("invalid qhmat block flag");
        r.admissible = adm == 1;
        store.check_rect(r, dense == 1);
        HostBlock b;
        b.rect = r;
        const std::uint64_t rows = r.row1 - r.row0, cols = r.col1 - r.col0;
        if (dense) {
            if (cols > std::numeric_limits<std::uint64_t>::max() / rows)
                throw std::runtime_error("dense block size overflow");
            in.need_items(rows * cols, 8);
            b.dense.resize(rows * cols);
            for (auto& v : b.dense) {
                v = in.f64();
                if (!std::isfinite(v)) throw std::runtime_error("nonfinite dense value");
            }
        } else {
            const std::uint64_t rank = in.u64();
            b.menu_size = in.u64();
            if (rank > std::min(rows, cols) || b.menu_size == 0 || b.menu_size > 31)
                throw std::runtime_error("invalid qhmat rank or precision menu");
            b.truncation_error = in.f64();
            b.quantization_bound = in.f64();
            if (!std::isfinite(b.truncation_error) || b.truncation_error < 0 ||
                !std::isfinite(b.quantization_bound) || b.quantization_bound < 0)
                throw std::runtime_error("invalid qhmat error bound");
            in.need_items(rank, 17);
            b.widths.resize(rank);
            for (auto& w : b.widths) {
                w = static_cast<int>(in.u8());
                if (w < 2 || w > 32) throw std::runtime_error("invalid qhmat bit width");
            }
            b.left.scales.resize(rank);
            b.right.scales.resize(rank);
            for (auto& s : b.left.scales) s = in.f64();
            for (auto& s : b.right.scales) s = in.f64();
            for (const auto* sc : {&b.left.scales, &b.right.scales})
                for (double s : *sc)
                    if (!std::isfinite(s) || s < 0) throw std::runtime_error("invalid qhmat scale");
            std::uint64_t lw = 0, rw = 0;
            for (int w : b.widths) {
                b.left.offsets.push_back(lw);
                b.right.offsets.push_back(rw);
                const std::uint64_t a = words_for(rows, w), c = words_for(cols, w);
                if (a > std::numeric_limits<std::uint64_t>::max() - lw ||
                    c > std::numeric_limits<std::uint64_t>::max() - rw)
                    throw std::runtime_error("packed block size overflow");
                lw += a;
                rw += c;
            }
            if (rw > std::numeric_limits<std::uint64_t>::max() - lw)
                throw std::runtime_error("packed block size overflow");
            in.need_items(lw + rw, 4);
            b.left.words.resize(lw);
            b.right.words.resize(rw);
            for (auto& w : b.left.words) w = in.u32();
            for (auto& w : b.right.words) w = in.u32();
        }
        store.blocks_.push_back(std::move(b));
    }
    if (in.at != in.size) throw std::runtime_error("trailing qhmat data");
    std::uint64_t stored = 0;
    for (int i = 0; i < 8; ++i) stored |= std::uint64_t(data[in.size + i]) << (8 * i);
    if (stored != fnv1a(data.data(), in.size)) throw std::runtime_error("qhmat checksum mismatch");
    store.finalize();
    return store;
}

}  // namespace qhamd
