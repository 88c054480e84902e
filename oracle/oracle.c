/*
 * CPU oracle for the qhmat H-matrix apply -- TEST INFRASTRUCTURE ONLY.
 * Restates the reference FP64 CPU path; see oracle.h for the function list,
 * the reference lines each one follows and what pins it.
 */
#include "oracle.h"

#include <math.h>
#include <string.h>

/* core.cpp:32-35 */
uint64_t orc_limit(int bits) { return ((uint64_t)1 << (bits - 1)) - 1; }

/* core.cpp:56-59 */
double orc_coefficient(int bits) {
    double q = (double)orc_limit(bits);
    return 1.0 / q + 1.0 / (4.0 * q * q);
}

/* core.cpp:249-250: ceil(rows * bits / 32) words per component. */
size_t orc_words_for(size_t rows, int bits) { return (rows * (size_t)bits + 31) / 32; }

/* core.cpp:241-264 for one component j.  The scale is max|x| / q (zero column
 * -> scale 0 and all-zero codes); codes are nearbyint(x / scale) clamped to
 * [-q, q] and stored as b-bit two's complement, LSB first, code i at bit i*b. */
void orc_pack_component(const double *column, size_t rows, int bits,
                        uint32_t *words, double *scale) {
    const int64_t q = (int64_t)orc_limit(bits);
    double peak = 0.0;
    for (size_t i = 0; i < rows; ++i) {
        double a = fabs(column[i]);
        if (a > peak) peak = a;
    }
    const double s = peak == 0 ? 0 : peak / (double)q;
    *scale = s;
    const uint32_t mask = bits == 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
    for (size_t i = 0; i < rows; ++i) {
        double v = s == 0 ? 0 : column[i] / s;
        int64_t code = (int64_t)nearbyint(v);
        if (code > q) code = q;
        if (code < -q) code = -q;
        uint32_t raw = (uint32_t)code & mask;
        size_t bit = i * (size_t)bits;
        size_t w = bit / 32;
        unsigned sh = (unsigned)(bit % 32);
        words[w] |= raw << sh;
        if (sh + (unsigned)bits > 32) words[w + 1] |= raw >> (32 - sh);
    }
}

/* core.cpp:271-280: the b bits at i*b, sign-extended from bit b-1. */
int64_t orc_decode_code(const uint32_t *cw, size_t row, int bits) {
    size_t bit = row * (size_t)bits;
    size_t w = bit / 32;
    unsigned sh = (unsigned)(bit % 32);
    uint32_t raw = cw[w] >> sh;
    if (sh + (unsigned)bits > 32) raw |= cw[w + 1] << (32 - sh);
    if (bits < 32) raw &= (1u << bits) - 1u;
    if (raw & (1u << (bits - 1))) return (int64_t)raw - ((int64_t)1 << bits);
    return (int64_t)raw;
}

/* All codes of one packed factor (components back to back, each starting on
 * the word after the previous one, core.cpp:248-251): out[j*rows + i]. */
void orc_decode_factor(const uint32_t *words, size_t rows, const int *widths, size_t rank,
                       int64_t *out) {
    size_t off = 0;
    for (size_t j = 0; j < rank; ++j) {
        for (size_t i = 0; i < rows; ++i)
            out[j * rows + i] = orc_decode_code(words + off, i, widths[j]);
        off += orc_words_for(rows, widths[j]);
    }
}

/* core.cpp:281 */
double orc_decode(const uint32_t *cw, double scale, size_t row, int bits) {
    return (double)orc_decode_code(cw, row, bits) * scale;
}

/* core.cpp:360-390.  Same loop nest and the same operation order, so the
 * FP64 result is reproducible bit for bit against the reference core built
 * with the same compiler flags:
 *   dense:    for i, for k, for p: y[perm[r0+i], p] += D[i,k] * x[perm[c0+k], p]
 *   low-rank: for p, for j: dot = sum_k R(k,j) x[perm[c0+k], p];
 *                           for i: y[perm[r0+i], p] += L(i,j) * dot          */
int orc_matmul(const orc_store *s, const double *x, size_t rhs, double *y) {
    const size_t n = s->order;
    for (size_t i = 0; i < n * rhs; ++i)
        if (!isfinite(x[i])) return -1;
    memset(y, 0, n * rhs * sizeof(double));
    for (size_t b = 0; b < s->nblocks; ++b) {
        const size_t r0 = s->rect[4 * b], r1 = s->rect[4 * b + 1];
        const size_t c0 = s->rect[4 * b + 2], c1 = s->rect[4 * b + 3];
        const size_t m = r1 - r0, nc = c1 - c0;
        if (s->is_dense[b]) {
            const double *d = s->dense + s->dense_off[b];
            for (size_t i = 0; i < m; ++i)
                for (size_t k = 0; k < nc; ++k)
                    for (size_t p = 0; p < rhs; ++p)
                        y[s->perm[r0 + i] * rhs + p] +=
                            d[i * nc + k] * x[s->perm[c0 + k] * rhs + p];
            continue;
        }
        const size_t base = s->comp_off[b];
        for (size_t p = 0; p < rhs; ++p)
            for (size_t j = 0; j < s->rank[b]; ++j) {
                const size_t c = base + j;
                const int bits = s->widths[c];
                const uint32_t *rw = s->rwords + s->roff[c];
                const uint32_t *lw = s->lwords + s->loff[c];
                double dot = 0.0;
                for (size_t k = 0; k < nc; ++k)
                    dot += orc_decode(rw, s->rscale[c], k, bits) *
                           x[s->perm[c0 + k] * rhs + p];
                for (size_t i = 0; i < m; ++i)
                    y[s->perm[r0 + i] * rhs + p] += orc_decode(lw, s->lscale[c], i, bits) * dot;
            }
    }
    return 0;
}

/* core.cpp:392-410: every entry of every block, written in source order. */
void orc_reconstruct(const orc_store *s, double *out) {
    const size_t n = s->order;
    for (size_t b = 0; b < s->nblocks; ++b) {
        const size_t r0 = s->rect[4 * b], r1 = s->rect[4 * b + 1];
        const size_t c0 = s->rect[4 * b + 2], c1 = s->rect[4 * b + 3];
        const size_t m = r1 - r0, nc = c1 - c0;
        const size_t rank = s->is_dense[b] ? 0 : s->rank[b];
        const size_t base = s->is_dense[b] ? 0 : s->comp_off[b];
        for (size_t i = 0; i < m; ++i)
            for (size_t k = 0; k < nc; ++k) {
                double v = s->is_dense[b] ? s->dense[s->dense_off[b] + i * nc + k] : 0.0;
                for (size_t j = 0; j < rank; ++j) {
                    const size_t c = base + j;
                    v += orc_decode(s->lwords + s->loff[c], s->lscale[c], i, s->widths[c]) *
                         orc_decode(s->rwords + s->roff[c], s->rscale[c], k, s->widths[c]);
                }
                out[s->perm[r0 + i] * n + s->perm[c0 + k]] = v;
            }
    }
}
