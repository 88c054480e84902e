/*
 * CPU oracle for the qhmat H-matrix apply -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's FP64 CPU path
 * (src/qhmat/core.cpp) used to check the HIP product path.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker or as the timed CPU baseline.  The product library never links
 * or calls it.
 *
 * Functions and the reference lines they restate:
 *   orc_limit            core.cpp:32-35    q = 2^(b-1) - 1
 *   orc_coefficient      core.cpp:56-59    c(b) = 1/q + 1/(4 q^2)
 *   orc_pack_component   core.cpp:235-267  one component: scale, round, clamp, LSB-first bits
 *   orc_decode_code      core.cpp:269-280  integer code at (row) of one component stream
 *   orc_decode           core.cpp:269-282  code * scale
 *   orc_decode_factor    core.cpp:248-251,269-280  all codes of a factor
 *   orc_matmul           core.cpp:360-390  y = H x, block insertion order, same loop nest
 *   orc_reconstruct      core.cpp:392-410  dense FP64 in source order
 *
 * Parity pins (tests/test_oracle.py): the reference's own known answers
 * (test_core.py:44-60 2x2 matrix, :154-167 rank-0 blocks, :170-192 random
 * orthogonal factors at every width), hand-computed code words, brute-force
 * long-double sums on tiny stores, and -- where oracle/_ref builds -- word-for-
 * word and bit-for-bit agreement with the reference's compiled core.
 */
#ifndef QHMAT_ORACLE_H
#define QHMAT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Flat view of a finalized store.  Coordinates are clustered indices;
 * perm[clustered] = source index (core.hpp:92-93).  For block b:
 *   rect[4b..4b+3] = row0,row1,col0,col1 (half-open)
 *   is_dense[b] != 0 -> dense[dense_off[b] .. + m*n) row-major
 *   otherwise rank[b] components starting at comp_off[b] in widths/lscale/
 *   rscale/loff/roff; loff/roff are absolute word offsets into lwords/rwords. */
typedef struct {
    size_t order;
    const size_t *perm;
    size_t nblocks;
    const size_t *rect;
    const unsigned char *is_dense;
    const double *dense;
    const size_t *dense_off;
    const size_t *rank;
    const size_t *comp_off;
    const int *widths;
    const double *lscale;
    const double *rscale;
    const uint32_t *lwords;
    const uint32_t *rwords;
    const size_t *loff;
    const size_t *roff;
} orc_store;

uint64_t orc_limit(int bits);
double orc_coefficient(int bits);
size_t orc_words_for(size_t rows, int bits);
/* column: rows values of one balanced factor component; words must hold
 * orc_words_for(rows, bits) zeroed words on entry. */
void orc_pack_component(const double *column, size_t rows, int bits,
                        uint32_t *words, double *scale);
int64_t orc_decode_code(const uint32_t *component_words, size_t row, int bits);
double orc_decode(const uint32_t *component_words, double scale, size_t row,
                  int bits);
/* Every code of a packed factor, component-major: out[j*rows + i]. */
void orc_decode_factor(const uint32_t *words, size_t rows, const int *widths, size_t rank,
                       int64_t *out);
/* x, y: order x rhs row-major in source order.  Returns 0, or -1 on a
 * non-finite x (core.cpp:365). */
int orc_matmul(const orc_store *s, const double *x, size_t rhs, double *y);
void orc_reconstruct(const orc_store *s, double *out);

#ifdef __cplusplus
}
#endif
#endif
