// Fused batches of independent contractions (fcc_batch.hip).
//
// n pairs (A_i, B_i) with the same orders and mode lists run as ONE batch-mode
// contraction: each operand is concatenated with an extra trailing mode that
// holds the pair index, and that mode is a batch pair of the contraction. The
// pair index then leads every contraction key (so entries of different pairs
// never meet) and every left row (so a pair's rows are one contiguous group),
// and all pairs' units share the persistent accumulate launch.
#pragma once
#include "fcc_internal.h"

namespace fcc {

constexpr uint32_t kMaxBatchPairs = 4096;  // pairs per fused launch (LDS-resident tables)

struct PairSrc {
    const uint32_t *coords;  // SoA [order][nnz] of this pair
    const double *vals;
    uint64_t nnz;
};

// out[i * order + m] = max coordinate of mode m of pair i (0 for an empty pair);
// max_nnz: the largest pair's entry count (sizes the launch).
void pair_max(const PairSrc *src, uint32_t n, int order, uint64_t max_nnz, uint32_t *out, hipStream_t s);

// dst_coords: (order + 1) rows of `total`; row `order` receives the pair
// index. offs: n + 1 exclusive offsets (device). dmax[m] |= max coordinate of
// mode m over all pairs (atomicMax; zero it first).
void gather_pairs(const PairSrc *src, const unsigned long long *offs, uint32_t n, int order, uint64_t total,
                  uint32_t *dst_coords, double *dst_vals, uint32_t *dmax, hipStream_t s);

// Per-pair output counts of a fused result whose coordinate row 0 is the pair
// index (counts zeroed by the caller).
void split_count(const uint32_t *pair_row, uint64_t nnz, uint32_t n, unsigned long long *counts, hipStream_t s);

// Scatters a fused result into per-pair SoA blocks: pair i's block starts at
// coords dst_c + (dC - 1) * offs[i] (rows of counts[i]) and vals dst_v +
// offs[i]. cursor: n zeroed counters.
void split_scatter(const uint32_t *coords, uint64_t stride, const double *vals, uint64_t nnz, int dC, uint32_t n,
                   const unsigned long long *offs, const unsigned long long *counts, unsigned long long *cursor,
                   uint32_t *dst_c, double *dst_v, hipStream_t s);

// left[g] = number of units whose first row lies in group g (zeroed by the caller).
void group_units(const uint32_t *unit_tile, const uint32_t *atile_id, uint32_t nunits, int logTL,
                 uint64_t group_rows, unsigned int *left, hipStream_t s);

}  // namespace fcc
