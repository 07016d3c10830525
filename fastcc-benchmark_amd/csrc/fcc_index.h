// Index-build entry points (fcc_index.hip).
#pragma once
#include <functional>

#include "fcc_internal.h"

namespace fcc {

void fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s);
void fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s);
void fill_f64(double *p, size_t n, double v, hipStream_t s);
// up to 8 fills (4- or 8-byte elements) in one launch
struct FillOp {
    void *p;
    size_t n;
    unsigned long long v;  // the element's bits
    uint32_t bytes;        // 4 or 8
};
struct FillSet {
    FillOp op[8];
    int n = 0;
    void add(void *p, size_t count, unsigned long long v, uint32_t bytes) {
        if (p && count) op[n++] = FillOp{p, count, v, bytes};
    }
};
void fill_many(const FillSet &f, hipStream_t s);
// dst[p] = src[perm[p]]
void perm_gather(const double *src, const uint32_t *perm, size_t n, double *dst, hipStream_t s);

// out[m] = max coordinate of mode m (m < order); out[order] |= 1 when a value
// is zero or tiny, |= 2 when a value is infinite or NaN. out must be zeroed by
// the caller.
void launch_extents(const Operand &t, uint32_t *dev_out, hipStream_t s);

// Left join read off the right operand's sort (self-contraction).
struct SelfJoin {
    uint32_t *rowA;  // row and fiber per original left entry; null: not written
    uint32_t *fidA;
    uint32_t *bfib;  // fiber of each right-sorted position (set by build_right: the index's Bfib)
    uint32_t *vsB;   // original entry of each right-sorted position
    uint64_t products;
};

// Right operand: linearise, sort by (key, col), fibers. Synchronises to read
// nfib. With self != null also fills the left join of a self-contraction.
void build_right(DevPool &pool, const Geom &g, const Operand &B, Index &ix,
                 std::vector<std::shared_ptr<void>> &hold, hipStream_t s, SelfJoin *self = nullptr);

// Left operand join: row and matching fiber per entry; total products and
// matched entry count (synchronises).
void join_left(DevPool &pool, const Geom &g, const Operand &A, const Index &ix, uint32_t *rowA, uint32_t *fidA,
               uint64_t *products, uint32_t *nvalid, hipStream_t s);

// Self-contraction: left tiles from the right operand's order (2 radix passes).
// after_sort (optional) runs once the left entries are sorted, before the
// rest of the build is queued (a point to start independent work on another
// stream that should not compete with the sort). records_late: the entries'
// source records are written after after_sort instead of before the sort
// (beside a long block rather than ahead of it).
void build_left_self(DevPool &pool, const Geom &g, const Operand &A, const SelfJoin &sj, int logTL, Index &ix,
                     std::vector<std::shared_ptr<void>> &hold, hipStream_t s,
                     const std::function<void()> *after_sort = nullptr, bool records_late = false);

// Left operand tiles of 2^logTL rows: sort, groups, active tiles, statistics.
void build_left(DevPool &pool, const Geom &g, const Operand &A, const uint32_t *rowA, const uint32_t *fidA,
                uint32_t nvalid, int logTL, Index &ix, std::vector<std::shared_ptr<void>> &hold, hipStream_t s,
                const std::function<void()> *after_sort = nullptr);

// Heavy-group segments (hdesc, atile_h0) from the group descriptors; run
// again after descriptors change (new buffers are kept in hold).
void build_hseg(DevPool &pool, Index &ix, std::vector<std::shared_ptr<void>> &hold, hipStream_t s);

}  // namespace fcc
