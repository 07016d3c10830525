// Fused batches of independent contractions: operand concatenation with a
// pair-index mode, per-pair split of the fused output, per-group unit counts.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fcc_batch.h"

namespace fcc {

namespace {

unsigned blocks_for(uint64_t n, unsigned cap) {
    uint64_t b = (n + 255) / 256;
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

// Wave-aggregated add of 1 per active lane to s_h[key]: one LDS atomic per
// wave when every active lane carries the same key (the common case: fused
// outputs are grouped by pair), else one per lane. Returns this lane's slot:
// s_h[key] before the add plus the lane's rank among the wave's lanes on key.
__device__ __forceinline__ uint32_t wave_add(uint32_t *s_h, bool t, uint32_t key) {
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t act = __ballot(t);
    if (!act) return 0;
    const int first = __ffsll((unsigned long long)act) - 1;
    const uint32_t k0 = __shfl(key, first, 64);
    if (__ballot(t && key != k0) == 0) {
        uint32_t base = 0;
        if (lane == (uint32_t)first) base = atomicAdd(&s_h[k0], (uint32_t)__popcll(act));
        base = __shfl(base, first, 64);
        const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
        return base + (uint32_t)__popcll(act & lt);
    }
    return t ? atomicAdd(&s_h[key], 1u) : 0u;
}

__global__ void k_zero_u32(uint32_t *p, uint64_t n) {
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) p[i] = 0u;
}

// The maximum coordinate of each mode per pair (0 when empty), atomically
// into out (zeroed by the caller): blockIdx.y is the pair, blockIdx.x one of
// gridDim.x equal slices of its entries (slices past its end exit at once).
__global__ __launch_bounds__(256) void k_pair_max(const PairSrc *src, int order, uint32_t *out) {
    __shared__ uint32_t s_m[4];
    const PairSrc ps = src[blockIdx.y];
    const uint64_t per = (ps.nnz + gridDim.x - 1) / gridDim.x;
    const uint64_t j0 = (uint64_t)blockIdx.x * per, j1 = min(ps.nnz, j0 + per);
    if (j0 >= j1) return;
    for (int m = 0; m < order; m++) {
        const uint32_t *c = ps.coords + (uint64_t)m * ps.nnz;
        uint32_t v = 0;
#pragma unroll 4
        for (uint64_t j = j0 + threadIdx.x; j < j1; j += 256) v = max(v, c[j]);
        for (int d = 32; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
        if ((threadIdx.x & 63) == 0) s_m[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0)
            atomicMax(&out[(uint64_t)blockIdx.y * order + m], max(max(s_m[0], s_m[1]), max(s_m[2], s_m[3])));
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_gather_pairs(const PairSrc *src, const unsigned long long *offs, uint32_t n,
                                                      int order, uint64_t total, uint32_t *dst, double *dv,
                                                      uint32_t *dmax) {
    __shared__ unsigned long long s_off[kMaxBatchPairs + 1];
    __shared__ uint32_t s_max[4][kMaxOrder];
    for (uint32_t i = threadIdx.x; i <= n; i += 256) s_off[i] = offs[i];
    __syncthreads();
    uint32_t mx[kMaxOrder];
#pragma unroll
    for (int m = 0; m < kMaxOrder; m++) mx[m] = 0;
    for (uint64_t p = blockIdx.x * 256ull + threadIdx.x; p < total; p += (uint64_t)gridDim.x * 256) {
        // s_off[lo] <= p < s_off[hi]: pair lo is the one holding p (empty pairs skipped)
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (s_off[mid] <= p) lo = mid;
            else hi = mid;
        }
        const PairSrc ps = src[lo];
        const uint64_t j = p - s_off[lo];
#pragma unroll
        for (int m = 0; m < kMaxOrder; m++)
            if (m < order) {
                uint32_t c = ps.coords[(uint64_t)m * ps.nnz + j];
                dst[(uint64_t)m * total + p] = c;
                mx[m] = max(mx[m], c);
            }
        dst[(uint64_t)order * total + p] = lo;
        dv[p] = ps.vals[j];
    }
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int m = 0; m < kMaxOrder; m++) {
        uint32_t v = mx[m];
        for (int d = 32; d > 0; d >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, d, 64));
        if (lane == 0) s_max[wave][m] = v;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)order) {
        const int m = threadIdx.x;
        uint32_t v = max(max(s_max[0][m], s_max[1][m]), max(s_max[2][m], s_max[3][m]));
        if (v) atomicMax(&dmax[m], v);
    }
}

// contiguous chunk of the fused output handled by block b (same in both passes)
__device__ __forceinline__ void chunk_of(uint64_t nnz, uint64_t *lo, uint64_t *hi) {
    uint64_t per = (nnz + gridDim.x - 1) / gridDim.x;
    per = (per + 255) & ~255ull;
    const uint64_t a = (uint64_t)blockIdx.x * per;
    *lo = a < nnz ? a : nnz;
    *hi = *lo + per < nnz ? *lo + per : nnz;
}

__global__ __launch_bounds__(256) void k_split_count(const uint32_t *pair, uint64_t nnz, uint32_t n,
                                                     unsigned long long *counts) {
    __shared__ uint32_t s_h[kMaxBatchPairs];
    for (uint32_t i = threadIdx.x; i < n; i += 256) s_h[i] = 0;
    __syncthreads();
    uint64_t lo, hi;
    chunk_of(nnz, &lo, &hi);
    for (uint64_t b = lo; b < hi; b += 256) {
        const uint64_t p = b + threadIdx.x;
        const uint32_t q = p < hi ? pair[p] : n;
        wave_add(s_h, q < n, q < n ? q : 0u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += 256)
        if (s_h[i]) atomicAdd(&counts[i], (unsigned long long)s_h[i]);
}

__global__ __launch_bounds__(256) void k_split_scatter(const uint32_t *coords, uint64_t stride, const double *vals,
                                                       uint64_t nnz, int dC, uint32_t n,
                                                       const unsigned long long *offs,
                                                       const unsigned long long *counts, unsigned long long *cursor,
                                                       uint32_t *dst_c, double *dst_v) {
    __shared__ uint32_t s_h[kMaxBatchPairs];
    __shared__ unsigned long long s_base[kMaxBatchPairs];
    for (uint32_t i = threadIdx.x; i < n; i += 256) s_h[i] = 0;
    __syncthreads();
    uint64_t lo, hi;
    chunk_of(nnz, &lo, &hi);
    for (uint64_t b = lo; b < hi; b += 256) {
        const uint64_t p = b + threadIdx.x;
        const uint32_t q = p < hi ? coords[p] : n;
        wave_add(s_h, q < n, q < n ? q : 0u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += 256) {
        if (s_h[i]) s_base[i] = atomicAdd(&cursor[i], (unsigned long long)s_h[i]);
        s_h[i] = 0;
    }
    __syncthreads();
    for (uint64_t b = lo; b < hi; b += 256) {
        const uint64_t p = b + threadIdx.x;
        uint32_t q = p < hi ? coords[p] : n;
        const bool t = q < n;  // a pair index is always < n; the test only guards the tables
        if (!t) q = 0;
        const uint32_t rank = wave_add(s_h, t, q);
        if (t) {
            const uint64_t pos = s_base[q] + rank;
            const uint64_t cnt = counts[q], off = offs[q];
            uint32_t *dc = dst_c + (uint64_t)(dC - 1) * off;
            for (int m = 1; m < dC; m++) dc[(uint64_t)(m - 1) * cnt + pos] = coords[(uint64_t)m * stride + p];
            dst_v[off + pos] = vals[p];
        }
    }
}

// units arrive in tile order, so a wavefront's units mostly share a group:
// one atomic per wavefront then
__global__ void k_group_units(const uint32_t *unit_tile, const uint32_t *atile_id, uint32_t nunits, int logTL,
                              uint64_t group_rows, unsigned int *left) {
    const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    const bool t = u < nunits;
    const uint32_t g = t ? (uint32_t)(((uint64_t)atile_id[unit_tile[u]] << logTL) / group_rows) : 0u;
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t act = __ballot(t);
    if (!act) return;
    const int first = __ffsll((unsigned long long)act) - 1;
    const uint32_t g0 = __shfl(g, first, 64);
    if (__ballot(t && g != g0) == 0) {
        if (lane == (uint32_t)first) atomicAdd(&left[g0], (unsigned int)__popcll(act));
    } else if (t) {
        atomicAdd(&left[g], 1u);
    }
}

}  // namespace

void pair_max(const PairSrc *src, uint32_t n, int order, uint64_t max_nnz, uint32_t *out, hipStream_t s) {
    if (!n) return;
    k_zero_u32<<<blocks_for((uint64_t)n * order, 1024), 256, 0, s>>>(out, (uint64_t)n * order);
    // about 64k entries per slice, at most 256 slices of the largest pair;
    // at most 65535 pairs per launch (grid y)
    const uint32_t slices = (uint32_t)std::min<uint64_t>(256, std::max<uint64_t>(1, (max_nnz + 65535) / 65536));
    for (uint32_t p0 = 0; p0 < n; p0 += 65535) {
        const uint32_t np = std::min<uint32_t>(65535, n - p0);
        k_pair_max<<<dim3(slices, np), 256, 0, s>>>(src + p0, order, out + (uint64_t)p0 * order);
    }
}

void gather_pairs(const PairSrc *src, const unsigned long long *offs, uint32_t n, int order, uint64_t total,
                  uint32_t *dst_coords, double *dst_vals, uint32_t *dmax, hipStream_t s) {
    if (!total) return;
    k_gather_pairs<<<blocks_for(total, 4096), 256, 0, s>>>(src, offs, n, order, total, dst_coords, dst_vals, dmax);
}

void split_count(const uint32_t *pair_row, uint64_t nnz, uint32_t n, unsigned long long *counts, hipStream_t s) {
    if (!nnz) return;
    k_split_count<<<blocks_for(nnz / 16, 2048), 256, 0, s>>>(pair_row, nnz, n, counts);
}

void split_scatter(const uint32_t *coords, uint64_t stride, const double *vals, uint64_t nnz, int dC, uint32_t n,
                   const unsigned long long *offs, const unsigned long long *counts, unsigned long long *cursor,
                   uint32_t *dst_c, double *dst_v, hipStream_t s) {
    if (!nnz) return;
    k_split_scatter<<<blocks_for(nnz / 16, 2048), 256, 0, s>>>(coords, stride, vals, nnz, dC, n, offs, counts, cursor,
                                                               dst_c, dst_v);
}

void group_units(const uint32_t *unit_tile, const uint32_t *atile_id, uint32_t nunits, int logTL,
                 uint64_t group_rows, unsigned int *left, hipStream_t s) {
    if (!nunits) return;
    k_group_units<<<(nunits + 255) / 256, 256, 0, s>>>(unit_tile, atile_id, nunits, logTL, group_rows, left);
}

}  // namespace fcc
