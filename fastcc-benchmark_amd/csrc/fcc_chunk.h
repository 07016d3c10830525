// Chunking of a fused batch (fcc_contract_batch): pure host C++, no HIP.
//
// The pairs of one fused launch are concatenated with a pair-index mode that
// leads both the contraction key and the left external index. A chunk of cnt
// pairs therefore has
//   left span   cnt * prod(A's other extents)      (must stay below 2^32)
//   right span  prod(B's other extents)            (must stay below 2^32)
//   key span    cnt * prod(paired key radices)     (key bits + column bits <= 64)
//   left composite key: row bits + fiber bits + 2 <= 64 (fibers <= B entries)
// and at most `max_pairs` pairs and fewer than 2^32 entries per side, with
// extents the per-mode maximum over the chunk's pairs. batch_chunks() splits
// pairs [0, n) greedily into consecutive chunks that satisfy all of this; a
// pair that violates it alone forms a chunk of its own (the contraction then
// reports the same error as for that pair on its own).
#pragma once
#include <algorithm>
#include <cstdint>
#include <vector>

namespace fcc {

inline int chunk_ceil_log2(uint64_t x) {
    int b = 0;
    while (b < 64 && (1ULL << b) < x) b++;
    return b;
}

// a * b, saturating at 2^64 - 1
inline uint64_t chunk_mul(uint64_t a, uint64_t b) {
    uint64_t z;
    return __builtin_mul_overflow(a, b, &z) ? ~0ull : z;
}

// extA: n * dA extents (max coordinate + 1, or the caller's), extB: n * dB;
// nnzA / nnzB: entries per pair; contraction modes am / bm (r pairs).
// Returns the chunk ends: chunk k is [ends[k-1], ends[k]) with ends[-1] = 0.
inline std::vector<uint32_t> batch_chunks(uint32_t n, int dA, int dB, const int32_t *am, const int32_t *bm, int r,
                                          const uint32_t *extA, const uint32_t *extB, const uint64_t *nnzA,
                                          const uint64_t *nnzB, uint32_t max_pairs) {
    std::vector<uint32_t> ends;
    auto contracted = [](const int32_t *l, int k, int m) {
        for (int i = 0; i < k; i++)
            if (l[i] == m) return true;
        return false;
    };
    auto fits = [&](uint32_t c0, uint32_t c1) {
        std::vector<uint64_t> EA(dA, 0), EB(dB, 0);
        uint64_t na = 0, nb = 0;
        for (uint32_t i = c0; i < c1; i++) {
            for (int m = 0; m < dA; m++) EA[m] = std::max<uint64_t>(EA[m], extA[(uint64_t)i * dA + m]);
            for (int m = 0; m < dB; m++) EB[m] = std::max<uint64_t>(EB[m], extB[(uint64_t)i * dB + m]);
            na += nnzA[i];
            nb += nnzB[i];
        }
        const uint64_t cnt = c1 - c0;
        if (cnt > max_pairs || na >= (1ull << 32) || nb >= (1ull << 32)) return false;
        uint64_t spanA = cnt, spanB = 1, K = cnt;
        for (int m = 0; m < dA; m++)
            if (!contracted(am, r, m)) spanA = chunk_mul(spanA, EA[m]);
        for (int m = 0; m < dB; m++)
            if (!contracted(bm, r, m)) spanB = chunk_mul(spanB, EB[m]);
        for (int i = 0; i < r; i++) K = chunk_mul(K, std::max(EA[am[i]], EB[bm[i]]));
        if (spanA >= (1ull << 32) || spanB >= (1ull << 32) || K == ~0ull) return false;
        const int kbits = std::max(1, chunk_ceil_log2(K)), cbits = std::max(1, chunk_ceil_log2(spanB));
        const int rbits = std::max(1, chunk_ceil_log2(spanA)), fbits = std::max(1, chunk_ceil_log2(nb + 1));
        return kbits + cbits <= 64 && rbits + fbits + 2 <= 64;
    };
    uint32_t c0 = 0;
    while (c0 < n) {
        // grow by doubling, then bisect the largest chunk that fits
        uint32_t good = c0 + 1, step = 1;
        while (good < n) {
            const uint32_t cand = std::min<uint64_t>((uint64_t)good + step, n);
            if (!fits(c0, cand)) break;
            good = cand;
            step *= 2;
        }
        uint32_t lo = good, hi = std::min<uint64_t>((uint64_t)good + step, n);  // fits(c0, lo); hi may not
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (fits(c0, mid)) lo = mid;
            else hi = mid;
        }
        if (hi > lo && fits(c0, hi)) lo = hi;
        ends.push_back(lo);
        c0 = lo;
    }
    return ends;
}

}  // namespace fcc
