// Dense contraction keys (fcc_dense.hip).
//
// A contraction key k whose left fiber A(:, k) and right fiber B(k, :) are
// both dense contributes nA(k) * nB(k) products that cover a large part of
// the output block spanA x spanB. Such keys are taken out of the sparse
// traversal and contracted together as one dense FP64 GEMM on the matrix
// cores:
//     C_dk[row][col] = sum over selected keys h of AT[h][row] * Bd[h][col]
// Structural support stays exact: bit h of rowbits[row] says A(row, key h)
// is stored, bit h of colbits[col] says B(key h, col) is stored, and the
// presence map P[row][col] = (rowbits[row] & colbits[col]) != 0. Every unit
// of work then starts its accumulator from (P, C_dk) over its rows and column
// window instead of from zero.
#pragma once
#include "fcc_internal.h"

namespace fcc {

constexpr uint32_t DK_BM = 128;  // GEMM tile rows (left rows)
constexpr uint32_t DK_BN = 128;  // GEMM tile columns (right columns)
constexpr uint32_t DK_FIBER_SPLIT = 8;  // workgroups per selected fiber in the panel and bitset scatters
constexpr uint32_t DK_BK = 16;   // GEMM k step (selected keys)

struct DenseKeys {
    uint32_t nh = 0;          // selected keys (0: the path is off)
    uint32_t nw = 0;          // 64-bit words per key bitset
    uint32_t kp = 0;          // nh rounded up to DK_BK
    uint64_t spanA = 0, spanB = 0;
    uint64_t ra = 0, rb = 0;  // spanA, spanB rounded up to DK_BM, DK_BN
    uint64_t products = 0;    // products of the selected keys
    float tau = 0;            // density threshold used
    bool self = false;        // self-contraction: AT is Bd, row bitsets are column bitsets
    bool upper = false;       // C's lower triangle is never read (mirror mode reads the cells on and
                              // above the diagonal only): the symmetric GEMM stores its tiles once
    bool full = false;        // a selected key is stored at every left row and every right column:
                              // every cell of the block is present (P all ones, no bitsets)
    uint32_t splits = 1;      // GEMM k splits: C holds this many ra x rb slabs, summed by the readers
    uint32_t *hkey = nullptr;              // per fiber: selected key index, or ~0
    uint32_t *hfib = nullptr;              // per selected key: its fiber
    uint8_t *tile_dk = nullptr;            // per active tile: 1 if it holds a selected key
    unsigned long long *rowbits = nullptr; // [spanA][nw] (null for a self-contraction)
    unsigned long long *colbits = nullptr; // [nw][rb] (word-major: coalesced per column)
    uint32_t *P = nullptr;                 // presence [ra][rb / 32] (+2 words)
    double *AT = nullptr;                  // [kp][ra]
    double *Bd = nullptr;                  // [kp][rb]
    double *C = nullptr;                   // [splits][ra][rb]
};

// Chooses the dense keys once the right fibers and the left join are known
// (before the left tiles are built): keys with nA * nB >= tau * spanA *
// spanB, at most 16384 of them, the block within max_cells cells; with
// require_gain only when their products average four or more per block cell.
// fidA: the join's fiber per left entry (nA entries), or null for a
// self-contraction (left fibers = right fibers). Fills nh, products, hkey,
// hfib, full and the sizes; synchronises. Leaves dk.nh = 0 when none qualify.
void dense_keys_select(DevPool &pool, const Geom &g, const Index &ix, const uint32_t *fidA, size_t nA, float tau,
                       uint64_t max_cells, bool require_gain, DenseKeys &dk, std::vector<std::shared_ptr<void>> &hold,
                       hipStream_t s);

// Allocates the bitsets, presence map, value panels and block (host only,
// no kernels: everything the block needs exists before it is launched, so it
// can run on another stream while the left tiles are built).
void dense_keys_alloc(DevPool &pool, const Geom &g, DenseKeys &dk, int num_cus,
                      std::vector<std::shared_ptr<void>> &hold);

// The block from the right fibers and the left join (fiber and row per left
// entry, nA entries; unused for a self-contraction): column (and row)
// bitsets, presence map, value panels, GEMM. Reads nothing the left-tile
// build writes; fidA/rowA/Avals must stay valid until it completes.
// gemm0 / gemm1 (optional) are recorded around the GEMM launch. When a full
// key is selected the presence map is a plain fill, queued on stream sp (the
// caller's stream, which reads the map next) so that it does not delay the
// GEMM on s.
void dense_keys_block(const Index &ix, const DenseKeys &dk, const uint32_t *fidA, const uint32_t *rowA,
                      const double *Avals, size_t nA, hipStream_t s, hipStream_t sp, hipEvent_t gemm0 = nullptr,
                      hipEvent_t gemm1 = nullptr);

// Values again (plans): scatters the selected keys' current values (ix.Aval,
// ix.Bval, index order) into the panels and runs the GEMM into dk.C. No memset nodes and no
// allocation: safe to capture into a graph.
void dense_keys_values(const Index &ix, const DenseKeys &dk, hipStream_t s);

// C[z][ra][rb] (z < splits) = partial sums of AT^T * Bd for row-major
// AT[kp][ra], Bd[kp][rb]; ra, rb multiples of DK_BM / DK_BN, kp a multiple
// of DK_BK; split z sums its share of the k steps. sym (AT == Bd, ra == rb:
// C is symmetric) computes the tiles on and above the diagonal and mirrors
// them, half the matrix-core work; with upper the mirror copies are not
// stored (the 128 x 128 tiles on and above the diagonal hold every cell
// with column >= row).
void launch_dgemm_tn(const double *AT, const double *Bd, double *C, uint64_t ra, uint64_t rb, uint32_t kp,
                     uint32_t splits, bool sym, bool upper, hipStream_t s);
// k splits that fill the GPU's workgroup slots for a small grid (1 for large).
uint32_t dgemm_splits(uint64_t ra, uint64_t rb, uint32_t kp, bool sym, int num_cus);

}  // namespace fcc
