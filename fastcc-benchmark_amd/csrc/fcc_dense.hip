// Dense contraction keys: selection, masking out of the sparse traversal,
// key bitsets, presence map, value panels and the FP64 MFMA GEMM.
//
// The reference contracts every key through its tile accumulators
// (contract.hpp:850-990: for each left tile, for each contraction index, every
// pair of left and right entries is multiplied into the accumulator). The
// products of a key whose left and right fibers are both dense form a rank-1
// update of a dense block; many such keys together are a dense GEMM, which
// the matrix cores run at the FP64 peak instead of one LDS atomic per
// product. The result, the set of outputs and their values, is the same.
#include "fcc_dense.h"

#include <algorithm>

#include "fcc_index.h"

namespace fcc {

namespace {

inline unsigned grid_of(uint64_t n, unsigned b = 256) { return (unsigned)std::max<uint64_t>(1, div_up(n, b)); }

typedef double f64x4 __attribute__((ext_vector_type(4)));

// left entries per fiber from the join (fiber per left entry, ~0 when
// unmatched); only fibers long enough on the right to reach the threshold
// (nB >= nb_min) are counted, which keeps the atomics off the many short ones
__global__ void k_dk_na(const uint32_t *fidA, size_t n, const uint32_t *fiber_ptr, uint32_t nb_min, uint32_t *na) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t f = fidA[i];
    if (f != 0xFFFFFFFFu && fiber_ptr[f + 1] - fiber_ptr[f] >= nb_min) atomicAdd(&na[f], 1u);
}

// a fiber is selected when nA * nB >= tau * spanA * spanB; tot[0] counts the
// selected fibers, tot[1] sums their products (wave-aggregated), tot[2] is set
// when a selected fiber is full on both sides (nA = spanA, nB = spanB: its
// products alone reach every cell of the block). na == null: a
// self-contraction, whose left fibers are the right ones (nA = nB).
__global__ void k_dk_flag(const uint32_t *fiber_ptr, const uint32_t *na, uint32_t nfib, double inv_area, double tau,
                          uint64_t spanA, uint64_t spanB, uint32_t *flag, unsigned long long *tot) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c = 0, prod = 0;
    bool full = false;
    if (f < nfib) {
        const uint64_t b = fiber_ptr[f + 1] - fiber_ptr[f], a = na ? na[f] : b;
        const bool sel = a && (double)a * (double)b * inv_area >= tau;
        flag[f] = sel ? 1u : 0u;
        c = sel ? 1ull : 0ull;
        prod = sel ? a * b : 0ull;
        full = sel && a == spanA && b == spanB;
    }
    for (int d = 32; d > 0; d >>= 1) {
        c += __shfl_xor(c, d, 64);
        prod += __shfl_xor(prod, d, 64);
    }
    const bool any_full = __ballot(full) != 0ull;
    if ((threadIdx.x & 63) == 0 && c) {
        atomicAdd(&tot[0], c);
        atomicAdd(&tot[1], prod);
        if (any_full) atomicOr(&tot[2], 1ull);
    }
}

// Presence map when a full key is selected: every cell (row < spanA, column
// < spanB) is present; the padding rows and columns of the block are not.
// One thread per four words (rows hold pw = rb / 32 words, a multiple of 4;
// the block holds at most 2^29 cells, so word indices fit 32 bits).
__global__ void k_dk_full_presence(uint4 *P4, uint32_t ra, uint32_t pw4, uint32_t spanA, uint32_t spanB) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ra * pw4) return;
    const uint32_t row = i / pw4, c0 = (i - row * pw4) * 128;
    uint32_t w[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t c = c0 + 32u * j;
        w[j] = (row < spanA && c < spanB) ? (spanB - c >= 32 ? 0xFFFFFFFFu : ((1u << (spanB - c)) - 1u)) : 0u;
    }
    P4[i] = make_uint4(w[0], w[1], w[2], w[3]);
}

// hkey[f] = the selected key index of fiber f (or none), hfib[h] = its fiber;
// hfib holds `cap` entries (a selection over the cap is discarded by the host)
__global__ void k_dk_keys(const uint32_t *flag, const uint32_t *idx, uint32_t nfib, uint32_t cap, uint32_t *hkey,
                          uint32_t *hfib) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfib) return;
    const bool sel = flag[f] != 0;
    hkey[f] = sel ? idx[f] : 0xFFFFFFFFu;
    if (sel && idx[f] < cap) hfib[idx[f]] = f;
}

// one workgroup per selected key: column bitsets
// A selected key's fiber is spread over gridDim.y workgroups (Zipf keys:
// the longest fibers hold a large share of the selected entries).
__global__ __launch_bounds__(256) void k_dk_colbits(const uint32_t *hfib, const uint32_t *fiber_ptr,
                                                   const uint32_t *Bcol, uint64_t rb, unsigned long long *colbits) {
    const uint32_t h = blockIdx.x, f = hfib[h];
    const unsigned long long bit = 1ull << (h & 63);
    unsigned long long *row = colbits + (uint64_t)(h >> 6) * rb;
    const uint32_t step = gridDim.y * blockDim.x;
    for (uint32_t p = fiber_ptr[f] + blockIdx.y * blockDim.x + threadIdx.x; p < fiber_ptr[f + 1]; p += step)
        atomicOr(&row[Bcol[p]], bit);
}

// P[row][c >> 5] bit (c & 31) = (rowbits[row] & colbits[c]) != 0.
// Workgroup: 64 rows x 64 columns; wavefront: 16 rows, a lane per column,
// its column's bitset words held in registers 8 at a time; the 64 rows'
// bitset words are staged in LDS 8 words per row at a time (broadcast reads).
// rowbits == null: a self-contraction, whose row bitsets are its column
// bitsets (colbits word-major, read transposed). Its map is symmetric: only
// blocks on and above the diagonal are computed, each also writing its
// mirror block (a 64 x 64 bit transpose through LDS).
__global__ __launch_bounds__(256) void k_dk_presence(const unsigned long long *rowbits,
                                                    const unsigned long long *colbits, uint64_t spanA, uint64_t rb,
                                                    uint32_t nw, uint32_t *P) {
    const bool sym = rowbits == nullptr;
    if (sym && blockIdx.y > blockIdx.x) return;  // written by block (x, y)'s mirror
    __shared__ unsigned long long s_rw[64][8];
    __shared__ uint32_t s_h[64][4];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t c = (uint64_t)blockIdx.x * 64 + lane;
    const uint64_t rb0 = (uint64_t)blockIdx.y * 64;
    const uint32_t lr0 = wave * 16;
    uint32_t hit = 0;
    for (uint32_t w0 = 0; w0 < nw; w0 += 8) {
        __syncthreads();
        for (uint32_t q = threadIdx.x; q < 64 * 8; q += 256) {
            const uint32_t rr = q >> 3, j = q & 7;
            const uint64_t row = rb0 + rr;
            unsigned long long v = 0ull;
            if (row < spanA && w0 + j < nw)
                v = rowbits ? rowbits[row * nw + w0 + j] : colbits[(uint64_t)(w0 + j) * rb + row];
            s_rw[rr][j] = v;
        }
        __syncthreads();
        unsigned long long cb[8];
#pragma unroll
        for (int j = 0; j < 8; j++) cb[j] = w0 + j < nw ? colbits[(uint64_t)(w0 + j) * rb + c] : 0ull;
#pragma unroll 4
        for (int i = 0; i < 16; i++) {
            unsigned long long acc = 0;
#pragma unroll
            for (int j = 0; j < 8; j++) acc |= s_rw[lr0 + i][j] & cb[j];
            if (acc) hit |= 1u << i;
        }
    }
    const uint64_t r0 = rb0 + lr0;
    const uint64_t pw = rb / 32;
    for (int i = 0; i < 16; i++) {
        const unsigned long long bal = __ballot((hit >> i) & 1u);
        if (lane == 0) {
            uint32_t *pr = P + (r0 + i) * pw + (uint64_t)blockIdx.x * 2;
            pr[0] = (uint32_t)bal;
            pr[1] = (uint32_t)(bal >> 32);
        }
    }
    if (sym && blockIdx.y != blockIdx.x) {
        // mirror row c: its bits for columns rb0 .. rb0 + 63 are this block's
        // column c, wavefront w holding rows rb0 + 16 w .. + 15
        s_h[lane][wave] = hit;
        __syncthreads();
        if (wave == 0) {
            uint32_t *pr = P + c * pw + rb0 / 32;
            pr[0] = s_h[lane][0] | (s_h[lane][1] << 16);
            pr[1] = s_h[lane][2] | (s_h[lane][3] << 16);
        }
    }
}

// From the left join (before the left tiles exist): row bitsets and the
// left panel, one thread per left entry.
__global__ void k_dk_join_A(const uint32_t *fidA, const uint32_t *rowA, const double *vals, size_t n,
                            const uint32_t *hkey, uint32_t nw, uint64_t ra, unsigned long long *rowbits, double *AT) {
    const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (e >= n) return;
    const uint32_t f = fidA[e];
    if (f == 0xFFFFFFFFu) return;
    const uint32_t h = hkey[f];
    if (h == 0xFFFFFFFFu) return;
    const uint32_t row = rowA[e];
    atomicOr(&rowbits[(uint64_t)row * nw + (h >> 6)], 1ull << (h & 63));
    atomicAdd(&AT[(uint64_t)h * ra + row], vals[e]);
}

__global__ void k_dk_scatter_A(uint32_t ngroups, const uint32_t *group_ptr, const uint32_t *group_fiber,
                               const uint32_t *hkey, const uint32_t *Arow, const double *Aval, uint64_t ra,
                               double *AT) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const uint32_t h = hkey[group_fiber[g]];
    if (h == 0xFFFFFFFFu) return;
    double *col = AT + (uint64_t)h * ra;
    for (uint32_t e = group_ptr[g]; e < group_ptr[g + 1]; e++) atomicAdd(&col[Arow[e]], Aval[e]);
}

__global__ __launch_bounds__(256) void k_dk_scatter_B(const uint32_t *hfib, const uint32_t *fiber_ptr,
                                                     const uint32_t *Bcol, const double *Bval, uint64_t rb,
                                                     double *Bd) {
    const uint32_t h = blockIdx.x, f = hfib[h];
    double *row = Bd + (uint64_t)h * rb;
    const uint32_t step = gridDim.y * blockDim.x;
    for (uint32_t p = fiber_ptr[f] + blockIdx.y * blockDim.x + threadIdx.x; p < fiber_ptr[f + 1]; p += step)
        atomicAdd(&row[Bcol[p]], Bval[p]);
}

// C = AT^T * Bd. Workgroup tile 128 x 128, four wavefronts of 64 x 64, each a
// 4 x 4 grid of v_mfma_f64_16x16x4 blocks (16 f64x4 accumulators). The k
// loop steps 16 keys at a time through double-buffered LDS panels
// sA[k][row], sB[k][col]; the next panels stream from global memory straight
// into LDS while the current ones feed the matrix cores, one barrier per step.
// MFMA operand layout (16x16x4, f64): lane l holds A[i = l % 16][k = l / 16]
// and B[k = l / 16][j = l % 16]; the result block D[i][j] sits in lane
// j + 16 * (i % 4), element i / 4.
// LDS banks: an operand read takes 16 consecutive doubles from each of four
// k rows; k rows are DK_LDS_ROW = 144 doubles apart (128 B past a multiple of
// 256 B), so two rows read in the same cycle fall on disjoint banks. A panel
// load writes one whole k row (1 KiB, 16 bytes per lane) per instruction.
constexpr uint32_t DK_LDS_ROW = DK_BM + 16;
static_assert(DK_BM == DK_BN, "one padded LDS row layout for both panels");

__global__ __launch_bounds__(256, 2) void k_dgemm_tn(const double *__restrict__ AT, const double *__restrict__ Bd,
                                                     double *__restrict__ C, uint32_t nbm, uint32_t nbn, uint64_t ra,
                                                     uint64_t rb, uint32_t kp, uint32_t splits, bool sym,
                                                     bool upper) {
    // two buffers as four distinct LDS objects: the compiler then knows that
    // reads of one buffer do not depend on panel loads still landing in the
    // other, and waits for those loads only at the step's end
    __shared__ __align__(16) double sA0[DK_BK][DK_LDS_ROW], sA1[DK_BK][DK_LDS_ROW];
    __shared__ __align__(16) double sB0[DK_BK][DK_LDS_ROW], sB1[DK_BK][DK_LDS_ROW];
    // Workgroups are dealt round-robin over the 8 XCDs (one L2 each): give
    // every XCD a contiguous run of logical tiles, and walk logical tiles down
    // bands of 8 row tiles so neighbours share AT rows and a Bd column panel.
    // With split k (small grids) workgroup z of a tile sums k steps
    // [z * nk / splits, (z + 1) * nk / splits) into its own slab of C.
    // sym (AT is Bd, so C is symmetric): only tiles tm <= tn, each band
    // holding the columns from its own first row on; an off-diagonal tile is
    // also stored transposed.
    const uint32_t band = 8;
    const uint32_t nblk = sym ? nbm * (nbm + 1) / 2 : nbm * nbn, total = nblk * splits;
    const uint32_t x = blockIdx.x & 7, jx = blockIdx.x >> 3;
    const uint32_t q = total >> 3, rem = total & 7;
    const uint32_t Lz = x * q + min(x, rem) + jx;
    const uint32_t z = Lz / nblk, L = Lz - z * nblk;
    uint32_t tm, tn;
    if (sym) {
        // band r0 / band: R rows, columns r0 .. nbm - 1, column r0 + j holding
        // min(R, j + 1) of them (the first R columns a triangle)
        uint32_t r0 = 0, w = L, R = min(band, nbm);
        for (;;) {
            const uint32_t cnt = R * (R + 1) / 2 + (nbm - r0 - R) * R;
            if (w < cnt) break;
            w -= cnt;
            r0 += band;
            R = min(band, nbm - r0);
        }
        const uint32_t tri = R * (R + 1) / 2;
        if (w < tri) {
            uint32_t j = 0;
            while ((j + 1) * (j + 2) / 2 <= w) j++;
            tm = r0 + w - j * (j + 1) / 2;
            tn = r0 + j;
        } else {
            w -= tri;
            tm = r0 + w % R;
            tn = r0 + R + w / R;
        }
    } else {
        const uint32_t per_band = band * nbn;
        const uint32_t bnd = L / per_band, w = L - bnd * per_band;
        const uint32_t rows_in = min(band, nbm - bnd * band);
        tm = bnd * band + w % rows_in;
        tn = w / rows_in;
    }

    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const uint32_t nk = kp / DK_BK;
    const uint32_t kb0 = (uint32_t)((uint64_t)z * nk / splits), kb1 = (uint32_t)((uint64_t)(z + 1) * nk / splits);
    C += (uint64_t)z * ra * rb;
    f64x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
    // Panels go from global memory straight into LDS (global_load_lds_dwordx4:
    // lane l's 16 bytes land at byte 16 l of the row), one k step ahead of
    // their use: one instruction moves a k row of 128 doubles, each wavefront
    // moves 4 rows of each panel. No staging registers, and the loads of step
    // k + 1 complete under step k's matrix ops.
    const double *gA = AT + (uint64_t)tm * DK_BM + 2 * lane;
    const double *gB = Bd + (uint64_t)tn * DK_BN + 2 * lane;
    using Panel = double[DK_BK][DK_LDS_ROW];
    auto issue_panel = [&](uint32_t k, Panel &pA, Panel &pB) {
        const uint64_t k0 = (uint64_t)(kb0 + k) * DK_BK;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t row = wave * 4 + r;
            __builtin_amdgcn_global_load_lds((const void *)(gA + (k0 + row) * ra),
                                             (__attribute__((address_space(3))) void *)&pA[row][0], 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const void *)(gB + (k0 + row) * rb),
                                             (__attribute__((address_space(3))) void *)&pB[row][0], 16, 0, 0);
        }
    };
    const uint32_t kr = lane >> 4, ci = lane & 15;
    auto compute = [&](const Panel &pA, const Panel &pB) {
#pragma unroll
        for (uint32_t kk = 0; kk < DK_BK; kk += 4) {
            double a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = pA[kk + kr][wm + i * 16 + ci];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = pB[kk + kr][wn + j * 16 + ci];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    };
    // this wavefront's panel loads have landed, then every wavefront's
    auto land = [] {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };
    // Step k reads buffer k % 2 while step k + 1's panels load into the other,
    // whose last readers passed the previous step's barrier. Two steps per
    // trip keep each buffer a fixed object.
    const uint32_t nks = kb1 - kb0;
    if (nks) issue_panel(0, sA0, sB0);
    land();
    uint32_t kb = 0;
    for (; kb + 2 <= nks; kb += 2) {
        issue_panel(kb + 1, sA1, sB1);
        compute(sA0, sB0);
        land();
        if (kb + 2 < nks) issue_panel(kb + 2, sA0, sB0);
        compute(sA1, sB1);
        land();
    }
    if (kb < nks) compute(sA0, sB0);
    const uint64_t row0 = (uint64_t)tm * DK_BM + wm + kr, col0 = (uint64_t)tn * DK_BN + wn + ci;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) C[(row0 + i * 16 + 4 * r) * rb + col0 + j * 16] = acc[i][j][r];
    if (sym && !upper && tm != tn) {
        // The mirror tile (ra == rb), transposed through LDS so that every
        // store writes whole 512-byte column runs: each wavefront takes one
        // of the four panel buffers (free once all have passed the barrier)
        // for 32 of its 64 columns at a time, columns TS doubles apart.
        constexpr uint32_t TS = 66;
        static_assert(32 * TS <= DK_BK * DK_LDS_ROW, "a mirror half fits one panel buffer");
        __syncthreads();
        double *T = wave == 0 ? &sA0[0][0] : wave == 1 ? &sA1[0][0] : wave == 2 ? &sB0[0][0] : &sB1[0][0];
        const uint64_t mrow0 = (uint64_t)tn * DK_BN + wn, mcol0 = (uint64_t)tm * DK_BM + wm;
#pragma unroll
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int jj = 0; jj < 2; jj++)
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int r = 0; r < 4; r++) T[(jj * 16 + ci) * TS + i * 16 + kr + 4 * r] = acc[i][2 * h + jj][r];
            __syncthreads();
            const uint32_t rr = (lane & 31) * 2;
#pragma unroll
            for (int it = 0; it < 16; it++) {
                const uint32_t c = 2 * it + (lane >> 5);
                const double2 v = *reinterpret_cast<const double2 *>(&T[c * TS + rr]);
                *reinterpret_cast<double2 *>(&C[(mrow0 + 32 * h + c) * rb + mcol0 + rr]) = v;
            }
            __syncthreads();
        }
    }
}

}  // namespace

void launch_dgemm_tn(const double *AT, const double *Bd, double *C, uint64_t ra, uint64_t rb, uint32_t kp,
                     uint32_t splits, bool sym, bool upper, hipStream_t s) {
    const uint32_t nbm = (uint32_t)(ra / DK_BM), nbn = (uint32_t)(rb / DK_BN);
    if (!nbm || !nbn || !kp) return;
    if (sym && (AT != Bd || ra != rb)) throw LibError{ERR_INVALID, "launch_dgemm_tn: sym needs AT == Bd and ra == rb"};
    const uint32_t nblk = sym ? nbm * (nbm + 1) / 2 : nbm * nbn;
    k_dgemm_tn<<<nblk * splits, 256, 0, s>>>(AT, Bd, C, nbm, nbn, ra, rb, kp, splits, sym, sym && upper);
}

uint32_t dgemm_splits(uint64_t ra, uint64_t rb, uint32_t kp, bool sym, int num_cus) {
    // two workgroups per CU. The GEMM runs beside the left-tile build, and
    // every split is one more partial block each unit reads back (measured:
    // 210 tiles split in two on 256 CUs made the accumulate kernel 13% slower
    // and the GEMM no faster), so k is split only for small blocks, tiles on
    // under a quarter of the CUs; then the split makes the last wave of
    // workgroups nearly full, keeping at least 8 k steps per split
    const uint64_t nbm = ra / DK_BM, nblk = sym ? nbm * (nbm + 1) / 2 : nbm * (rb / DK_BN);
    const uint64_t slots = 2ull * (uint64_t)std::max(1, num_cus);
    const uint32_t nk = kp / DK_BK;
    if (4 * nblk >= (uint64_t)std::max(1, num_cus)) return 1;
    uint32_t best = 1;
    double best_eff = 0;
    for (uint32_t sp = 1; sp <= 8 && nk / sp >= 8; sp++) {
        const uint64_t t = nblk * sp;
        const double eff = (double)t / (double)(div_up(t, slots) * slots);
        if (eff > best_eff + 0.02) {
            best_eff = eff;
            best = sp;
        }
    }
    return best;
}

void dense_keys_select(DevPool &pool, const Geom &g, const Index &ix, const uint32_t *fidA, size_t nA, float tau,
                       uint64_t max_cells, bool require_gain, DenseKeys &dk, std::vector<std::shared_ptr<void>> &hold,
                       hipStream_t s) {
    dk = DenseKeys();
    if (!ix.nfib || tau <= 0.0f) return;
    const uint64_t ra = div_up(g.spanA, DK_BM) * DK_BM, rb = div_up(g.spanB, DK_BN) * DK_BN;
    if (ra * rb > max_cells || ra / 64 > 65535) return;
    auto keep = [&](size_t bytes) -> void * {
        void *p = pool.get(bytes + 16);
        DevPool *pl = &pool;
        hold.emplace_back(p, [pl](void *q) { pl->put(q); });
        return p;
    };
    const uint32_t nfib = ix.nfib;
    Tmp<uint32_t> na(pool, fidA ? nfib : 0), flag(pool, nfib), idx(pool, nfib), dtot(pool, 1);
    Tmp<unsigned long long> tot(pool, 3);
    if (fidA) {
        // nA <= spanA, so a selected key has nB >= tau * spanB
        const double nbm = std::floor((double)tau * (double)g.spanB);
        fill_u32(na, nfib, 0, s);
        if (nA)
            k_dk_na<<<grid_of(nA), 256, 0, s>>>(fidA, nA, ix.fiber_ptr,
                                                 (uint32_t)std::min(nbm, 4294967295.0), na);
    }
    const double inv_area = 1.0 / ((double)g.spanA * (double)g.spanB);
    // at most this many keys, and their panels within 2 GiB
    const uint64_t max_keys = std::min<uint64_t>(16384, (2ull << 30) / (8 * (ra + rb)));
    // the key numbering is queued with the selection, before the host waits
    // for the count: a selection within max_keys needs no second pass
    uint32_t *hkey = (uint32_t *)keep((size_t)nfib * 4);
    uint32_t *hfib = (uint32_t *)keep((size_t)max_keys * 4);
    double t = tau;
    unsigned long long h[3] = {0, 0, 0};
    for (int it = 0; it < 40; it++) {
        fill_u64(tot, 3, 0, s);
        k_dk_flag<<<grid_of(nfib), 256, 0, s>>>(ix.fiber_ptr, fidA ? na.get() : nullptr, nfib, inv_area, t, g.spanA,
                                                 g.spanB, flag, tot);
        scan_u32(pool, flag, idx, nfib, dtot, s);
        k_dk_keys<<<grid_of(nfib), 256, 0, s>>>(flag, idx, nfib, (uint32_t)max_keys, hkey, hfib);
        pool.read(h, tot.get(), 24, s);
        pool.sync(s);
        if (h[0] <= max_keys) break;
        t *= 1.25;
    }
    if (!h[0] || h[0] > max_keys) return;
    if (require_gain && (double)h[1] < 4.0 * (double)g.spanA * (double)g.spanB) return;
    dk.nh = (uint32_t)h[0];
    dk.products = h[1];
    dk.tau = (float)t;
    dk.nw = (dk.nh + 63) / 64;
    dk.kp = (uint32_t)div_up(dk.nh, DK_BK) * DK_BK;
    dk.spanA = g.spanA;
    dk.spanB = g.spanB;
    dk.ra = ra;
    dk.rb = rb;
    dk.self = fidA == nullptr;
    dk.full = h[2] != 0;
    dk.hkey = hkey;
    dk.hfib = hfib;
}

void dense_keys_alloc(DevPool &pool, const Geom &g, DenseKeys &dk, int num_cus,
                      std::vector<std::shared_ptr<void>> &hold) {
    if (!dk.nh) return;
    auto keep = [&](size_t bytes) -> void * {
        void *p = pool.get(bytes + 16);
        DevPool *pl = &pool;
        hold.emplace_back(p, [pl](void *q) { pl->put(q); });
        return p;
    };
    const uint64_t ra = dk.ra, rb = dk.rb;
    dk.colbits = (unsigned long long *)keep((size_t)dk.nw * rb * 8);
    if (!dk.self) dk.rowbits = (unsigned long long *)keep((size_t)g.spanA * dk.nw * 8);
    dk.P = (uint32_t *)keep((size_t)(ra * (rb / 32) + 2) * 4);
    dk.Bd = (double *)keep((size_t)dk.kp * rb * 8);
    dk.AT = dk.self ? dk.Bd : (double *)keep((size_t)dk.kp * ra * 8);  // self: the same panel
    dk.splits = dgemm_splits(ra, rb, dk.kp, dk.self, num_cus);
    dk.C = (double *)keep((size_t)dk.splits * ra * rb * 8);
}

void dense_keys_block(const Index &ix, const DenseKeys &dk, const uint32_t *fidA, const uint32_t *rowA,
                      const double *Avals, size_t nA, hipStream_t s, hipStream_t sp, hipEvent_t gemm0,
                      hipEvent_t gemm1) {
    if (!dk.nh) return;
    const uint64_t ra = dk.ra, rb = dk.rb;
    fill_u32(dk.P + ra * (rb / 32), 2, 0, s);
    fill_f64(dk.Bd, (size_t)dk.kp * rb, 0.0, s);
    if (!dk.full) {
        fill_u64(dk.colbits, (size_t)dk.nw * rb, 0, s);
        k_dk_colbits<<<dim3(dk.nh, DK_FIBER_SPLIT), 256, 0, s>>>(dk.hfib, ix.fiber_ptr, ix.Bcol, rb, dk.colbits);
    }
    k_dk_scatter_B<<<dim3(dk.nh, DK_FIBER_SPLIT), 256, 0, s>>>(dk.hfib, ix.fiber_ptr, ix.Bcol, ix.Bval, rb, dk.Bd);
    if (!dk.self) {
        fill_u64(dk.rowbits, (size_t)dk.spanA * dk.nw, 0, s);
        fill_f64(dk.AT, (size_t)dk.kp * ra, 0.0, s);
        if (nA)
            k_dk_join_A<<<grid_of(nA), 256, 0, s>>>(fidA, rowA, Avals, nA, dk.hkey, dk.nw, ra, dk.rowbits, dk.AT);
    }
    if (dk.full)
        k_dk_full_presence<<<grid_of(ra * (rb / 128)), 256, 0, sp>>>((uint4 *)dk.P, (uint32_t)ra, (uint32_t)(rb / 128),
                                                                     (uint32_t)dk.spanA, (uint32_t)dk.spanB);
    else
        k_dk_presence<<<dim3((unsigned)(rb / 64), (unsigned)(ra / 64)), 256, 0, s>>>(dk.rowbits, dk.colbits, dk.spanA,
                                                                                     rb, dk.nw, dk.P);
    if (gemm0) FCC_HIP(hipEventRecord(gemm0, s));
    launch_dgemm_tn(dk.AT, dk.Bd, dk.C, ra, rb, dk.kp, dk.splits, dk.self, dk.upper, s);
    if (gemm1) FCC_HIP(hipEventRecord(gemm1, s));
}

void dense_keys_values(const Index &ix, const DenseKeys &dk, hipStream_t s) {
    if (!dk.nh) return;
    fill_f64(dk.Bd, (size_t)dk.kp * dk.rb, 0.0, s);
    if (!dk.self) fill_f64(dk.AT, (size_t)dk.kp * dk.ra, 0.0, s);
    if (ix.ngroups && !dk.self)
        k_dk_scatter_A<<<grid_of(ix.ngroups), 256, 0, s>>>(ix.ngroups, ix.group_ptr, ix.group_fiber, dk.hkey, ix.Arow,
                                                            ix.Aval, dk.ra, dk.AT);
    k_dk_scatter_B<<<dim3(dk.nh, DK_FIBER_SPLIT), 256, 0, s>>>(dk.hfib, ix.fiber_ptr, ix.Bcol, ix.Bval, dk.rb, dk.Bd);
    launch_dgemm_tn(dk.AT, dk.Bd, dk.C, dk.ra, dk.rb, dk.kp, dk.splits, dk.self, dk.upper, s);
}

}  // namespace fcc
