// C ABI (include/fcc.h): context, orchestration of one contraction, results.
//
// fcc_contract stands in for Tensor<double>::fastcc_multiply
// (contract.hpp:824-912). Phases:
//   index   extents (contract.hpp:124-144), right CSR by contraction key and
//           left join (replacing both TileIndexedTensor builds,
//           contract.hpp:852-862), left tiles
//   plan    per-tile kind/window choice from the exact bounds or the MLP
//   count   exact planner only: structural count per unit, scan -> offsets
//   numeric accumulate + drain straight into the SoA output (replacing the
//           per-thread TileAccumulator + ListTensor + heaps,
//           contract.hpp:864-911)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <random>
#include <string>
#include <vector>

#include "../../include/fcc.h"
#include "fcc_accum.h"
#include "fcc_batch.h"
#include "fcc_chunk.h"
#include "fcc_dense.h"
#include "fcc_mlp.h"

using namespace fcc;

// Device scratch of fcc_contract_noalloc: every device temporary of such a
// call is carved from it. Each device the library runs on holds its own copy
// until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (fcc_ctx::scratch).
static __device__ __align__(256) unsigned char scratch[FCC_NOALLOC_SCRATCH_BYTES];

struct fcc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // second stream (default priority): the dense-key block beside the left tiles
    DevPool pool;
    std::string err;
    int num_cus = 256;
    size_t mem_total = 0;  // device memory (read once: a driver query per call costs tens of microseconds)
    int clock_khz = 100000;  // wall_clock64 rate
    // learned planner
    std::vector<float> mlp_host;
    float *mlp_dev = nullptr;
    bool mlp_ready = false;
    // autotuner
    std::vector<fcc_tune_entry> tune;
    // statistics of the last collect_stats contraction
    std::vector<float> feat, labels;
    hipEvent_t ev[12] = {};
    void *scratch = nullptr;  // this device's copy of the no-allocation scratch (null: lookup failed)
};

namespace {

// Device storage shared by the per-pair results of one fused batch; returned
// to the pool when the last of them is freed.
struct SharedBlock {
    DevPool *pool = nullptr;
    void *coords = nullptr;
    void *vals = nullptr;
    ~SharedBlock() {
        if (coords) pool->put(coords);
        if (vals) pool->put(vals);
    }
};

struct ResultPriv {
    void *coords = nullptr;
    void *vals = nullptr;
    std::shared_ptr<SharedBlock> block;  // set instead of coords/vals for batch results
};

int fail(fcc_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0.f;
    }
    return ms;
}

// Key modes: nb batch pairs first, then the contraction pairs (ka/kb, r
// entries in all). A's external index leads with the batch coordinates (radix
// = the key radix of their pair), then A's other modes ascending; B's external
// index is B's modes outside kb, ascending. Outputs therefore carry (batch,
// A externals, B externals), the batch-mode order of index_tensors.hpp:322-325.
Geom make_geom(const fcc_tensor *a, const int32_t *ka, const fcc_tensor *b, const int32_t *kb, int r, int nb,
               const uint32_t *EA, const uint32_t *EB) {
    Geom g;
    memset(&g, 0, sizeof(g));
    g.dA = a->order;
    g.dB = b->order;
    g.r = r;
    g.dC = a->order + b->order - 2 * r + nb;
    for (int i = 0; i < r; i++) {
        g.am[i] = ka[i];
        g.bm[i] = kb[i];
        g.R[i] = std::max<uint64_t>(EA[ka[i]], EB[kb[i]]);
    }
    auto in = [](const int32_t *l, int n, int m) {
        for (int i = 0; i < n; i++)
            if (l[i] == m) return true;
        return false;
    };
    // spans saturate at 2^64 - 1 instead of wrapping (an external index
    // span of 2^32 or more is rejected by the caller)
    auto mul_sat = [](uint64_t x, uint64_t y) {
        uint64_t z;
        return __builtin_mul_overflow(x, y, &z) ? ~0ull : z;
    };
    g.spanA = 1;
    g.spanB = 1;
    for (int i = 0; i < nb; i++) {
        g.xa[g.nxa] = ka[i];
        g.xaE[g.nxa] = (uint32_t)g.R[i];
        g.spanA = mul_sat(g.spanA, g.R[i]);
        g.nxa++;
    }
    for (int m = 0; m < a->order; m++)
        if (!in(ka, r, m)) {
            g.xa[g.nxa] = m;
            g.xaE[g.nxa] = EA[m];
            g.spanA = mul_sat(g.spanA, EA[m]);
            g.nxa++;
        }
    for (int m = 0; m < b->order; m++)
        if (!in(kb, r, m)) {
            g.xb[g.nxb] = m;
            g.xbE[g.nxb] = EB[m];
            g.spanB = mul_sat(g.spanB, EB[m]);
            g.nxb++;
        }
    g.K = 1;
    long double K = 1.0L;
    for (int i = 0; i < r; i++) {
        K *= (long double)g.R[i];
        g.K *= g.R[i];
    }
    if (K >= 9.0e18L) g.K = ~0ull;
    g.rbits = std::max(1, ceil_log2(std::min<uint64_t>(g.spanA, 1ull << 63)));
    g.cbits = std::max(1, ceil_log2(std::min<uint64_t>(g.spanB, 1ull << 63)));
    g.kbits = std::max(1, ceil_log2(g.K));
    return g;
}

int validate(const fcc_tensor *t, const char *name, std::string &msg) {
    if (!t) { msg = std::string(name) + " is NULL"; return FCC_ERR_INVALID; }
    if (t->order < 1 || t->order > FCC_MAX_ORDER) { msg = std::string(name) + ": order out of range"; return FCC_ERR_INVALID; }
    if (t->nnz < 0 || t->nnz >= (int64_t)1 << 32) { msg = std::string(name) + ": nnz out of range"; return FCC_ERR_INVALID; }
    if (t->nnz > 0 && (!t->coords || !t->vals)) { msg = std::string(name) + ": NULL coords or vals"; return FCC_ERR_INVALID; }
    return FCC_OK;
}

// Rows per left tile (and dense window) from global statistics, the tuner
// table, or the caller's hint (the reference's tile_size, rounded up to a
// power of two like contract.hpp:844-846).
void choose_tiles(fcc_ctx *ctx, const fcc_options &o, const Geom &g, uint64_t products, uint32_t nvalid,
                  bool dense_keys, int *logTL, uint32_t *Wd, int32_t *bucket_out, int32_t *wg_out) {
    double cells = (double)g.spanA * (double)g.spanB;
    double dens = cells > 0 ? (double)products / cells : 0.0;
    double rows = (double)g.spanA * (1.0 - std::exp(-(double)nvalid / std::max(1.0, (double)g.spanA)));
    rows = std::max(1.0, rows);
    double prow = (double)products / rows;
    uint64_t TL = 1, W = 0;
    int32_t bucket = fcc_density_bucket(dens);
    const fcc_tune_entry *te = nullptr;
    for (auto &e : ctx->tune)
        if (e.bucket == bucket) te = &e;
    *bucket_out = bucket;
    *wg_out = (te && o.tile_hint <= 0 && te->workgroups > 0) ? te->workgroups : 0;
    if (o.tile_hint > 0) {
        TL = pow2_ceil((uint64_t)o.tile_hint);
    } else if (te && te->row_tile > 0) {
        TL = (uint64_t)te->row_tile;
    } else if (dense_keys) {
        // the dense keys' products are gone: what stays is many short
        // fibers, and every column window rescans the tile's entries: whole
        // rows in one window, as many rows as fit while the tiles still number
        // at least twice the accumulate kernels' workgroup slots (every unit
        // is a dense block-seeded tile: fewer, larger ones leave the last
        // workgroups waiting); else two-row tiles with the widest windows
        const uint64_t sp = (g.spanB + 31) & ~31ull;
        const uint64_t min_tiles = 4ull * (uint64_t)std::max(1, ctx->num_cus);
        TL = 2;
        while (TL * 2 * sp <= DENSE_SLOTS && TL < 16 && g.spanA / (TL * 2) >= min_tiles) TL *= 2;
    } else if (dens >= 0.25) {
        const uint64_t sp = (g.spanB + 31) & ~31ull;  // dense rows are padded to 32 columns
        uint64_t tf = 1;
        while (tf * 2 * sp <= DENSE_SLOTS) tf *= 2;
        // rows fit whole: as many rows as fit (16..64); else column windows,
        // fewer rows when the rows are long (each window rescans the tile's
        // groups, so wide outputs favour fewer, wider windows)
        if (tf * sp <= DENSE_SLOTS && tf >= 16) TL = std::min<uint64_t>(tf, 64);
        else TL = g.spanB > 8192 ? 8 : 16;
    } else {
        double t = (double)HASH_TARGET / std::max(1.0, prow);
        TL = 1;
        while ((double)(TL * 2) <= t && TL < 1024) TL *= 2;
    }
    if (o.col_window > 0) W = (uint64_t)o.col_window;
    else if (te && te->col_window > 0) W = (uint64_t)te->col_window;
    TL = std::min<uint64_t>(TL, 4096);  // the dense LDS tile must hold at least one column per row
    int lt = 0;
    while ((1ull << lt) < TL) lt++;
    lt = std::min(lt, g.rbits);
    TL = 1ull << lt;
    if (W == 0) {
        // the widest window the LDS tile holds (a multiple of 32 columns):
        // fewer windows per tile, so fewer window-clipped fiber segments
        // (order-5 shape, two-row tiles: 4096 -> 4992 columns, accumulate
        // 9.96 -> 8.98 ms)
        W = std::max<uint64_t>(32, (DENSE_SLOTS / TL) & ~31ull);
        // whole (padded) rows in one window when they fit
        const uint64_t sp = (g.spanB + 31) & ~31ull;
        if (sp <= W) W = sp;
    }
    W = (W + 31) & ~31ull;  // dense windows cover whole 32-column bitmap words
    while (W > 32 && W * TL > DENSE_SLOTS) W -= 32;
    *logTL = lt;
    *Wd = (uint32_t)W;
}

__global__ void k_sample_rows(const uint32_t *rowA, uint32_t *fidA, size_t n, int logT, uint32_t stride) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n && ((rowA[i] >> logT) % stride) != 0) fidA[i] = 0xFFFFFFFFu;
}

__global__ void k_count_valid(const uint32_t *fidA, const uint32_t *fiber_ptr, size_t n, unsigned long long *out) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    unsigned long long c = 0, p = 0;
    if (i < n && fidA[i] != 0xFFFFFFFFu) {
        c = 1;
        p = fiber_ptr[fidA[i] + 1] - fiber_ptr[fidA[i]];
    }
    for (int d = 32; d > 0; d >>= 1) {
        c += __shfl_xor(c, d, 64);
        p += __shfl_xor(p, d, 64);
    }
    if ((threadIdx.x & 63) == 0 && c) {
        atomicAdd(&out[0], c);
        atomicAdd(&out[1], p);
    }
}

// Per-tile output counts from per-unit counts; a unit of DIRECT tiles may
// span several tiles, so DIRECT tiles take their products instead.
__global__ void k_unit_to_tile(const unsigned long long *unit_count, const uint32_t *unit_tile,
                               const uint8_t *unit_kind, uint32_t n, unsigned long long *tile_count) {
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n && unit_kind[u] != KIND_DIRECT && unit_count[u]) atomicAdd(&tile_count[unit_tile[u]], unit_count[u]);
}

__global__ void k_direct_tile_count(const uint8_t *st_direct, const uint64_t *st_outd, uint32_t T, int force,
                                    unsigned long long *tile_count) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < T && !force && st_direct[a]) tile_count[a] = st_outd[a];
}

__global__ void k_kind_hist(const uint8_t *kind, uint32_t n, unsigned long long *h) {
    uint32_t c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) c[kind[u] & 7]++;
    for (int k = 0; k < 8; k++) {
        uint32_t v = c[k];
        for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
        if ((threadIdx.x & 63) == 0 && v) atomicAdd(&h[k], (unsigned long long)v);
    }
}

// Distinct output columns per active tile = |union of its groups' fiber
// columns| (training label). LDS bitmap when the column span allows, else an
// LDS hash set that saturates at its capacity.
constexpr uint32_t COLSET = 16384;
__global__ __launch_bounds__(256) void k_tile_cols(const uint32_t *atile_g0, const uint32_t *group_fiber,
                                                   const uint32_t *fiber_ptr, const uint32_t *Bcol, uint32_t natiles,
                                                   uint64_t spanB, unsigned long long *tile_cols) {
    __shared__ uint32_t set[COLSET];
    __shared__ uint32_t cnt;
    const bool bitmap = spanB <= (uint64_t)COLSET * 32;
    for (uint32_t a = blockIdx.x; a < natiles; a += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < COLSET; i += blockDim.x) set[i] = bitmap ? 0u : 0xFFFFFFFFu;
        if (threadIdx.x == 0) cnt = 0;
        __syncthreads();
        for (uint32_t g = atile_g0[a]; g < atile_g0[a + 1]; g++) {
            uint32_t f = group_fiber[g];
            for (uint32_t p = fiber_ptr[f] + threadIdx.x; p < fiber_ptr[f + 1]; p += blockDim.x) {
                uint32_t c = Bcol[p];
                if (bitmap) {
                    atomicOr(&set[c >> 5], 1u << (c & 31));
                } else {
                    uint32_t h = (c * 2654435761u) >> 18;
                    for (uint32_t pr = 0; pr < COLSET; pr++) {
                        if (cnt >= COLSET * 3 / 4) break;
                        uint32_t old = atomicCAS(&set[h], 0xFFFFFFFFu, c);
                        if (old == 0xFFFFFFFFu) { atomicAdd(&cnt, 1u); break; }
                        if (old == c) break;
                        h = (h + 1) & (COLSET - 1);
                    }
                }
            }
        }
        __syncthreads();
        if (bitmap) {
            uint32_t part = 0;
            for (uint32_t i = threadIdx.x; i < COLSET; i += blockDim.x) part += __popc(set[i]);
            atomicAdd(&cnt, part);
        }
        __syncthreads();
        if (threadIdx.x == 0) tile_cols[a] = cnt;
        __syncthreads();
    }
}

// Training labels: gaps (<= 0) between the log1p of the exact tile output
// nnz / distinct columns and their upper bounds taken from the features.
__global__ void k_labels(const unsigned long long *tile_nnz, const unsigned long long *tile_cols, const float *feat,
                         uint32_t n, float *labels) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const float *f = feat + (size_t)a * MLP_IN;
    labels[2 * a] = fminf(log1pf((float)tile_nnz[a]) - fminf(f[3], f[5]), 0.f);
    labels[2 * a + 1] = fminf(log1pf((float)tile_cols[a]) - f[2], 0.f);
}

// MAPE of predicted vs actual tile nnz: sum |p - a| / max(a, 1)
// Grid-stride over the tiles (a capped grid): one atomic per workgroup.
__global__ __launch_bounds__(256) void k_mape(const float *pred, const unsigned long long *actual, uint32_t n,
                                              double *acc) {
    __shared__ double s_e[4];
    double e = 0.0;
    for (uint32_t a = blockIdx.x * 256 + threadIdx.x; a < n; a += gridDim.x * 256) {
        const double act = (double)actual[a];
        e += fabs((double)pred[a] - act) / (act > 1.0 ? act : 1.0);
    }
    for (int d = 32; d > 0; d >>= 1) e += __shfl_xor(e, d, 64);
    if ((threadIdx.x & 63) == 0) s_e[threadIdx.x >> 6] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double t = s_e[0] + s_e[1] + s_e[2] + s_e[3];
        if (t != 0.0) unsafeAtomicAdd(acc, t);
    }
}

// The same over every S-th tile (acc[1] counts them).
__global__ void k_mape_sampled(const float *pred, const unsigned long long *actual, uint32_t n, uint32_t S,
                               double *acc) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0, c = 0.0;
    if (a < n && a % S == 0) {
        const double act = (double)actual[a];
        e = fabs((double)pred[a] - act) / (act > 1.0 ? act : 1.0);
        c = 1.0;
    }
    for (int d = 32; d > 0; d >>= 1) {
        e += __shfl_xor(e, d, 64);
        c += __shfl_xor(c, d, 64);
    }
    if ((threadIdx.x & 63) == 0 && c != 0.0) {
        unsafeAtomicAdd(&acc[0], e);
        unsafeAtomicAdd(&acc[1], c);
    }
}

// Keeps the windows of every S-th tile (DIRECT tiles excepted: their count is
// their products), clears the rest: the units of a sampled exact count.
__global__ void k_sample_tiles(uint32_t *t_nwin, const uint8_t *t_kind, uint32_t n, uint32_t S) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n && (a % S != 0 || t_kind[a] == KIND_DIRECT)) t_nwin[a] = 0;
}

// Sum of predicted tile outputs, each capped by its exact bounds (products, area).
__global__ __launch_bounds__(256) void k_pred_sum(const float *pred, const uint64_t *st_products, const uint64_t *st_maxgp,
                           const uint8_t *st_direct, const uint64_t *st_outd, double area, uint32_t n,
                           unsigned long long *sum) {
    __shared__ unsigned long long s_v[4];
    unsigned long long v = 0ull;
    for (uint32_t a = blockIdx.x * 256 + threadIdx.x; a < n; a += gridDim.x * 256) {
        if (st_direct[a]) {
            v += st_outd[a];  // DIRECT tiles: exactly their outputs
        } else {
            double p = ceil((double)pred[a]);
            p = p < (double)st_maxgp[a] ? (double)st_maxgp[a] : p;
            p = p > (double)st_products[a] ? (double)st_products[a] : p;
            p = p > area ? area : p;
            v += (unsigned long long)p;
        }
    }
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) s_v[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned long long t = s_v[0] + s_v[1] + s_v[2] + s_v[3];
        if (t) atomicAdd(sum, t);
    }
}

__global__ void k_summarize(const uint32_t *coords, const double *vals, uint64_t n, uint64_t stride, int d,
                            const uint32_t *E, unsigned long long *hsum, double *vs) {
    uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    unsigned long long h = 0;
    double s = 0.0, w = 0.0;
    if (i < n) {
        uint64_t lin = 0;
        for (int m = 0; m < d; m++) lin = lin * E[m] + coords[(uint64_t)m * stride + i];
        uint64_t z = lin + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z = z ^ (z >> 31);
        h = z;
        s = vals[i];
        w = vals[i] * ((double)(z >> 11) * 0x1.0p-53);
    }
    for (int dd = 32; dd > 0; dd >>= 1) {
        h += __shfl_xor(h, dd, 64);
        s += __shfl_xor(s, dd, 64);
        w += __shfl_xor(w, dd, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(hsum, h);
        unsafeAtomicAdd(&vs[0], s);
        unsafeAtomicAdd(&vs[1], w);
    }
}

unsigned grid256(size_t n) { return (unsigned)div_up(n ? n : 1, 256); }

struct Contraction {
    fcc_ctx *ctx;
    hipStream_t s;
    fcc_options o;
    fcc_result *out;
    // fused batch: the leading batch coordinate is a group index in [0, groups)
    // with extent groups; summaries and completion times are kept per group
    uint32_t groups = 0;
    int32_t bucket = 0;    // density bucket of the last run (autotuner key)
    int32_t tuned_wg = 0;  // workgroups from the tuner table, 0 = default
    std::vector<unsigned long long> group_sink;  // 4 per group: count, hash, val_sum, val_wsum bits
    std::vector<double> group_ms;                // per group: call start -> output complete
    // plans: keep what the numeric pass needs after run(), keep the sort
    // permutations (values can be gathered again) and use explicit-bitmap
    // dense accumulators (later values may hold zeros)
    bool keep = false;
    // caller-owned outputs (the no-allocation entry point): SoA coordinates
    // with row stride user_cap and user_cap values; null = library-owned
    uint32_t *user_c = nullptr;
    double *user_v = nullptr;
    uint64_t user_cap = 0;
    struct Kept {
        std::vector<std::shared_ptr<void>> hold;
        Index ix;
        Geom g;
        Units U;
        Tmp<uint32_t> u_tile, u_c0, u_W;
        Tmp<uint8_t> u_kind, u_multi;
        Tmp<unsigned long long> unit_count, unit_off, unit_cur;
        Tmp<unsigned int> work;
        Tmp<unsigned long long> gmeta, sinkbuf;
        Tmp<uint32_t> flags;
        Tmp<unsigned int> g_left;
        Tmp<unsigned long long> g_done;
        Tmp<unsigned long long> gk;
        Tmp<double> gv;
        Tmp<uint32_t> gu;
        Tmp<uint32_t> out_c;
        Tmp<double> out_v;
        Tmp<float> pred;  // per-tile MLP predictions (learned planner)
        Tmp<unsigned long long> tile_band;  // mirrored plans: exact outputs of each tile's band
        Tmp<uint32_t> fcut;  // window cuts of every fiber (dense column windows)
        DenseKeys dk;        // dense-key block (dk.nh == 0: none)
        AccArgs X;
        unsigned grid = 1;
        bool bitmap = false, sink = false, empty = true;
        size_t nsink = 4;
        uint64_t stride = 0, nnz_out = 0;
        int mode = 0;  // placement of the last numeric pass: 0 exact offsets, 1 atomic cursor
        int plan_force = 0;  // forced accumulator kind of the plan (0: planner's choice)
    } k;

    // ka/kb: r key pairs, the first nb of them batch pairs
    int run(const fcc_tensor *a, const int32_t *ka, const fcc_tensor *b, const int32_t *kb, int r, int nb);
};

int Contraction::run(const fcc_tensor *ta, const int32_t *am, const fcc_tensor *tb, const int32_t *bm, int r,
                     int nb) {
    DevPool &pool = ctx->pool;
    pool.discard();  // read-backs left queued by a call that failed
    hipEvent_t *ev = ctx->ev;
    FCC_HIP(hipEventRecord(ev[0], s));

    // ---- operands on the device
    Tmp<uint32_t> ca, cb;
    Tmp<double> va, vb;
    Operand A{ta->order, ta->nnz, ta->coords, ta->vals}, B{tb->order, tb->nnz, tb->coords, tb->vals};
    if (!ta->on_device && ta->nnz) {
        ca = Tmp<uint32_t>(pool, (size_t)ta->order * ta->nnz);
        va = Tmp<double>(pool, (size_t)ta->nnz);
        FCC_HIP(hipMemcpyAsync(ca.get(), ta->coords, 4ull * ta->order * ta->nnz, hipMemcpyHostToDevice, s));
        FCC_HIP(hipMemcpyAsync(va.get(), ta->vals, 8ull * ta->nnz, hipMemcpyHostToDevice, s));
        A.coords = ca;
        A.vals = va;
    }
    if (!tb->on_device && tb->nnz) {
        if (tb == ta || (tb->coords == ta->coords && tb->vals == ta->vals && tb->order == ta->order &&
                         tb->nnz == ta->nnz && !ta->on_device)) {
            B.coords = A.coords;
            B.vals = A.vals;
        } else {
            cb = Tmp<uint32_t>(pool, (size_t)tb->order * tb->nnz);
            vb = Tmp<double>(pool, (size_t)tb->nnz);
            FCC_HIP(hipMemcpyAsync(cb.get(), tb->coords, 4ull * tb->order * tb->nnz, hipMemcpyHostToDevice, s));
            FCC_HIP(hipMemcpyAsync(vb.get(), tb->vals, 8ull * tb->nnz, hipMemcpyHostToDevice, s));
            B.coords = cb;
            B.vals = vb;
        }
    }

    // ---- extents (contract.hpp:124-144) and the zero/tiny value flag
    Tmp<uint32_t> dext(pool, 2 * (kMaxOrder + 1));
    fill_u32(dext, 2 * (kMaxOrder + 1), 0, s);
    // one operand on both sides (a self-contraction) is scanned once
    const bool same_operand = A.coords == B.coords && A.vals == B.vals && A.order == B.order && A.nnz == B.nnz;
    launch_extents(A, dext.get(), s);
    if (!same_operand) launch_extents(B, dext.get() + kMaxOrder + 1, s);
    uint32_t hext[2 * (kMaxOrder + 1)];
    pool.read(hext, dext.get(), sizeof(hext), s);
    pool.sync(s);
    if (same_operand) memcpy(hext + kMaxOrder + 1, hext, (kMaxOrder + 1) * sizeof(uint32_t));
    uint32_t EA[kMaxOrder], EB[kMaxOrder];
    for (int m = 0; m < ta->order; m++) EA[m] = ta->extents ? ta->extents[m] : (ta->nnz ? hext[m] + 1 : 0);
    for (int m = 0; m < tb->order; m++)
        EB[m] = tb->extents ? tb->extents[m] : (tb->nnz ? hext[kMaxOrder + 1 + m] + 1 : 0);
    const uint32_t vflags = hext[ta->order] | hext[kMaxOrder + 1 + tb->order];
    const bool bitmap = keep || (vflags & 1u) != 0;
    const bool nonfinite = (vflags & 2u) != 0;
    k.bitmap = bitmap;
    if (ta->extents)
        for (int m = 0; m < ta->order; m++)
            if (ta->nnz && hext[m] >= EA[m]) throw LibError{FCC_ERR_INVALID, "a: coordinate exceeds given extent"};
    if (tb->extents)
        for (int m = 0; m < tb->order; m++)
            if (tb->nnz && hext[kMaxOrder + 1 + m] >= EB[m])
                throw LibError{FCC_ERR_INVALID, "b: coordinate exceeds given extent"};

    Geom g = make_geom(ta, am, tb, bm, r, nb, EA, EB);
    k.g = g;
    out->order = g.dC;
    for (int m = 0; m < g.nxa; m++) out->extents[m] = g.xaE[m];
    for (int m = 0; m < g.nxb; m++) out->extents[g.nxa + m] = g.xbE[m];
    fcc_stats &st = out->stats;
    st.mlp_mape = -1.0;
    if (ta->nnz == 0 || tb->nnz == 0 || g.spanA == 0 || g.spanB == 0) {
        FCC_HIP(hipEventRecord(ev[1], s));
        FCC_HIP(hipEventSynchronize(ev[1]));
        st.ms_total = ev_ms(ev[0], ev[1]);
        return FCC_OK;
    }
    if (g.spanA >= (1ull << 32) || g.spanB >= (1ull << 32))
        throw LibError{FCC_ERR_INVALID, "external index span must stay below 2^32"};
    if (g.K == ~0ull || g.kbits + g.cbits > 64)
        throw LibError{FCC_ERR_INVALID, "contraction key and column exceed 64 bits"};

    // ---- index
    auto &hold = k.hold;
    Index &ix = k.ix;
    ix.keep_perm = keep;
    Tmp<uint32_t> rowA(pool, (size_t)A.nnz), fidA(pool, (size_t)A.nnz);
    uint64_t products = 0;
    uint32_t nvalid = 0;
    // self-contraction shortcut: A's external index is B's column index
    bool self = nb == 0 && A.coords == B.coords && A.vals == B.vals && ta->order == tb->order && ta->nnz == tb->nnz;
    for (int i = 0; i < r && self; i++) self = am[i] == bm[i];
    for (int m = 0; m < ta->order && self; m++) self = EA[m] == EB[m];
    Tmp<uint32_t> vsB;
    // the per-entry row and fiber of a self-contraction are only read when
    // left rows are sampled (the tiles come from the right operand's order)
    const bool need_join = o.left_tile_stride > 1;
    SelfJoin sj{need_join ? rowA.get() : nullptr, need_join ? fidA.get() : nullptr, nullptr, nullptr, 0};
    if (self) {
        vsB = Tmp<uint32_t>(pool, (size_t)A.nnz);
        sj.vsB = vsB;
        build_right(pool, g, B, ix, hold, s, &sj);
        products = sj.products;
        nvalid = (uint32_t)A.nnz;
    } else {
        build_right(pool, g, B, ix, hold, s);
        join_left(pool, g, A, ix, rowA, fidA, &products, &nvalid, s);
    }
    if (o.left_tile_stride > 1) {
        uint64_t T = pow2_ceil(o.tile_hint > 0 ? (uint64_t)o.tile_hint : 128);
        int lT = ceil_log2(T);
        k_sample_rows<<<grid256(A.nnz), 256, 0, s>>>(rowA, fidA, (size_t)A.nnz, lT, (uint32_t)o.left_tile_stride);
        Tmp<unsigned long long> cv(pool, 2);
        fill_u64(cv, 2, 0, s);
        k_count_valid<<<grid256(A.nnz), 256, 0, s>>>(fidA, ix.fiber_ptr, (size_t)A.nnz, cv);
        unsigned long long h[2];
        pool.read(h, cv.get(), 16, s);
        pool.sync(s);
        nvalid = (uint32_t)h[0];
        products = h[1];
    }
    st.products = products;

    // ---- dense keys: keys dense on both sides leave the tile traversal and
    // are contracted by one FP64 GEMM; units start from its block
    DenseKeys &dk = k.dk;
    dk = DenseKeys();
    // (not with infinite or NaN values: the GEMM's zero padding would spread them)
    if (!groups && o.dense_keys >= 0.0f && products && !nonfinite) {
        // the block and its panels within an eighth of device memory
        const uint64_t max_cells = std::min<uint64_t>(1ull << 29, ctx->mem_total / 64);
        // the default threshold also asks the block to pay for itself; an
        // explicit one is taken as given
        const bool dflt = o.dense_keys == 0.0f;
        const bool self_fib = self && o.left_tile_stride <= 1;  // left fibers = right fibers
        dense_keys_select(pool, g, ix, self_fib ? nullptr : fidA.get(), (size_t)A.nnz, dflt ? 0.02f : o.dense_keys,
                          max_cells, dflt, dk, hold, s);
    }
    // The block depends only on the right fibers and the join: it runs on
    // the side stream beside the left-tile build, forked once the left sort
    // is queued (the sort and the GEMM both want the LDS; what follows the
    // sort -- gathers, groups, statistics -- is latency-bound and shares the
    // CUs well). Its buffers are all allocated before the fork; the join
    // arrays live until the join below (side_join synchronises the side
    // stream on any early exit first).
    struct SideJoin {
        hipStream_t st = nullptr;
        ~SideJoin() {
            if (st) (void)hipStreamSynchronize(st);
        }
    } side_join;
    // dense column windows: every fiber's window cuts, once, instead of a
    // binary search per group per window unit (bounded table size). They
    // depend on the right index alone; with a dense-key block they are queued
    // right after its fork, so the block starts as early as it can (the
    // accumulate pass waits for the block's GEMM anyway).
    uint32_t fcut_n = 0, fcut_W = 0;
    auto make_fiber_cuts = [&](uint32_t Wd) {
        if (g.spanB <= Wd) return;
        const uint64_t nw = (g.spanB + Wd - 1) / Wd + 1;
        if ((uint64_t)ix.nfib * nw > (1ull << 28)) return;
        k.fcut = Tmp<uint32_t>(pool, (size_t)ix.nfib * nw);
        fcut_n = (uint32_t)nw;
    };
    std::function<void()> fork_dense = [&]() {
        dense_keys_alloc(pool, g, dk, ctx->num_cus, hold);
        FCC_HIP(hipEventRecord(ev[9], s));
        FCC_HIP(hipStreamWaitEvent(ctx->side, ev[9], 0));
        side_join.st = ctx->side;
        FCC_HIP(hipEventRecord(ev[7], ctx->side));
        dense_keys_block(ix, dk, fidA.get(), rowA.get(), A.vals, (size_t)A.nnz, ctx->side, s, ev[10], ev[11]);
        FCC_HIP(hipEventRecord(ev[8], ctx->side));
        if (k.fcut.get()) launch_fiber_cuts(ix, fcut_n, fcut_W, k.fcut, s);
    };
    if (!dk.nh)
        for (int e : {7, 8, 10, 11}) FCC_HIP(hipEventRecord(ev[e], s));
    int logTL;
    uint32_t Wd;
    choose_tiles(ctx, o, g, products, nvalid, dk.nh > 0, &logTL, &Wd, &bucket, &tuned_wg);
    make_fiber_cuts(Wd);
    fcut_W = Wd;
    if (k.fcut.get() && !dk.nh) launch_fiber_cuts(ix, fcut_n, Wd, k.fcut, s);
    // mirror mode: one tensor, the same mode lists, every left tile (a
    // sample of rows has no symmetric output) and no statistics collection
    // (its labels are whole-tile output counts)
    const bool mirror = self && o.left_tile_stride <= 1 && !o.collect_stats && o.mirror >= 0 && !groups;
    ix.mirror = mirror;
    dk.upper = mirror && dk.self;  // mirror mode reads the block on and above its diagonal only
    if (const char *hm = getenv("FCC_HEAVY_MIN")) ix.heavy_min = (uint32_t)std::max(1, atoi(hm));
    if (const char *hs = getenv("FCC_HSEG")) ix.hseg = (uint32_t)std::max(64, atoi(hs));
    ix.dk_hkey = dk.nh ? dk.hkey : nullptr;  // the left build takes these keys out of the tiles
    ix.seg_below = 2 * (ACC_THREADS / 64);
    if (const char *sb = getenv("FCC_SEG_BELOW")) ix.seg_below = (uint32_t)std::max(0, atoi(sb));
    ix.features = o.planner == FCC_PLAN_LEARNED || o.collect_stats;
    // A very large GEMM outlasts the whole left-tile build, so it starts
    // before the left sort; a smaller one waits for the sort to be queued
    // (they compete for the LDS: beside the GEMM a sort pass ran 6x slower).
    // Measured on the order-5 self-contraction: its full 340 GFLOP GEMM was
    // better forked early, its symmetric 170 GFLOP half 0.7 ms better late.
    const uint64_t dk_nbm = dk.ra / DK_BM;
    const double dk_tiles = dk.self ? (double)(dk_nbm * (dk_nbm + 1) / 2) : (double)dk_nbm * (double)(dk.rb / DK_BN);
    const bool early_fork = dk.nh && 2.0 * dk_tiles * DK_BM * DK_BN * (double)dk.kp > 2.5e11;
    if (early_fork) fork_dense();
    const std::function<void()> *after_sort = (dk.nh && !early_fork) ? &fork_dense : nullptr;
    // a long block (the order-5 self-contraction's 1.7e11 flops, not the
    // order-4 Zipf 1.1 one's 1.5e10) takes the left records beside it rather
    // than waiting for them
    const bool records_late = after_sort && 2.0 * dk_tiles * DK_BM * DK_BN * (double)dk.kp >= 5e10;
    if (self && o.left_tile_stride <= 1) build_left_self(pool, g, A, sj, logTL, ix, hold, s, after_sort, records_late);
    else build_left(pool, g, A, rowA, fidA, nvalid, logTL, ix, hold, s, after_sort);
    if (dk.nh && !side_join.st) throw LibError{FCC_ERR_INVALID, "dense-key block was not started"};
    vsB.reset();
    st.row_tile = (int32_t)ix.TL;
    st.col_window = (int32_t)Wd;
    st.left_tiles = ix.natiles;
    st.mirrored = mirror ? 1 : 0;

    dk.tile_dk = ix.tile_dk;
    // Planning and the count pass read only the block's presence map (ready
    // before the GEMM starts, event 10); the numeric pass reads its values and
    // joins the GEMM itself (event 8), so the GEMM runs on beside the count.
    if (dk.nh) FCC_HIP(hipStreamWaitEvent(s, ev[10], 0));
    auto join_block = [&]() {
        if (dk.nh) FCC_HIP(hipStreamWaitEvent(s, ev[8], 0));
    };
    rowA.reset();
    fidA.reset();
    st.dense_keys = dk.nh;
    st.dense_key_products = dk.products;
    FCC_HIP(hipEventRecord(ev[1], s));

    const bool learned = o.planner == FCC_PLAN_LEARNED;
    const bool sink = o.sink == FCC_SINK_COUNT_CHECKSUM;
    if (learned && !ctx->mlp_ready) throw LibError{FCC_ERR_INVALID, "learned planner requested but no MLP parameters set"};

    uint64_t nnz_out = 0;
    const size_t nsink = 4 * (size_t)std::max<uint32_t>(groups, 1);
    k.nsink = nsink;
    k.sink = sink;
    auto &sinkbuf = k.sinkbuf;
    sinkbuf = Tmp<unsigned long long>(pool, nsink);
    fill_u64(sinkbuf, nsink, 0, s);
    if (ix.natiles == 0) {
        FCC_HIP(hipEventRecord(ev[2], s));
        FCC_HIP(hipEventSynchronize(ev[2]));
        st.ms_index = ev_ms(ev[0], ev[1]);
        st.ms_total = ev_ms(ev[0], ev[2]);
        st.bytes_alg = (uint64_t)ta->nnz * (4 * ta->order + 8) + (uint64_t)tb->nnz * (4 * tb->order + 8);
        return FCC_OK;
    }

    // ---- plan
    const uint32_t T = ix.natiles;
    Tmp<float> feat, predc_t;
    auto &pred_t = k.pred;
    if (learned || o.collect_stats) {
        feat = Tmp<float>(pool, (size_t)T * MLP_IN);
        mlp_tile_features(ix, g.spanB, feat, s);
    }
    if (learned) {
        pred_t = Tmp<float>(pool, T);
        predc_t = Tmp<float>(pool, T);
        mlp_infer(ctx->mlp_dev, feat, T, pred_t, predc_t, s);
    }
    PlanParams P;
    P.force = o.force_acc;
    P.TL = ix.TL;
    P.spanB = g.spanB;
    P.Wd = Wd;
    P.margin = 0.15f;
    // per-tile counts are wanted when collecting planner statistics: no runs then
    P.run = o.collect_stats ? 0 : DIRECT_RUN;
    k.plan_force = P.force;
    P.group_rows = groups ? g.spanA / groups : 0;
    P.mirror = mirror ? 1 : 0;
    P.atile_id = ix.atile_id;
    P.logTL = ix.logTL;
    Tmp<uint32_t> t_nwin(pool, T), t_W(pool, T), t_ubase(pool, T), t_w0(pool, T), d_nunits(pool, 1);
    Tmp<uint8_t> t_kind(pool, T);
    Units &U = k.U;
    auto &u_tile = k.u_tile, &u_c0 = k.u_c0, &u_W = k.u_W;
    auto &u_kind = k.u_kind, &u_multi = k.u_multi;
    auto &unit_count = k.unit_count, &unit_off = k.unit_off, &unit_cur = k.unit_cur;
    Tmp<unsigned long long> d_total(pool, 1);

    // ---- accumulate arguments
    AccArgs &X = k.X;
    memset(&X, 0, sizeof(X));
    X.Bcol = ix.Bcol;
    X.Bval = ix.Bval;
    X.fiber_ptr = ix.fiber_ptr;
    X.Ali = ix.Ali;
    X.Aval = ix.Aval;
    X.Afp = ix.Afp;
    X.Afib = ix.Afib;
    auto &fcut = k.fcut;
    if (fcut.get()) {
        X.fcut = fcut;
        X.fcut_n = fcut_n;
        X.fcut_W = Wd;
    }
    X.group_ptr = ix.group_ptr;
    X.gdesc = ix.gdesc;
    X.hdesc = ix.hdesc;
    X.atile_h0 = ix.atile_h0;
    X.group_fiber = ix.group_fiber;
    X.atile_id = ix.atile_id;
    X.atile_g0 = ix.atile_g0;
    X.atile_lg = ix.atile_lg;
    X.st_products = (const uint64_t *)ix.st_products;
    X.st_pprefix = (const uint64_t *)ix.st_pprefix;
    X.Arow = ix.Arow;
    X.natiles = ix.natiles;
    X.unit_order = nullptr;
    X.row_own = (getenv("FCC_ROW_OWN") && !mirror) ? 1 : 0;
    X.mirror = mirror ? 1 : 0;
    X.claim_div = 2 * (ACC_THREADS / 64);
    if (const char *cd = getenv("FCC_CLAIM_DIV")) X.claim_div = (uint32_t)std::max(0, atoi(cd));
    X.seg_below = ix.seg_below;
    X.logTL = ix.logTL;
    X.spanB = g.spanB;
    X.spanA = g.spanA;
    if (dk.nh) {
        X.dk_P = dk.P;
        X.dk_C = dk.C;
        X.dk_tile = dk.tile_dk;
        X.dk_pw = dk.rb / 32;
        X.dk_rb = dk.rb;
        X.dk_splits = dk.splits;
        X.dk_slab = dk.ra * dk.rb;
    }
    X.nxa = g.nxa;
    X.nxb = g.nxb;
    for (int m = 0; m < kMaxOrder; m++) {
        X.xaE[m] = g.xaE[m];
        X.xbE[m] = g.xbE[m];
    }
    auto &work = k.work;
    auto &gmeta = k.gmeta;  // [0]=gcount [1]=cursor [2]=gtotal [3]=units counted from the dense block
    auto &flags = k.flags;  // [0]=gflag [1]=fail
    work = Tmp<unsigned int>(pool, 4);
    gmeta = Tmp<unsigned long long>(pool, 4);
    flags = Tmp<uint32_t>(pool, 4);
    X.work = work;
    X.gcount = gmeta.get();
    X.cursor = gmeta.get() + 1;
    X.block_counted = gmeta.get() + 3;
    X.gflag = flags.get();
    X.fail_flag = flags.get() + 1;
    X.sink = sinkbuf.get();
    X.sinkf = (double *)(sinkbuf.get() + 2);
    auto &g_left = k.g_left;
    auto &g_done = k.g_done;
    if (groups) {
        X.group_rows = g.spanA / groups;
        g_left = Tmp<unsigned int>(pool, groups);
        g_done = Tmp<unsigned long long>(pool, groups + 1);
        X.group_left = g_left;
        X.group_done = g_done;
        X.clock_start = g_done.get() + groups;
    }
    // per-group unit counts, consumed by each numeric launch
    auto arm_groups = [&]() {
        if (!groups) return;
        fill_u32(g_left, groups, 0, s);
        fill_u64(g_done, groups + 1, 0, s);
        group_units(X.unit_tile, ix.atile_id, X.nunits, ix.logTL, X.group_rows, g_left, s);
    };
    unsigned &grid = k.grid;
    grid = 1;

    // units from the exact bounds (pred == null) or from predictions; the
    // unit count comes back with the predicted output total (psum, when
    // given) in one synchronisation
    // tiles per accumulator kind: with no LDS-hash tile the count pass needs
    // LDS only for the dense byte map
    Tmp<unsigned long long> tk_hist(pool, 8);
    unsigned long long tkh[8] = {};
    auto make_units = [&](const float *pred, const unsigned long long *psum, uint64_t *psum_h) {
        P.dense_ratio = pred ? 0.15f : 0.5f;
        launch_plan_tiles(ix, pred, P, t_nwin, t_W, t_kind, t_w0, s);
        fill_u64(tk_hist, 8, 0, s);
        k_kind_hist<<<std::min(grid256(T), 1024u), 256, 0, s>>>(t_kind, T, tk_hist);
        scan_u32(pool, t_nwin, t_ubase, T, d_nunits, s);
        pool.read(&U.n, d_nunits.get(), 4, s);
        pool.read(tkh, tk_hist.get(), sizeof(tkh), s);
        if (psum) pool.read(psum_h, psum, 8, s);
        pool.sync(s);
        u_tile = Tmp<uint32_t>(pool, U.n);
        u_c0 = Tmp<uint32_t>(pool, U.n);
        u_W = Tmp<uint32_t>(pool, U.n);
        u_kind = Tmp<uint8_t>(pool, U.n);
        u_multi = Tmp<uint8_t>(pool, U.n);
        U.tile = u_tile;
        U.c0 = u_c0;
        U.W = u_W;
        U.kind = u_kind;
        U.multi = u_multi;
        U.pred = nullptr;
        launch_emit_units(t_nwin, t_W, t_kind, t_ubase, t_w0, T, g.spanB, pred, U, s);
        unit_count = Tmp<unsigned long long>(pool, U.n);
        unit_off = Tmp<unsigned long long>(pool, U.n);
        unit_cur = Tmp<unsigned long long>(pool, U.n);
        X.nunits = U.n;
        X.unit_tile = U.tile;
        X.unit_c0 = U.c0;
        X.unit_W = U.W;
        X.unit_kind = U.kind;
        X.unit_multi = U.multi;
        X.unit_count = unit_count;
        X.unit_off = unit_off;
        X.unit_cursor = unit_cur;
        grid = (unsigned)(o.workgroups > 0 ? o.workgroups : (tuned_wg > 0 ? tuned_wg : ctx->num_cus * 2));
        grid = std::max(1u, std::min<unsigned>(grid, U.n));
        st.units = U.n;
    };
    // learned: predicted outputs summed over tiles (each capped by its exact
    // bounds) -- the output capacity
    uint64_t predicted = 0;
    Tmp<unsigned long long> psum(pool, 1);
    if (learned) {
        fill_u64(psum, 1, 0, s);
        k_pred_sum<<<std::min(grid256(T), 2048u), 256, 0, s>>>(pred_t, (const uint64_t *)ix.st_products,
                                               (const uint64_t *)ix.st_maxgp, ix.st_direct,
                                               (const uint64_t *)ix.st_outd, (double)ix.TL * (double)g.spanB, T,
                                               psum);
    }
    make_units(learned ? pred_t.get() : nullptr, learned ? psum.get() : nullptr, &predicted);
    FCC_HIP(hipEventRecord(ev[2], s));

    auto &gk = k.gk;
    auto &gv = k.gv;
    auto &gu = k.gu;
    auto alloc_ghash = [&](uint64_t want) {
        uint64_t cap = pow2_ceil(std::max<uint64_t>(want, 1u << 16));
        gk = Tmp<unsigned long long>(pool, cap);
        gv = Tmp<double>(pool, cap);
        gu = Tmp<uint32_t>(pool, cap);
        X.gkeys = gk;
        X.gvals = gv;
        X.gunit = gu;
        X.gmask = cap - 1;
    };
    // statistics read back with the results: [0, 8) units per accumulator
    // kind, [8] the learned planner's summed tile error (double bits)
    Tmp<unsigned long long> tail(pool, 9);
    // counters, flags, the global hash and the statistics: one launch
    auto reset_ghash = [&](bool with_sink) {
        FillSet f;
        f.add(X.gkeys, X.gmask + 1, ~0ull, 8);
        f.add(X.gvals, X.gmask + 1, 0ull, 8);  // +0.0
        f.add(gmeta.get(), 4, 0, 8);
        f.add(flags.get(), 4, 0, 4);
        f.add(work.get(), 4, 0, 4);
        f.add(tail.get(), 9, 0, 8);
        if (with_sink) f.add(sinkbuf.get(), nsink, 0, 8);
        fill_many(f, s);
    };
    // After a numeric pass: the statistics kernels and every small read-back
    // (flags, global-hash meta, summaries, statistics, group completion) are
    // queued together and covered by one synchronisation.
    unsigned long long sk[4] = {}, th[9] = {}, hm[4] = {};
    uint32_t hf[2] = {};
    std::vector<unsigned long long> gdone;
    auto fetch_tail = [&]() {
        k_kind_hist<<<std::min(grid256(U.n), 1024u), 256, 0, s>>>(U.kind, U.n, tail);
        pool.read(hf, flags.get(), 8, s);
        pool.read(hm, gmeta.get(), 32, s);
        pool.read(sk, sinkbuf.get(), 32, s);
        pool.read(th, tail.get(), sizeof(th), s);
        if (groups) {
            group_sink.resize(nsink);
            gdone.resize(groups + 1);
            pool.read(group_sink.data(), sinkbuf.get(), 8 * nsink, s);
            pool.read(gdone.data(), g_done.get(), 8 * (groups + 1), s);
        }
        pool.sync(s);
    };

    auto &out_c = k.out_c;
    auto &out_v = k.out_v;
    uint64_t stride = 0;
    // Exact output counts of the whole bands of every S-th tile (0 for the
    // others) into tn2: a count pass without mirroring over unclipped fiber
    // ranges, on units planned from the exact bounds. Under mirroring a
    // tile's own units write only part of its band (cells with column >= row,
    // and their transposes), so per-tile counts need this pass. With grow,
    // an overflowing global key set is enlarged and the pass repeated;
    // without, overflow returns false.
    auto band_counts = [&](uint32_t S, unsigned long long *tn2, bool grow) -> bool {
        PlanParams P2 = P;
        P2.mirror = 0;
        P2.run = 0;
        P2.dense_ratio = 0.5f;
        Tmp<uint32_t> n2(pool, T), W2(pool, T), b2(pool, T), w2(pool, T), tot2(pool, 1);
        Tmp<uint8_t> k2(pool, T);
        launch_plan_tiles(ix, nullptr, P2, n2, W2, k2, w2, s);
        k_sample_tiles<<<grid256(T), 256, 0, s>>>(n2, k2, T, S);
        scan_u32(pool, n2, b2, T, tot2, s);
        uint32_t nu = 0;
        pool.read(&nu, tot2.get(), 4, s);
        pool.sync(s);
        Units U2;
        Tmp<uint32_t> u2t(pool, nu), u2c(pool, nu), u2w(pool, nu);
        Tmp<uint8_t> u2k(pool, nu), u2m(pool, nu);
        U2.n = nu;
        U2.tile = u2t;
        U2.c0 = u2c;
        U2.W = u2w;
        U2.kind = u2k;
        U2.multi = u2m;
        Tmp<unsigned long long> cnt2(pool, nu);
        Tmp<unsigned int> work2(pool, 4);
        if (nu) launch_emit_units(n2, W2, k2, b2, w2, T, g.spanB, nullptr, U2, s);
        for (int attempt = 0;; attempt++) {
            AccArgs XS = X;
            XS.mirror = 0;
            XS.unclipped = 1;
            XS.seg_below = 0;  // no heavy segments: they were cut from clipped ranges
            XS.nunits = nu;
            XS.unit_tile = U2.tile;
            XS.unit_c0 = U2.c0;
            XS.unit_W = U2.W;
            XS.unit_kind = U2.kind;
            XS.unit_multi = U2.multi;
            XS.unit_order = nullptr;
            XS.unit_count = cnt2;
            XS.work = work2;
            XS.set_words = SYM_SLOTS;
            FillSet f;
            f.add(XS.gkeys, XS.gmask + 1, ~0ull, 8);
            f.add(gmeta.get(), 4, 0, 8);
            f.add(flags.get(), 4, 0, 4);
            f.add(work2.get(), 4, 0, 4);
            if (nu) f.add(cnt2.get(), nu, 0, 8);
            f.add(tn2, T, 0, 8);
            fill_many(f, s);
            if (nu) launch_count(XS, std::max(1u, std::min<unsigned>((unsigned)ctx->num_cus * 2, nu)), s);
            if (nu) k_unit_to_tile<<<grid256(nu), 256, 0, s>>>(cnt2, U2.tile, U2.kind, nu, tn2);
            // DIRECT tiles: every product is a distinct output of the band
            k_direct_tile_count<<<grid256(T), 256, 0, s>>>(ix.st_direct, (const uint64_t *)ix.st_products, T,
                                                           P.force, tn2);
            uint32_t gf2 = 0;
            pool.read(&gf2, flags.get(), 4, s);
            pool.sync(s);
            fill_u32(flags, 4, 0, s);
            if (!gf2) return true;
            if (!grow) return false;
            if (attempt >= 3) throw LibError{FCC_ERR_OVERFLOW, "band count global set overflow"};
            alloc_ghash((X.gmask + 1) * 16);
        }
    };
    bool use_exact = !learned;
    bool mape_pending = false;

    if (learned) {
        // ---- learned planning: predicted capacity, single numeric pass
        // never more than the products bound: no tile can emit more outputs than products
        uint64_t capacity = std::min<uint64_t>((uint64_t)((double)predicted * 1.10) + 4096, products + 16);
        if (user_c || user_v) capacity = user_cap;  // the caller's buffers, whole
        alloc_ghash(1u << 20);
        join_block();
        FCC_HIP(hipEventRecord(ev[3], s));
        Tmp<unsigned long long> tn(pool, T);
        for (int attempt = 0;; attempt++) {
            if (!sink && !user_v) {
                out_c = Tmp<uint32_t>(pool, (size_t)g.dC * capacity);
                out_v = Tmp<double>(pool, (size_t)capacity);
            }
            stride = capacity;
            X.ocoords = user_v ? user_c : out_c.get();
            X.ovals = user_v ? user_v : out_v.get();
            X.ostride = stride;
            X.capacity = sink ? ~0ull : capacity;
            reset_ghash(sink);
            arm_groups();
            if (attempt == 0) FCC_HIP(hipEventRecord(ev[4], s));
            launch_numeric(X, grid, bitmap, sink, 1, s);
            if (attempt == 0) FCC_HIP(hipEventRecord(ev[5], s));
            launch_gdrain(X, sink, 1, s);
            FCC_HIP(hipEventRecord(ev[6], s));
            // tile-level MAPE against the actual per-unit outputs (kept only
            // when this pass is the final one)
            if (!mirror) {
                fill_u64(tn, T, 0, s);
                k_unit_to_tile<<<grid256(U.n), 256, 0, s>>>(unit_cur, U.tile, U.kind, U.n, tn);
                k_direct_tile_count<<<grid256(T), 256, 0, s>>>(ix.st_direct, (const uint64_t *)ix.st_outd, T,
                                                               P.force, tn);
                k_mape<<<std::min(grid256(T), 2048u), 256, 0, s>>>(pred_t, tn, T, reinterpret_cast<double *>(tail.get() + 8));
            }
            fetch_tail();
            st.ghash_entries = hm[0];
            if (hf[0]) {
                // the predictions left more spill than the global hash holds:
                // fall back to the exact planner (re-planned from the bounds)
                st.reruns++;
                use_exact = true;
                out_c.reset();
                out_v.reset();
                break;
            }
            if (hf[1]) {
                if (user_v) {
                    out->nnz = (int64_t)hm[1];
                    throw LibError{FCC_ERR_OVERFLOW, "output capacity " + std::to_string(user_cap) + " below the " +
                                                         std::to_string(hm[1]) + " outputs (out->nnz)"};
                }
                capacity = hm[1];
                st.reruns++;
                if (attempt >= 4) throw LibError{FCC_ERR_OVERFLOW, "learned capacity re-run did not converge"};
                continue;
            }
            nnz_out = sink ? 0 : hm[1];
            mape_pending = !mirror;  // per-tile outputs = the whole tile band only without mirroring
            break;
        }
        if (!use_exact && mirror && o.judge) {
            // A mirrored tile's outputs are not its whole band, the quantity
            // the network predicts: the exact planner judges the predictions
            // on a sample of about 256 tiles instead.
            const uint32_t S = std::max<uint32_t>(1, T / 256);
            Tmp<unsigned long long> tn2(pool, T);
            Tmp<double> acc2(pool, 2);
            fill_f64(acc2, 2, 0.0, s);
            if (band_counts(S, tn2, false)) {
                k_mape_sampled<<<grid256(T), 256, 0, s>>>(pred_t, tn2, T, S, acc2);
                double m2[2] = {0.0, 0.0};
                pool.read(m2, acc2.get(), 16, s);
                pool.sync(s);
                if (m2[1] > 0) {
                    st.mlp_mape = m2[0] / m2[1];
                    st.mape_tiles = (uint64_t)m2[1];
                }
            }
        }
    }
    if (use_exact) {
        if (learned) {
            make_units(nullptr, nullptr, nullptr);
            fill_u64(sinkbuf, nsink, 0, s);
        }
        // ---- exact symbolic planning: count, scan, then fill
        alloc_ghash(1u << 16);
        // the key set's LDS (see make_units) and, when it is only the dense
        // byte map, a default grid of four workgroups per CU
        const bool small_set = tkh[KIND_LHASH] == 0 && tkh[KIND_GHASH] == 0;
        X.set_words = small_set ? DENSE_SLOTS / 4 : SYM_SLOTS;
        unsigned cgrid = grid;
        if (small_set && o.workgroups <= 0 && tuned_wg <= 0)
            cgrid = std::max(1u, std::min<unsigned>((unsigned)ctx->num_cus * 4, U.n));
        for (int attempt = 0;; attempt++) {
            reset_ghash(false);
            fill_u64(unit_count, U.n, 0, s);
            launch_count(X, cgrid, s);
            launch_refine_kinds(unit_count, U.W, U.n, o.force_acc, mirror ? 1 : 0, ix.TL, U.kind, gmeta.get() + 2,
                                s);
            scan_u64(pool, (const uint64_t *)unit_count.get(), (uint64_t *)unit_off.get(), U.n,
                     (uint64_t *)d_total.get(), s);
            uint32_t gf = 0;
            pool.read(&gf, flags.get(), 4, s);
            pool.read(hm, gmeta.get(), 32, s);
            pool.read(&nnz_out, d_total.get(), 8, s);
            pool.sync(s);
            if (!gf) break;
            if (attempt >= 3) throw LibError{FCC_ERR_OVERFLOW, "count-pass global set overflow"};
            alloc_ghash((X.gmask + 1) * 16);
        }
        uint64_t gtotal = hm[2];
        st.units_block_counted = hm[3];
        join_block();
        FCC_HIP(hipEventRecord(ev[3], s));
        if (o.collect_stats) {
            // labels: exact tile nnz and distinct output columns
            Tmp<unsigned long long> tn(pool, T), tc(pool, T);
            fill_u64(tn, T, 0, s);
            k_unit_to_tile<<<grid256(U.n), 256, 0, s>>>(unit_count, U.tile, U.kind, U.n, tn);
            k_direct_tile_count<<<grid256(T), 256, 0, s>>>(ix.st_direct, (const uint64_t *)ix.st_outd, T,
                                                           P.force, tn);
            k_tile_cols<<<std::min<unsigned>(T, 4096), 256, 0, s>>>(ix.atile_g0, ix.group_fiber, ix.fiber_ptr,
                                                                     ix.Bcol, T, g.spanB, tc);
            Tmp<float> lab(pool, (size_t)T * 2);
            k_labels<<<grid256(T), 256, 0, s>>>(tn, tc, feat, T, lab);
            ctx->feat.resize((size_t)T * MLP_IN);
            ctx->labels.resize((size_t)T * 2);
            FCC_HIP(hipMemcpyAsync(ctx->feat.data(), feat.get(), ctx->feat.size() * 4, hipMemcpyDeviceToHost, s));
            FCC_HIP(hipMemcpyAsync(ctx->labels.data(), lab.get(), ctx->labels.size() * 4, hipMemcpyDeviceToHost, s));
            FCC_HIP(hipStreamSynchronize(s));
        }
        if (!sink && user_v && nnz_out > user_cap) {
            out->nnz = (int64_t)nnz_out;
            throw LibError{FCC_ERR_OVERFLOW, "output capacity " + std::to_string(user_cap) + " below the " +
                                                 std::to_string(nnz_out) + " outputs (out->nnz)"};
        }
        if (!sink && !user_v) {
            out_c = Tmp<uint32_t>(pool, (size_t)g.dC * nnz_out);
            out_v = Tmp<double>(pool, (size_t)nnz_out);
        }
        stride = user_v ? user_cap : nnz_out;
        X.ocoords = user_v ? user_c : out_c.get();
        X.ovals = user_v ? user_v : out_v.get();
        X.ostride = stride;
        X.capacity = nnz_out;
        if (gtotal > (X.gmask + 1) / 2) alloc_ghash(2 * gtotal);
        reset_ghash(false);
        arm_groups();
        FCC_HIP(hipEventRecord(ev[4], s));
        launch_numeric(X, grid, bitmap, sink, 0, s);
        FCC_HIP(hipEventRecord(ev[5], s));
        launch_gdrain(X, sink, 0, s);
        FCC_HIP(hipEventRecord(ev[6], s));
        fetch_tail();
        if (hf[0]) throw LibError{FCC_ERR_OVERFLOW, "global hash overflow in the numeric pass"};
        if (hf[1]) throw LibError{FCC_ERR_OVERFLOW, "numeric pass disagreed with the count pass"};
        st.ghash_entries = hm[0];
    }

    if (keep && mirror) {
        // a plan answers per-tile output counts (fcc_plan_tiles): under
        // mirroring these are counted once here, over every tile's band
        k.tile_band = Tmp<unsigned long long>(pool, T);
        band_counts(1, k.tile_band, true);
    }

    // ---- results (read back by fetch_tail)
    st.units_dense = th[KIND_DENSE];
    st.units_lhash = th[KIND_LHASH];
    st.units_ghash = th[KIND_GHASH];
    st.units_direct = th[KIND_DIRECT];
    if (mape_pending) {
        double m;
        memcpy(&m, &th[8], 8);
        st.mlp_mape = m / (double)T;
        st.mape_tiles = T;
    }
    if (sink) {
        nnz_out = sk[0];
        out->coord_hash = sk[0] ? sk[1] : 0;
        double d[2];
        memcpy(d, &sk[2], 16);
        out->val_sum = d[0];
        out->val_wsum = d[1];
    }
    out->nnz = (int64_t)nnz_out;
    out->coord_stride = (int64_t)stride;
    k.stride = stride;
    k.nnz_out = nnz_out;
    k.mode = (learned && !use_exact) ? 1 : 0;
    k.empty = false;
    if (!sink) {
        out->coords = user_v ? user_c : out_c.p;
        out->vals = user_v ? user_v : out_v.p;
        if (!keep && !user_v) {  // the result owns the output; a plan keeps it
            ResultPriv *rp = new ResultPriv;
            rp->coords = out_c.p;
            rp->vals = out_v.p;
            out_c.p = nullptr;
            out_v.p = nullptr;
            out->priv = rp;
        }
        if (o.summarize && nnz_out) {
            Tmp<unsigned long long> hs(pool, 1);
            Tmp<double> vs(pool, 2);
            Tmp<uint32_t> E(pool, kMaxOrder);
            fill_u64(hs, 1, 0, s);
            fill_f64(vs, 2, 0.0, s);
            FCC_HIP(hipMemcpyAsync(E.get(), out->extents, 4 * kMaxOrder, hipMemcpyHostToDevice, s));
            k_summarize<<<grid256(nnz_out), 256, 0, s>>>(out->coords, out->vals, nnz_out, stride, g.dC, E, hs, vs);
            unsigned long long hh;
            double dv[2];
            pool.read(&hh, hs.get(), 8, s);
            pool.read(dv, vs.get(), 16, s);
            pool.sync(s);
            out->coord_hash = hh;
            out->val_sum = dv[0];
            out->val_wsum = dv[1];
        }
    }
    st.ms_index = ev_ms(ev[0], ev[1]);
    st.ms_dense = ev_ms(ev[7], ev[8]);
    st.ms_dense_gemm = ev_ms(ev[10], ev[11]);
    if (dk.nh) {
        // the tiles the GEMM computes: all, or (self) those on and above the diagonal
        const uint64_t nbm = dk.ra / DK_BM, tiles = dk.self ? nbm * (nbm + 1) / 2 : nbm * (dk.rb / DK_BN);
        st.dense_flops = 2ull * tiles * DK_BM * DK_BN * dk.kp;
        // the tiles' units read the block's cells once (every k split) plus the presence words
        st.dense_block_bytes = dk.splits * dk.spanA * dk.spanB * 8 + dk.spanA * dk.spanB / 8;
    }
    st.ms_plan = ev_ms(ev[1], ev[2]);
    st.ms_symbolic = (learned && !use_exact) ? 0.0 : ev_ms(ev[2], ev[3]);
    st.ms_numeric = ev_ms(ev[4], ev[5]);
    st.ms_numeric_kernel = st.ms_numeric;
    st.ms_overflow = ev_ms(ev[5], ev[6]);
    st.ms_total = ev_ms(ev[0], ev[6]);
    st.ms_done = st.ms_total;
    st.bytes_alg = (uint64_t)ta->nnz * (4 * ta->order + 8) + (uint64_t)tb->nnz * (4 * tb->order + 8) +
                   nnz_out * (4 * (uint64_t)g.dC + 8);
    if (groups) {
        // a group's output is complete when its last unit ends; entries that
        // went through the global hash only after the drain
        group_ms.assign(groups, st.ms_total);
        const double t0 = ev_ms(ev[0], ev[4]);
        const unsigned long long c0 = gdone[groups];
        if (!st.reruns && !st.ghash_entries && c0)
            for (uint32_t q = 0; q < groups; q++)
                if (gdone[q] >= c0) group_ms[q] = t0 + (double)(gdone[q] - c0) / (double)ctx->clock_khz;
    }
    return FCC_OK;
}

// ------------------------------------------------------------ fused batches

// Per-pair extents (maximum coordinate + 1, or the caller's, whichever is
// larger) decide how many consecutive pairs one fused launch can hold: the
// returned chunk ends (fcc_chunk.h).
std::vector<uint32_t> chunk_pairs(fcc_ctx *ctx, int32_t n, const fcc_tensor *a, const fcc_tensor *b,
                                  const int32_t *a_modes, const int32_t *b_modes, int32_t n_contr) {
    DevPool &pool = ctx->pool;
    hipStream_t s = ctx->stream;
    const int dA = a[0].order, dB = b[0].order;
    std::vector<uint32_t> pext[2];
    std::vector<uint64_t> pnnz[2];
    for (int k = 0; k < 2; k++) {
        const fcc_tensor *src = k ? b : a;
        const int d = k ? dB : dA;
        pext[k].assign((size_t)n * d, 0);
        pnnz[k].assign(n, 0);
        std::vector<PairSrc> dev;
        std::vector<int32_t> dev_idx;
        for (int32_t i = 0; i < n; i++) {
            const fcc_tensor &t = src[i];
            pnnz[k][i] = (uint64_t)t.nnz;
            if (t.nnz && t.on_device) {
                dev.push_back(PairSrc{t.coords, t.vals, (uint64_t)t.nnz});
                dev_idx.push_back(i);
            } else if (t.nnz) {
                for (int m = 0; m < d; m++) {
                    uint32_t mx = 0;
                    for (int64_t j = 0; j < t.nnz; j++) mx = std::max(mx, t.coords[(size_t)m * t.nnz + j]);
                    pext[k][(size_t)i * d + m] = mx + 1;
                }
            }
        }
        if (!dev.empty()) {
            Tmp<unsigned char> dps(pool, sizeof(PairSrc) * dev.size());
            Tmp<uint32_t> dmx(pool, dev.size() * d);
            FCC_HIP(hipMemcpyAsync(dps.get(), dev.data(), sizeof(PairSrc) * dev.size(), hipMemcpyHostToDevice, s));
            uint64_t mx_nnz = 0;
            for (const PairSrc &q : dev) mx_nnz = std::max<uint64_t>(mx_nnz, q.nnz);
            pair_max((const PairSrc *)dps.get(), (uint32_t)dev.size(), d, mx_nnz, dmx, s);
            std::vector<uint32_t> hm(dev.size() * d);
            FCC_HIP(hipMemcpyAsync(hm.data(), dmx.get(), 4 * hm.size(), hipMemcpyDeviceToHost, s));
            FCC_HIP(hipStreamSynchronize(s));
            for (size_t q = 0; q < dev.size(); q++)
                for (int m = 0; m < d; m++) pext[k][(size_t)dev_idx[q] * d + m] = hm[q * d + m] + 1;
        }
        for (int32_t i = 0; i < n; i++)
            if (src[i].extents)
                for (int m = 0; m < d; m++)
                    pext[k][(size_t)i * d + m] = std::max(pext[k][(size_t)i * d + m], src[i].extents[m]);
    }
    return batch_chunks((uint32_t)n, dA, dB, a_modes, b_modes, n_contr, pext[0].data(), pext[1].data(),
                        pnnz[0].data(), pnnz[1].data(), kMaxBatchPairs);
}

// One operand side of a fused batch: the cnt pairs' entries concatenated in
// pair order on the device, with a trailing pair-index mode; extents the
// bounding box of the pairs' (explicit extents included), pair mode = cnt.
struct FusedSide {
    int d = 0;
    std::vector<unsigned long long> off;  // pair i's entries: [off[i], off[i + 1])
    Tmp<uint32_t> c;
    Tmp<double> v;
    uint32_t E[kMaxOrder + 1] = {};
};

void fuse_side(fcc_ctx *ctx, uint32_t cnt, const fcc_tensor *src, int d, FusedSide &S) {
    DevPool &pool = ctx->pool;
    hipStream_t s = ctx->stream;
    S.d = d;
    S.off.assign(cnt + 1, 0);
    uint64_t host_n = 0;
    for (uint32_t i = 0; i < cnt; i++) {
        S.off[i + 1] = S.off[i] + (uint64_t)src[i].nnz;
        if (!src[i].on_device) host_n += (uint64_t)src[i].nnz;
    }
    const uint64_t total = S.off[cnt];
    if (total >= (1ull << 32)) throw LibError{FCC_ERR_INVALID, "batch operand exceeds 2^32 nonzeros"};
    Tmp<uint32_t> stc(pool, (size_t)d * host_n);
    Tmp<double> stv(pool, (size_t)host_n);
    std::vector<PairSrc> ps(cnt);
    uint64_t ho = 0;
    for (uint32_t i = 0; i < cnt; i++) {
        const fcc_tensor &t = src[i];
        ps[i].nnz = (uint64_t)t.nnz;
        if (t.on_device || !t.nnz) {
            ps[i].coords = t.coords;
            ps[i].vals = t.vals;
        } else {
            FCC_HIP(hipMemcpyAsync(stc.get() + (size_t)d * ho, t.coords, 4ull * d * t.nnz, hipMemcpyHostToDevice, s));
            FCC_HIP(hipMemcpyAsync(stv.get() + ho, t.vals, 8ull * t.nnz, hipMemcpyHostToDevice, s));
            ps[i].coords = stc.get() + (size_t)d * ho;
            ps[i].vals = stv.get() + ho;
            ho += (uint64_t)t.nnz;
        }
    }
    Tmp<unsigned char> dps(pool, sizeof(PairSrc) * cnt);
    Tmp<unsigned long long> doff(pool, cnt + 1);
    Tmp<uint32_t> dmax(pool, kMaxOrder);
    FCC_HIP(hipMemcpyAsync(dps.get(), ps.data(), sizeof(PairSrc) * cnt, hipMemcpyHostToDevice, s));
    FCC_HIP(hipMemcpyAsync(doff.get(), S.off.data(), 8ull * (cnt + 1), hipMemcpyHostToDevice, s));
    fill_u32(dmax, kMaxOrder, 0, s);
    S.c = Tmp<uint32_t>(pool, (size_t)(d + 1) * total);
    S.v = Tmp<double>(pool, (size_t)total);
    gather_pairs((const PairSrc *)dps.get(), doff, cnt, d, total, S.c, S.v, dmax, s);
    uint32_t hmax[kMaxOrder];
    FCC_HIP(hipMemcpyAsync(hmax, dmax.get(), sizeof(hmax), hipMemcpyDeviceToHost, s));
    FCC_HIP(hipStreamSynchronize(s));
    for (int m = 0; m < d; m++) {
        uint32_t e = total ? hmax[m] + 1 : 0;
        for (uint32_t i = 0; i < cnt; i++)
            if (src[i].extents) e = std::max(e, src[i].extents[m]);
        S.E[m] = e;
    }
    S.E[d] = cnt;
}

}  // namespace

// ====================================================================== ABI

extern "C" {

void fcc_options_default(fcc_options *o) {
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->planner = FCC_PLAN_EXACT;
    o->sink = FCC_SINK_MATERIALIZE;
    o->force_acc = FCC_ACC_AUTO;
    o->left_tile_stride = 1;
}

int fcc_create(fcc_ctx **out, int device) {
    if (!out) return FCC_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return FCC_ERR_HIP;
    }
    if (device < 0 || device >= n) return FCC_ERR_INVALID;
    fcc_ctx *c = new (std::nothrow) fcc_ctx;
    if (!c) return FCC_ERR_OOM;
    c->device = device;
    try {
        FCC_HIP(hipSetDevice(device));
        FCC_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        FCC_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        for (auto &e : c->ev) FCC_HIP(hipEventCreate(&e));
        hipDeviceProp_t prop;
        FCC_HIP(hipGetDeviceProperties(&prop, device));
        c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
        c->mem_total = prop.totalGlobalMem;
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            c->clock_khz = khz;
        FCC_HIP(hipMalloc(&c->mlp_dev, sizeof(float) * MLP_PARAMS));
        c->pool.init_pinned();
        // the no-allocation scratch on this device (the device is current)
        if (hipGetSymbolAddress(&c->scratch, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            c->scratch = nullptr;
        }
    } catch (HipError &e) {
        for (auto &ev : c->ev)
            if (ev) (void)hipEventDestroy(ev);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        if (c->side) (void)hipStreamDestroy(c->side);
        delete c;
        return FCC_ERR_HIP;
    }
    *out = c;
    return FCC_OK;
}

void fcc_destroy(fcc_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    c->pool.release_all();
    if (c->mlp_dev) (void)hipFree(c->mlp_dev);
    for (auto &e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->side) (void)hipStreamDestroy(c->side);
    delete c;
}

const char *fcc_last_error(const fcc_ctx *c) { return c ? c->err.c_str() : "no context"; }

int fcc_set_memory_budget(fcc_ctx *c, uint64_t bytes) {
    if (!c) return FCC_ERR_INVALID;
    c->pool.set_budget((size_t)bytes);
    return FCC_OK;
}

int fcc_memory_stats(const fcc_ctx *c, uint64_t *held, uint64_t *in_use) {
    if (!c) return FCC_ERR_INVALID;
    if (held) *held = c->pool.bytes_cached();
    if (in_use) *in_use = c->pool.bytes_in_use();
    return FCC_OK;
}

void *fcc_stream(fcc_ctx *c) { return c ? (void *)c->stream : nullptr; }

namespace {
int check_modes(const fcc_tensor *a, const fcc_tensor *b, const int32_t *am, const int32_t *bm, int32_t r,
                const int32_t *ab, const int32_t *bb, int32_t nb, std::vector<int32_t> &ka, std::vector<int32_t> &kb,
                std::string &msg) {
    if (r < 0 || nb < 0 || r + nb > a->order || r + nb > b->order || (r > 0 && (!am || !bm)) ||
        (nb > 0 && (!ab || !bb))) {
        msg = "bad contraction/batch mode count";
        return FCC_ERR_INVALID;
    }
    ka.assign(ab, ab + nb);
    kb.assign(bb, bb + nb);
    ka.insert(ka.end(), am, am + r);
    kb.insert(kb.end(), bm, bm + r);
    for (size_t i = 0; i < ka.size(); i++) {
        if (ka[i] < 0 || ka[i] >= a->order || kb[i] < 0 || kb[i] >= b->order) {
            msg = "contraction or batch mode out of range";
            return FCC_ERR_INVALID;
        }
        for (size_t k = 0; k < i; k++)
            if (ka[k] == ka[i] || kb[k] == kb[i]) {
                msg = "repeated contraction or batch mode";
                return FCC_ERR_INVALID;
            }
    }
    if (a->order + b->order - 2 * (r + nb) + nb > FCC_MAX_ORDER) {
        msg = "output order too large";
        return FCC_ERR_INVALID;
    }
    return FCC_OK;
}

int contract_impl(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                  const fcc_tensor *b, const int32_t *b_modes, const int32_t *b_batch, int32_t n_contr,
                  int32_t n_batch, const fcc_options *options, fcc_result *out) {
    if (!ctx) return FCC_ERR_INVALID;
    if (!out) return fail(ctx, FCC_ERR_INVALID, "out is NULL");
    memset(out, 0, sizeof(*out));
    std::string msg;
    int rc = validate(a, "a", msg);
    if (!rc) rc = validate(b, "b", msg);
    std::vector<int32_t> ka, kb;
    if (!rc) rc = check_modes(a, b, a_modes, b_modes, n_contr, a_batch, b_batch, n_batch, ka, kb, msg);
    if (rc) return fail(ctx, rc, msg);
    Contraction c;
    c.ctx = ctx;
    c.s = ctx->stream;
    fcc_options_default(&c.o);
    if (options) c.o = *options;
    if (c.o.left_tile_stride < 1) c.o.left_tile_stride = 1;
    c.out = out;
    try {
        FCC_HIP(hipSetDevice(ctx->device));
        rc = c.run(a, ka.data(), b, kb.data(), n_contr + n_batch, n_batch);
        ctx->err.clear();
        return rc;
    } catch (HipError &e) {
        (void)hipGetLastError();
        fcc_result_free(ctx, out);
        return fail(ctx, FCC_ERR_HIP, e.where + ": " + hipGetErrorString(e.code));
    } catch (LibError &e) {
        fcc_result_free(ctx, out);
        return fail(ctx, e.code, e.msg);
    } catch (std::bad_alloc &) {
        fcc_result_free(ctx, out);
        return fail(ctx, FCC_ERR_OOM, "host allocation failed");
    }
}
}  // namespace

int fcc_contract(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const fcc_tensor *b,
                 const int32_t *b_modes, int32_t n_contr, const fcc_options *options, fcc_result *out) {
    return contract_impl(ctx, a, a_modes, nullptr, b, b_modes, nullptr, n_contr, 0, options, out);
}

int fcc_contract_noalloc(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const fcc_tensor *b,
                         const int32_t *b_modes, int32_t n_contr, const fcc_options *options, uint32_t *out_coords,
                         double *out_vals, int64_t out_capacity, fcc_result *out) {
    if (!ctx) return FCC_ERR_INVALID;
    if (!out) return fail(ctx, FCC_ERR_INVALID, "out is NULL");
    memset(out, 0, sizeof(*out));
    std::string msg;
    int rc = validate(a, "a", msg);
    if (!rc) rc = validate(b, "b", msg);
    std::vector<int32_t> ka, kb;
    if (!rc) rc = check_modes(a, b, a_modes, b_modes, n_contr, nullptr, nullptr, 0, ka, kb, msg);
    if (rc) return fail(ctx, rc, msg);
    fcc_options o;
    fcc_options_default(&o);
    if (options) o = *options;
    if (o.left_tile_stride < 1) o.left_tile_stride = 1;
    if (o.collect_stats) return fail(ctx, FCC_ERR_INVALID, "collect_stats needs host storage: not without allocation");
    // the dense-key block would need spanA x spanB cells of scratch
    o.dense_keys = -1.0f;
    const bool sink = o.sink == FCC_SINK_COUNT_CHECKSUM;
    const int dC = a->order + b->order - 2 * n_contr;
    if (!sink && (out_capacity < 0 || (out_capacity > 0 && (!out_vals || (dC > 0 && !out_coords)))))
        return fail(ctx, FCC_ERR_INVALID, "materialised outputs need out_vals, out_coords and out_capacity >= 0");
    if (!ctx->scratch) return fail(ctx, FCC_ERR_HIP, "no-allocation scratch not found on this device");
    DevPool &pool = ctx->pool;
    struct ArenaEnd {  // declared before the contraction: its temporaries go first
        DevPool &p;
        ~ArenaEnd() { p.end_arena(); }
    } arena_end{pool};
    pool.begin_arena(ctx->scratch, FCC_NOALLOC_SCRATCH_BYTES);
    Contraction c;
    c.ctx = ctx;
    c.s = ctx->stream;
    c.o = o;
    c.out = out;
    if (!sink) {
        c.user_c = out_coords;
        c.user_v = out_vals ? out_vals : reinterpret_cast<double *>(ctx->scratch);  // capacity 0: never written
        c.user_cap = (uint64_t)out_capacity;
    }
    try {
        FCC_HIP(hipSetDevice(ctx->device));
        rc = c.run(a, ka.data(), b, kb.data(), n_contr, 0);
        if (!sink && out->nnz == 0) {
            out->coords = nullptr;
            out->vals = nullptr;
        }
        ctx->err.clear();
        return rc;
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where + ": " + hipGetErrorString(e.code));
    } catch (LibError &e) {
        out->coords = nullptr;
        out->vals = nullptr;
        return fail(ctx, e.code, e.msg);
    } catch (std::bad_alloc &) {
        return fail(ctx, FCC_ERR_OOM, "host allocation failed");
    }
}

int fcc_contract_ex(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                    const fcc_tensor *b, const int32_t *b_modes, const int32_t *b_batch, int32_t n_contr,
                    int32_t n_batch, const fcc_options *options, fcc_result *out) {
    return contract_impl(ctx, a, a_modes, a_batch, b, b_modes, b_batch, n_contr, n_batch, options, out);
}

int fcc_result_to_host(fcc_ctx *ctx, const fcc_result *r, uint32_t *coords, double *vals) {
    if (!ctx || !r) return FCC_ERR_INVALID;
    if (r->nnz == 0) return FCC_OK;
    if (!r->coords || !r->vals) return fail(ctx, FCC_ERR_INVALID, "result holds no materialised output");
    try {
        for (int m = 0; m < r->order && coords; m++)
            FCC_HIP(hipMemcpyAsync(coords + (size_t)m * r->nnz, r->coords + (size_t)m * r->coord_stride,
                                   4ull * r->nnz, hipMemcpyDeviceToHost, ctx->stream));
        if (vals)
            FCC_HIP(hipMemcpyAsync(vals, r->vals, 8ull * r->nnz, hipMemcpyDeviceToHost, ctx->stream));
        FCC_HIP(hipStreamSynchronize(ctx->stream));
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where);
    }
    return FCC_OK;
}

int fcc_result_as_tensor(fcc_ctx *ctx, fcc_result *r, fcc_tensor *t) {
    if (!ctx) return FCC_ERR_INVALID;
    if (!r || !t) return fail(ctx, FCC_ERR_INVALID, "NULL result or tensor");
    if (r->nnz > 0 && (!r->coords || !r->vals))
        return fail(ctx, FCC_ERR_INVALID, "result holds no materialised output");
    if (r->order < 1) return fail(ctx, FCC_ERR_INVALID, "an order-0 result is no operand");
    try {
        if (r->nnz > 0 && r->coord_stride != r->nnz) {
            ResultPriv *rp = (ResultPriv *)r->priv;
            if (!rp || rp->block) return fail(ctx, FCC_ERR_INVALID, "a view with a padded stride cannot be compacted");
            FCC_HIP(hipSetDevice(ctx->device));
            uint32_t *nc = (uint32_t *)ctx->pool.get(4ull * r->order * r->nnz);
            FCC_HIP(hipMemcpy2DAsync(nc, 4ull * r->nnz, r->coords, 4ull * r->coord_stride, 4ull * r->nnz, r->order,
                                     hipMemcpyDeviceToDevice, ctx->stream));
            FCC_HIP(hipStreamSynchronize(ctx->stream));
            ctx->pool.put(rp->coords);
            rp->coords = nc;
            r->coords = nc;
            r->coord_stride = r->nnz;
        }
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where);
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    }
    t->order = r->order;
    t->nnz = r->nnz;
    t->coords = r->coords;
    t->vals = r->vals;
    t->extents = r->extents;
    t->on_device = 1;
    return FCC_OK;
}

void fcc_result_free(fcc_ctx *ctx, fcc_result *r) {
    if (!r) return;
    ResultPriv *rp = (ResultPriv *)r->priv;
    if (rp && ctx && !rp->block) {
        if (rp->coords) ctx->pool.put(rp->coords);
        if (rp->vals) ctx->pool.put(rp->vals);
    }
    delete rp;
    r->priv = nullptr;
    r->coords = nullptr;
    r->vals = nullptr;
}

int32_t fcc_density_bucket(double d) {
    if (!(d > 0)) return 0;
    int b = (int)std::floor(std::log10(d)) + 6;  // 1e-6 -> 0 ... 1e3 -> 9
    return std::max(0, std::min(11, b));
}

int fcc_tune_set(fcc_ctx *ctx, const fcc_tune_entry *e, int32_t n) {
    if (!ctx || (n > 0 && !e) || n < 0) return FCC_ERR_INVALID;
    ctx->tune.assign(e, e + n);
    return FCC_OK;
}

int fcc_tune_get(fcc_ctx *ctx, fcc_tune_entry *e, int32_t cap, int32_t *n) {
    if (!ctx || !n) return FCC_ERR_INVALID;
    *n = (int32_t)ctx->tune.size();
    for (int32_t i = 0; i < std::min<int32_t>(cap, *n); i++) e[i] = ctx->tune[i];
    return FCC_OK;
}

int64_t fcc_mlp_param_count(void) { return MLP_PARAMS; }

int fcc_mlp_set_params(fcc_ctx *ctx, const float *p) {
    if (!ctx || !p) return FCC_ERR_INVALID;
    ctx->mlp_host.assign(p, p + MLP_PARAMS);
    if (hipMemcpy(ctx->mlp_dev, p, sizeof(float) * MLP_PARAMS, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, "mlp upload");
    }
    ctx->mlp_ready = true;
    return FCC_OK;
}

int fcc_mlp_get_params(fcc_ctx *ctx, float *p) {
    if (!ctx || !p) return FCC_ERR_INVALID;
    if (!ctx->mlp_ready) return fail(ctx, FCC_ERR_INVALID, "no MLP parameters");
    if (hipMemcpy(p, ctx->mlp_dev, sizeof(float) * MLP_PARAMS, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, "mlp download");
    }
    return FCC_OK;
}

int64_t fcc_tile_stats_count(fcc_ctx *ctx) { return ctx ? (int64_t)ctx->labels.size() / 2 : 0; }

int fcc_tile_stats_copy(fcc_ctx *ctx, float *features, float *labels, int64_t n) {
    if (!ctx || n < 0 || n > fcc_tile_stats_count(ctx)) return FCC_ERR_INVALID;
    if (features) memcpy(features, ctx->feat.data(), sizeof(float) * n * MLP_IN);
    if (labels) memcpy(labels, ctx->labels.data(), sizeof(float) * n * 2);
    return FCC_OK;
}

int fcc_mlp_forward_backward(fcc_ctx *ctx, const float *features, const float *labels, int64_t n, float *pred,
                             float *loss, float *grad) {
    if (!ctx || !features || !labels || n <= 0 || !ctx->mlp_ready) return FCC_ERR_INVALID;
    try {
        FCC_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        Tmp<float> f(ctx->pool, (size_t)n * MLP_IN), l(ctx->pool, (size_t)n * MLP_OUT), g(ctx->pool, MLP_PARAMS),
            ls(ctx->pool, 1), pr(ctx->pool, (size_t)n * MLP_OUT);
        FCC_HIP(hipMemcpyAsync(f.get(), features, 4ull * n * MLP_IN, hipMemcpyHostToDevice, s));
        FCC_HIP(hipMemcpyAsync(l.get(), labels, 4ull * n * MLP_OUT, hipMemcpyHostToDevice, s));
        mlp_fwd_bwd(ctx->mlp_dev, f, l, nullptr, (uint32_t)n, g, ls, pr, s);
        std::vector<float> gh(MLP_PARAMS, 0.f);
        FCC_HIP(hipMemcpyAsync(gh.data(), g.get(), 4ull * mlp_trainable_count(), hipMemcpyDeviceToHost, s));
        if (pred) FCC_HIP(hipMemcpyAsync(pred, pr.get(), 4ull * n * MLP_OUT, hipMemcpyDeviceToHost, s));
        if (loss) FCC_HIP(hipMemcpyAsync(loss, ls.get(), 4, hipMemcpyDeviceToHost, s));
        FCC_HIP(hipStreamSynchronize(s));
        if (grad) memcpy(grad, gh.data(), sizeof(float) * MLP_PARAMS);
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where);
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    }
    return FCC_OK;
}

int fcc_mlp_train(fcc_ctx *ctx, const float *features, const float *labels, int64_t n, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, float *loss_out) {
    if (!ctx || !features || !labels || n <= 0 || epochs <= 0 || batch <= 0) return FCC_ERR_INVALID;
    try {
        FCC_HIP(hipSetDevice(ctx->device));
        hipStream_t s = ctx->stream;
        // normalisation from the training set (host), then Xavier-uniform init
        std::vector<float> P(MLP_PARAMS, 0.f);
        const int OFF_IM = MLP_PARAMS - 2 * (MLP_IN + MLP_OUT);
        std::mt19937_64 rng(seed);
        for (int i = 0; i < MLP_IN; i++) {
            double m = 0, v = 0;
            for (int64_t k = 0; k < n; k++) m += features[k * MLP_IN + i];
            m /= n;
            for (int64_t k = 0; k < n; k++) v += (features[k * MLP_IN + i] - m) * (features[k * MLP_IN + i] - m);
            v = std::sqrt(v / n);
            P[OFF_IM + i] = (float)m;
            P[OFF_IM + MLP_IN + i] = (float)(v > 1e-3 ? v : 1.0);
        }
        for (int o = 0; o < MLP_OUT; o++) {
            double m = 0, v = 0;
            for (int64_t k = 0; k < n; k++) m += labels[k * MLP_OUT + o];
            m /= n;
            for (int64_t k = 0; k < n; k++) v += (labels[k * MLP_OUT + o] - m) * (labels[k * MLP_OUT + o] - m);
            v = std::sqrt(v / n);
            P[OFF_IM + 2 * MLP_IN + o] = (float)m;
            P[OFF_IM + 2 * MLP_IN + MLP_OUT + o] = (float)(v > 1e-3 ? v : 1.0);
        }
        auto init = [&](int off, int fan_in, int fan_out) {
            float lim = std::sqrt(6.f / (fan_in + fan_out));
            std::uniform_real_distribution<float> U(-lim, lim);
            for (int i = 0; i < fan_in * fan_out; i++) P[off + i] = U(rng);
        };
        const int W1 = 0, B1 = W1 + MLP_IN * MLP_HID, W2 = B1 + MLP_HID, B2 = W2 + MLP_HID * MLP_HID,
                  W3 = B2 + MLP_HID;
        init(W1, MLP_IN, MLP_HID);
        init(W2, MLP_HID, MLP_HID);
        init(W3, MLP_HID, MLP_OUT);
        FCC_HIP(hipMemcpyAsync(ctx->mlp_dev, P.data(), 4ull * MLP_PARAMS, hipMemcpyHostToDevice, s));
        Tmp<float> f(ctx->pool, (size_t)n * MLP_IN), l(ctx->pool, (size_t)n * MLP_OUT), g(ctx->pool, MLP_PARAMS),
            ls(ctx->pool, 1), m1(ctx->pool, MLP_PARAMS), m2(ctx->pool, MLP_PARAMS);
        Tmp<uint32_t> perm(ctx->pool, (size_t)n);
        FCC_HIP(hipMemcpyAsync(f.get(), features, 4ull * n * MLP_IN, hipMemcpyHostToDevice, s));
        FCC_HIP(hipMemcpyAsync(l.get(), labels, 4ull * n * MLP_OUT, hipMemcpyHostToDevice, s));
        std::vector<float> zeros(MLP_PARAMS, 0.f);
        FCC_HIP(hipMemcpyAsync(m1.get(), zeros.data(), 4ull * MLP_PARAMS, hipMemcpyHostToDevice, s));
        FCC_HIP(hipMemcpyAsync(m2.get(), zeros.data(), 4ull * MLP_PARAMS, hipMemcpyHostToDevice, s));
        std::vector<uint32_t> ph((size_t)n);
        for (int64_t i = 0; i < n; i++) ph[i] = (uint32_t)i;
        int step = 0;
        float last = 0.f;
        for (int e = 0; e < epochs; e++) {
            std::shuffle(ph.begin(), ph.end(), rng);
            FCC_HIP(hipMemcpyAsync(perm.get(), ph.data(), 4ull * n, hipMemcpyHostToDevice, s));
            for (int64_t b0 = 0; b0 < n; b0 += batch) {
                uint32_t bn = (uint32_t)std::min<int64_t>(batch, n - b0);
                mlp_fwd_bwd(ctx->mlp_dev, f, l, perm.get() + b0, bn, g, ls, nullptr, s);
                mlp_adam(ctx->mlp_dev, g, m1, m2, lr, ++step, s);
            }
        }
        // final full-set loss
        mlp_fwd_bwd(ctx->mlp_dev, f, l, nullptr, (uint32_t)n, g, ls, nullptr, s);
        FCC_HIP(hipMemcpyAsync(&last, ls.get(), 4, hipMemcpyDeviceToHost, s));
        FCC_HIP(hipStreamSynchronize(s));
        if (loss_out) *loss_out = last;
        ctx->mlp_host.resize(MLP_PARAMS);
        FCC_HIP(hipMemcpy(ctx->mlp_host.data(), ctx->mlp_dev, 4ull * MLP_PARAMS, hipMemcpyDeviceToHost));
        ctx->mlp_ready = true;
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where);
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    }
    return FCC_OK;
}

int fcc_contract_batch(fcc_ctx *ctx, int32_t n, const fcc_tensor *a, const int32_t *a_modes, const fcc_tensor *b,
                       const int32_t *b_modes, int32_t n_contr, const fcc_options *options, fcc_result *outs) {
    if (!ctx) return FCC_ERR_INVALID;
    if (n < 0 || (n > 0 && (!a || !b || !outs))) return fail(ctx, FCC_ERR_INVALID, "bad batch arguments");
    for (int32_t i = 0; i < n; i++) memset(&outs[i], 0, sizeof(fcc_result));
    if (n == 0) return FCC_OK;
    std::string msg;
    std::vector<int32_t> ka, kb;
    int rc = FCC_OK;
    for (int32_t i = 0; i < n && !rc; i++) {
        rc = validate(&a[i], "a[i]", msg);
        if (!rc) rc = validate(&b[i], "b[i]", msg);
        if (!rc && (a[i].order != a[0].order || b[i].order != b[0].order)) {
            msg = "batch pairs differ in order";
            rc = FCC_ERR_INVALID;
        }
    }
    if (!rc) rc = check_modes(&a[0], &b[0], a_modes, b_modes, n_contr, nullptr, nullptr, 0, ka, kb, msg);
    if (rc) return fail(ctx, rc, msg);
    const int dA = a[0].order, dB = b[0].order, dC = dA + dB - 2 * n_contr;
    if (dA + 1 > FCC_MAX_ORDER || dB + 1 > FCC_MAX_ORDER || dC + 1 > FCC_MAX_ORDER) {
        // no room for the pair-index mode: one call per pair
        for (int32_t i = 0; i < n; i++) {
            rc = fcc_contract(ctx, &a[i], a_modes, &b[i], b_modes, n_contr, options, &outs[i]);
            if (rc) {
                for (int32_t k = 0; k < i; k++) fcc_result_free(ctx, &outs[k]);
                return rc;
            }
        }
        return FCC_OK;
    }
    fcc_options o;
    fcc_options_default(&o);
    if (options) o = *options;
    if (o.left_tile_stride < 1) o.left_tile_stride = 1;
    const bool sink = o.sink == FCC_SINK_COUNT_CHECKSUM;
    DevPool &pool = ctx->pool;
    hipStream_t s = ctx->stream;
    try {
        FCC_HIP(hipSetDevice(ctx->device));
        const std::vector<uint32_t> ends = chunk_pairs(ctx, n, a, b, a_modes, b_modes, n_contr);
        for (size_t ci = 0; ci < ends.size(); ci++) {
            const int32_t c0 = ci ? (int32_t)ends[ci - 1] : 0;
            const uint32_t cnt = ends[ci] - (uint32_t)c0;
            const fcc_tensor *pa = a + c0, *pb = b + c0;
            fcc_result *po = outs + c0;
            // ---- concatenate each operand with a trailing pair-index mode
            FusedSide side[2];
            fuse_side(ctx, cnt, pa, dA, side[0]);
            fuse_side(ctx, cnt, pb, dB, side[1]);
            // ---- one batch-mode contraction over all pairs
            fcc_tensor ta{dA + 1, (int64_t)side[0].off[cnt], side[0].c.get(), side[0].v.get(), side[0].E, 1};
            fcc_tensor tb{dB + 1, (int64_t)side[1].off[cnt], side[1].c.get(), side[1].v.get(), side[1].E, 1};
            std::vector<int32_t> fa{dA}, fb{dB};
            fa.insert(fa.end(), a_modes, a_modes + n_contr);
            fb.insert(fb.end(), b_modes, b_modes + n_contr);
            fcc_result fused;
            memset(&fused, 0, sizeof(fused));
            struct FusedGuard {
                fcc_ctx *ctx;
                fcc_result *r;
                ~FusedGuard() { fcc_result_free(ctx, r); }
            } guard{ctx, &fused};
            Contraction c;
            c.ctx = ctx;
            c.s = s;
            c.o = o;
            c.out = &fused;
            c.groups = cnt;
            c.run(&ta, fa.data(), &tb, fb.data(), n_contr + 1, 1);
            side[0].c.reset();
            side[0].v.reset();
            side[1].c.reset();
            side[1].v.reset();
            // ---- per-pair results
            for (uint32_t i = 0; i < cnt; i++) {
                fcc_result &r = po[i];
                r.order = dC;
                for (int m = 0; m < dC; m++) r.extents[m] = fused.extents[m + 1];
                r.stats = fused.stats;
                r.stats.ms_done = c.group_ms.empty() ? fused.stats.ms_total : c.group_ms[i];
            }
            if (sink) {
                for (uint32_t i = 0; i < cnt && !c.group_sink.empty(); i++) {
                    const unsigned long long *g4 = &c.group_sink[4 * (size_t)i];
                    po[i].nnz = (int64_t)g4[0];
                    po[i].coord_hash = g4[0] ? g4[1] : 0;
                    memcpy(&po[i].val_sum, &g4[2], 8);
                    memcpy(&po[i].val_wsum, &g4[3], 8);
                }
            } else if (fused.nnz > 0) {
                const uint64_t fn = (uint64_t)fused.nnz;
                Tmp<unsigned long long> counts(pool, cnt), offs(pool, cnt), cursor(pool, cnt);
                fill_u64(counts, cnt, 0, s);
                fill_u64(cursor, cnt, 0, s);
                split_count(fused.coords, fn, cnt, counts, s);
                std::vector<unsigned long long> hc(cnt), ho(cnt);
                FCC_HIP(hipMemcpyAsync(hc.data(), counts.get(), 8ull * cnt, hipMemcpyDeviceToHost, s));
                FCC_HIP(hipStreamSynchronize(s));
                unsigned long long acc = 0;
                for (uint32_t i = 0; i < cnt; i++) {
                    ho[i] = acc;
                    acc += hc[i];
                }
                FCC_HIP(hipMemcpyAsync(offs.get(), ho.data(), 8ull * cnt, hipMemcpyHostToDevice, s));
                auto blk = std::make_shared<SharedBlock>();
                blk->pool = &pool;
                blk->coords = dC ? pool.get(4ull * dC * fn) : nullptr;
                blk->vals = pool.get(8ull * fn);
                split_scatter(fused.coords, (uint64_t)fused.coord_stride, fused.vals, fn, dC + 1, cnt, offs, counts,
                              cursor, (uint32_t *)blk->coords, (double *)blk->vals, s);
                for (uint32_t i = 0; i < cnt; i++) {
                    fcc_result &r = po[i];
                    r.nnz = (int64_t)hc[i];
                    r.coord_stride = (int64_t)hc[i];
                    r.coords = dC ? (uint32_t *)blk->coords + (size_t)dC * ho[i] : nullptr;
                    r.vals = (double *)blk->vals + ho[i];
                    ResultPriv *rp = new ResultPriv;
                    rp->block = blk;
                    r.priv = rp;
                    if (o.summarize && hc[i]) {
                        Tmp<unsigned long long> hs(pool, 1);
                        Tmp<double> vs(pool, 2);
                        Tmp<uint32_t> E(pool, kMaxOrder);
                        fill_u64(hs, 1, 0, s);
                        fill_f64(vs, 2, 0.0, s);
                        FCC_HIP(hipMemcpyAsync(E.get(), r.extents, 4 * kMaxOrder, hipMemcpyHostToDevice, s));
                        k_summarize<<<grid256(hc[i]), 256, 0, s>>>(r.coords, r.vals, hc[i], hc[i], dC, E, hs, vs);
                        double dv[2];
                        FCC_HIP(hipMemcpyAsync(&r.coord_hash, hs.get(), 8, hipMemcpyDeviceToHost, s));
                        FCC_HIP(hipMemcpyAsync(dv, vs.get(), 16, hipMemcpyDeviceToHost, s));
                        FCC_HIP(hipStreamSynchronize(s));
                        r.val_sum = dv[0];
                        r.val_wsum = dv[1];
                    }
                }
                FCC_HIP(hipStreamSynchronize(s));
            }
        }
        ctx->err.clear();
        return FCC_OK;
    } catch (HipError &e) {
        (void)hipGetLastError();
        rc = fail(ctx, FCC_ERR_HIP, e.where + ": " + hipGetErrorString(e.code));
    } catch (LibError &e) {
        rc = fail(ctx, e.code, e.msg);
    } catch (std::bad_alloc &) {
        rc = fail(ctx, FCC_ERR_OOM, "host allocation failed");
    }
    for (int32_t k = 0; k < n; k++) fcc_result_free(ctx, &outs[k]);
    return rc;
}

// ------------------------------------------------------------- autotuner

int fcc_autotune(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const fcc_tensor *b,
                 const int32_t *b_modes, int32_t n_contr, const fcc_options *options, int32_t reps,
                 fcc_tune_entry *best) {
    if (!ctx) return FCC_ERR_INVALID;
    if (!best) return fail(ctx, FCC_ERR_INVALID, "best is NULL");
    std::string msg;
    int rc = validate(a, "a", msg);
    if (!rc) rc = validate(b, "b", msg);
    std::vector<int32_t> ka, kb;
    if (!rc) rc = check_modes(a, b, a_modes, b_modes, n_contr, nullptr, nullptr, 0, ka, kb, msg);
    if (rc) return fail(ctx, rc, msg);
    fcc_options base;
    fcc_options_default(&base);
    if (options) base = *options;
    if (base.left_tile_stride < 1) base.left_tile_stride = 1;
    reps = std::max(1, reps);
    // time one candidate: the fastest of reps full contractions (ms_total)
    auto trial = [&](int32_t tl, int32_t w, int32_t wg, int32_t *bucket, int32_t *tl_used, int32_t *w_used,
                     uint64_t *dense_units) -> double {
        double bestms = 1e300;
        for (int i = 0; i < reps; i++) {
            Contraction c;
            c.ctx = ctx;
            c.s = ctx->stream;
            c.o = base;
            c.o.tile_hint = tl;
            c.o.col_window = w;
            c.o.workgroups = wg;
            fcc_result r;
            memset(&r, 0, sizeof(r));
            c.out = &r;
            c.run(a, ka.data(), b, kb.data(), n_contr, 0);
            bestms = std::min(bestms, r.stats.ms_total);
            *bucket = c.bucket;
            *tl_used = r.stats.row_tile;
            *w_used = r.stats.col_window;
            *dense_units = r.stats.units_dense;
            fcc_result_free(ctx, &r);
        }
        return bestms;
    };
    try {
        FCC_HIP(hipSetDevice(ctx->device));
        fcc_tune_entry win{0, 0, 0, 0, 1e300};
        uint64_t win_dense = 0;
        int32_t bucket = 0;
        // rows per tile: every power of two the dense tile can hold a column of
        for (int32_t tl = 1; tl <= 1024; tl *= 2) {
            int32_t tu, wu;
            uint64_t du;
            double ms = trial(tl, 0, 0, &bucket, &tu, &wu, &du);
            if (ms < win.ms) {
                win = fcc_tune_entry{bucket, tu, 0, 0, ms};
                win_dense = du;
            }
        }
        // dense windows narrower than the widest that fits, then the grid size
        if (win_dense) {
            const int32_t wmax = (int32_t)(DENSE_SLOTS / std::max(1, win.row_tile)) & ~31;
            for (int32_t w : {wmax / 2, wmax / 4}) {
                if (w < 32) continue;
                int32_t tu, wu;
                uint64_t du;
                double ms = trial(win.row_tile, w & ~31, 0, &bucket, &tu, &wu, &du);
                if (ms < win.ms) win = fcc_tune_entry{bucket, win.row_tile, wu, 0, ms};
            }
        }
        for (int32_t f : {1, 3, 4}) {
            int32_t tu, wu;
            uint64_t du;
            double ms = trial(win.row_tile, win.col_window, f * ctx->num_cus, &bucket, &tu, &wu, &du);
            if (ms < win.ms) win = fcc_tune_entry{bucket, win.row_tile, win.col_window, f * ctx->num_cus, ms};
        }
        win.bucket = bucket;
        bool found = false;
        for (auto &e : ctx->tune)
            if (e.bucket == win.bucket) {
                e = win;
                found = true;
            }
        if (!found) ctx->tune.push_back(win);
        *best = win;
        ctx->err.clear();
        return FCC_OK;
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where + ": " + hipGetErrorString(e.code));
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    } catch (std::bad_alloc &) {
        return fail(ctx, FCC_ERR_OOM, "host allocation failed");
    }
}

// ---------------------------------------------------------------- plans

struct fcc_plan {
    fcc_ctx *ctx = nullptr;
    std::unique_ptr<Contraction> c;
    bool self_values = false;  // b's values are a's (the same tensor on both sides)
    int64_t nA = 0, nB = 0;
    Tmp<double> a_stage, b_stage;  // value staging: the graph reads fixed addresses
    // one graph per accumulate variant (dense slots marked empty by -0.0, or
    // by a bitmap when a value is zero or tiny)
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    Tmp<uint32_t> vflag;  // value flags of an execution's inputs
    fcc_stats stats{};
    fcc_plan_info info{};
    int32_t order = 0;
    uint32_t extents[FCC_MAX_ORDER] = {};
    // batch plans (fcc_plan_create_batch): the pairs and the fixed per-pair
    // split of materialised outputs (counts and offsets do not change with
    // the values)
    uint32_t pairs = 0;
    int32_t pair_order = 0;
    std::vector<unsigned long long> pair_nnz, pair_off;
    Tmp<unsigned long long> split_counts, split_offs, split_cursor;
    Tmp<uint32_t> split_c;
    Tmp<double> split_v;
    ~fcc_plan() {
        for (int i = 0; i < 2; i++) {
            if (exec[i]) (void)hipGraphExecDestroy(exec[i]);
            if (graph[i]) (void)hipGraphDestroy(graph[i]);
        }
    }
};

namespace {

// The numeric pass of a plan on the values staged at asrc / bsrc: gather the
// values into index order, clear the counters and the global hash with fill
// kernels (no memset nodes: the sequence may be captured into a graph), run
// the persistent accumulate kernel and the global-hash drain.
__global__ void k_stamp_clock(unsigned long long *dst) { *dst = wall_clock64(); }

// flag |= 1 when a value is zero or below 2^-511 in magnitude (a product of
// two could be +-0), |= 2 when a value is not finite
__global__ __launch_bounds__(256) void k_value_flags(const double *v, uint64_t n, uint32_t *flag) {
    uint32_t f = 0;
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256) {
        const double x = v[i];
        if (!(fabs(x) >= 0x1.0p-511) && !isnan(x)) f |= 1u;
        if (!isfinite(x)) f |= 2u;
    }
    const uint64_t b1 = __ballot(f & 1u), b2 = __ballot(f & 2u);
    if ((threadIdx.x & 63) == 0 && (b1 | b2)) atomicOr(flag, (b1 ? 1u : 0u) | (b2 ? 2u : 0u));
}

// A batch plan also stamps the device clock (the pairs' completion times are
// taken against it), re-arms the per-pair unit counters, and after the drain
// splits the fused output into the per-pair blocks.
void plan_sequence(fcc_plan *P, const double *asrc, const double *bsrc, bool bitmap, hipStream_t s) {
    Contraction::Kept &k = P->c->k;
    AccArgs &X = k.X;
    if (P->pairs) {
        k_stamp_clock<<<1, 1, 0, s>>>(k.g_done.get() + P->pairs);
        fill_u32(k.g_left, P->pairs, 0, s);
        fill_u64(k.g_done, P->pairs, 0, s);
        group_units(X.unit_tile, k.ix.atile_id, X.nunits, k.ix.logTL, X.group_rows, k.g_left, s);
    }
    perm_gather(bsrc, k.ix.permB, k.ix.nB, k.ix.Bval, s);
    perm_gather(asrc, k.ix.permA, k.ix.nAv, k.ix.Aval, s);
    dense_keys_values(k.ix, k.dk, s);
    fill_u64(X.gkeys, X.gmask + 1, ~0ull, s);
    fill_f64(X.gvals, X.gmask + 1, 0.0, s);
    fill_u64(k.gmeta, 4, 0, s);
    fill_u32(k.flags, 4, 0, s);
    fill_u32(k.work, 4, 0, s);
    fill_u64(k.sinkbuf, k.nsink, 0, s);
    launch_numeric(X, k.grid, bitmap, k.sink, 0, s);
    launch_gdrain(X, k.sink, 0, s);
    if (P->pairs && P->split_v.get()) {
        fill_u64(P->split_cursor, P->pairs, 0, s);
        split_scatter(k.out_c.get(), k.stride, k.out_v.get(), k.nnz_out, P->pair_order + 1, P->pairs, P->split_offs,
                      P->split_counts, P->split_cursor, P->split_c.get(), P->split_v, s);
    }
}

// Stages new values (host values, or any values of a graph execution, go
// through the plan's fixed buffers), runs the sequence -- eagerly, or as the
// plan's graph, captured on first use -- and reads back the flags, the sink
// words and (batch plans) the pairs' completion clocks. st: device times.
void run_plan(fcc_plan *plan, const double *a_vals, const double *b_vals, int32_t on_device, int32_t use_graph,
              std::vector<unsigned long long> &sinkw, std::vector<unsigned long long> &gdone, fcc_stats &st) {
    fcc_ctx *ctx = plan->ctx;
    Contraction::Kept &k = plan->c->k;
    const bool self = plan->self_values;
    hipStream_t s = ctx->stream;
    hipEvent_t *ev = ctx->ev;
    FCC_HIP(hipSetDevice(ctx->device));
    FCC_HIP(hipEventRecord(ev[0], s));
    const double *asrc = a_vals, *bsrc = self ? a_vals : b_vals;
    if (use_graph || !on_device) {
        const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (plan->nA && a_vals != plan->a_stage.get())
            FCC_HIP(hipMemcpyAsync(plan->a_stage.get(), a_vals, 8ull * plan->nA, kind, s));
        asrc = plan->a_stage.get();
        if (self) {
            bsrc = asrc;
        } else {
            if (plan->nB && b_vals != plan->b_stage.get())
                FCC_HIP(hipMemcpyAsync(plan->b_stage.get(), b_vals, 8ull * plan->nB, kind, s));
            bsrc = plan->b_stage.get();
        }
    }
    // The values pick the accumulate variant: without zero or tiny values no
    // sum of products is -0.0, which then marks an empty dense slot (the
    // bitmap variant costs about 10% on sparse-output shapes).
    if (!plan->vflag.get()) plan->vflag = Tmp<uint32_t>(ctx->pool, 1);
    fill_u32(plan->vflag, 1, 0, s);
    const unsigned vgrid = (unsigned)std::min<uint64_t>(1024, std::max<int64_t>(1, (plan->nA + 4095) / 4096));
    if (plan->nA) k_value_flags<<<vgrid, 256, 0, s>>>(asrc, (uint64_t)plan->nA, plan->vflag);
    if (!self && plan->nB) {
        const unsigned g2 = (unsigned)std::min<uint64_t>(1024, std::max<int64_t>(1, (plan->nB + 4095) / 4096));
        k_value_flags<<<g2, 256, 0, s>>>(bsrc, (uint64_t)plan->nB, plan->vflag);
    }
    uint32_t vf = 0;
    FCC_HIP(hipMemcpyAsync(&vf, plan->vflag.get(), 4, hipMemcpyDeviceToHost, s));
    FCC_HIP(hipStreamSynchronize(s));
    if ((vf & 2u) && k.dk.nh)
        throw LibError{FCC_ERR_INVALID, "non-finite values in a plan that multiplies dense keys on the matrix cores "
                                        "(a zero-padded panel would spread them; plan with dense_keys < 0)"};
    const int variant = (vf & 1u) ? 1 : 0;
    FCC_HIP(hipEventRecord(ev[4], s));
    if (use_graph) {
        if (!plan->exec[variant]) {
            hipGraph_t g = nullptr;
            FCC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                plan_sequence(plan, asrc, bsrc, variant == 1, s);
            } catch (...) {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            FCC_HIP(hipStreamEndCapture(s, &g));
            plan->graph[variant] = g;
            FCC_HIP(hipGraphInstantiate(&plan->exec[variant], g, nullptr, nullptr, 0));
        }
        FCC_HIP(hipGraphLaunch(plan->exec[variant], s));
    } else {
        plan_sequence(plan, asrc, bsrc, variant == 1, s);
    }
    FCC_HIP(hipEventRecord(ev[5], s));
    uint32_t hf[2];
    unsigned long long hm = 0;
    FCC_HIP(hipMemcpyAsync(hf, k.flags.get(), 8, hipMemcpyDeviceToHost, s));
    FCC_HIP(hipMemcpyAsync(&hm, k.gmeta.get(), 8, hipMemcpyDeviceToHost, s));
    sinkw.assign(k.nsink, 0);
    FCC_HIP(hipMemcpyAsync(sinkw.data(), k.sinkbuf.get(), 8 * k.nsink, hipMemcpyDeviceToHost, s));
    if (plan->pairs) {
        gdone.assign(plan->pairs + 1, 0);
        FCC_HIP(hipMemcpyAsync(gdone.data(), k.g_done.get(), 8 * gdone.size(), hipMemcpyDeviceToHost, s));
    }
    FCC_HIP(hipStreamSynchronize(s));
    if (hf[0]) throw LibError{FCC_ERR_OVERFLOW, "global hash overflow in the numeric pass"};
    if (hf[1]) throw LibError{FCC_ERR_OVERFLOW, "numeric pass disagreed with the plan"};
    st = plan->stats;
    st.ms_index = 0;
    st.ms_plan = 0;
    st.ms_symbolic = 0;
    st.ms_numeric = ev_ms(ev[4], ev[5]);
    st.ms_numeric_kernel = st.ms_numeric;
    st.ms_overflow = 0;
    st.ms_total = ev_ms(ev[0], ev[5]);
    st.ms_done = st.ms_total;
    st.reruns = 0;
    st.ghash_entries = hm;
}

// After a plan's first contraction: a learned run's per-unit output counts
// become exact offsets, the global hash is sized for what the run put in it,
// the value staging buffers are allocated and the plan summary is filled.
void finish_plan(fcc_plan *P, const fcc_result &first) {
    fcc_ctx *ctx = P->ctx;
    Contraction::Kept &k = P->c->k;
    DevPool &pool = ctx->pool;
    hipStream_t s = ctx->stream;
    if (!k.empty && k.mode == 1) {
        // learned run: its per-unit output counts become exact offsets
        Tmp<unsigned long long> tot(pool, 1);
        scan_u64(pool, (const uint64_t *)k.unit_cur.get(), (uint64_t *)k.unit_off.get(), k.U.n,
                 (uint64_t *)tot.get(), s);
        k.X.unit_off = k.unit_off;
        k.X.capacity = k.sink ? ~0ull : k.stride;
        k.mode = 0;
    }
    if (!k.empty && first.stats.ghash_entries * 2 > k.X.gmask + 1) {
        uint64_t cap = pow2_ceil(first.stats.ghash_entries * 2);
        k.gk = Tmp<unsigned long long>(pool, cap);
        k.gv = Tmp<double>(pool, cap);
        k.gu = Tmp<uint32_t>(pool, cap);
        k.X.gkeys = k.gk;
        k.X.gvals = k.gv;
        k.X.gunit = k.gu;
        k.X.gmask = cap - 1;
    }
    P->a_stage = Tmp<double>(pool, (size_t)std::max<int64_t>(P->nA, 1));
    P->b_stage = Tmp<double>(pool, (size_t)std::max<int64_t>(P->nB, 1));
    // plan summary (exact count, bounds, shape)
    fcc_plan_info &I = P->info;
    I.out_nnz = first.nnz;
    I.products = first.stats.products;
    I.left_tiles = first.stats.left_tiles;
    I.units = first.stats.units;
    I.units_dense = first.stats.units_dense;
    I.units_lhash = first.stats.units_lhash;
    I.units_ghash = first.stats.units_ghash;
    I.row_tile = first.stats.row_tile;
    I.col_window = first.stats.col_window;
    I.order = first.order;
    memcpy(I.extents, first.extents, sizeof(I.extents));
    I.ms_plan = first.stats.ms_total;
    const uint32_t T = k.ix.natiles;
    if (!k.empty && T) {
        std::vector<unsigned long long> mg(T), pr(T);
        FCC_HIP(hipMemcpyAsync(mg.data(), k.ix.st_maxgp, 8ull * T, hipMemcpyDeviceToHost, s));
        FCC_HIP(hipMemcpyAsync(pr.data(), k.ix.st_products, 8ull * T, hipMemcpyDeviceToHost, s));
        FCC_HIP(hipStreamSynchronize(s));
        const double area = (double)k.ix.TL * (double)k.g.spanB;
        for (uint32_t t = 0; t < T; t++) {
            I.nnz_lower_bound += mg[t];
            I.nnz_upper_bound += (uint64_t)std::min<double>((double)pr[t], area);
        }
    }
    P->stats = first.stats;
    P->order = first.order;
    memcpy(P->extents, first.extents, sizeof(P->extents));
}

}  // namespace

int fcc_plan_create(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                    const fcc_tensor *b, const int32_t *b_modes, const int32_t *b_batch, int32_t n_contr,
                    int32_t n_batch, const fcc_options *options, fcc_plan **plan) {
    if (!ctx) return FCC_ERR_INVALID;
    if (!plan) return fail(ctx, FCC_ERR_INVALID, "plan is NULL");
    *plan = nullptr;
    std::string msg;
    int rc = validate(a, "a", msg);
    if (!rc) rc = validate(b, "b", msg);
    std::vector<int32_t> ka, kb;
    if (!rc) rc = check_modes(a, b, a_modes, b_modes, n_contr, a_batch, b_batch, n_batch, ka, kb, msg);
    if (rc) return fail(ctx, rc, msg);
    std::unique_ptr<fcc_plan> P(new fcc_plan);
    P->ctx = ctx;
    P->c.reset(new Contraction);
    Contraction &c = *P->c;
    c.ctx = ctx;
    c.s = ctx->stream;
    fcc_options_default(&c.o);
    if (options) c.o = *options;
    if (c.o.left_tile_stride < 1) c.o.left_tile_stride = 1;
    c.keep = true;
    fcc_result first;
    memset(&first, 0, sizeof(first));
    c.out = &first;
    P->self_values = a == b || (a->coords == b->coords && a->vals == b->vals && a->order == b->order &&
                                a->nnz == b->nnz && a->on_device == b->on_device);
    P->nA = a->nnz;
    P->nB = b->nnz;
    hipStream_t s = ctx->stream;
    try {
        FCC_HIP(hipSetDevice(ctx->device));
        c.run(a, ka.data(), b, kb.data(), n_contr + n_batch, n_batch);
        finish_plan(P.get(), first);
        *plan = P.release();
        ctx->err.clear();
        return FCC_OK;
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where + ": " + hipGetErrorString(e.code));
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    } catch (std::bad_alloc &) {
        return fail(ctx, FCC_ERR_OOM, "host allocation failed");
    }
}

int fcc_plan_info_get(const fcc_plan *plan, fcc_plan_info *info) {
    if (!plan || !info) return FCC_ERR_INVALID;
    *info = plan->info;
    return FCC_OK;
}

int64_t fcc_plan_tile_count(const fcc_plan *plan) {
    if (!plan || plan->c->k.empty) return 0;
    return (int64_t)plan->c->k.ix.natiles;
}

int fcc_plan_tiles(fcc_plan *plan, uint64_t *nnz, uint64_t *products, float *predicted, int64_t n) {
    if (!plan) return FCC_ERR_INVALID;
    fcc_ctx *ctx = plan->ctx;
    Contraction::Kept &k = plan->c->k;
    if (k.empty) return n == 0 ? FCC_OK : fail(ctx, FCC_ERR_INVALID, "plan has no tiles");
    const uint32_t T = k.ix.natiles;
    if (n != (int64_t)T) return fail(ctx, FCC_ERR_INVALID, "n must equal fcc_plan_tile_count");
    hipStream_t s = ctx->stream;
    try {
        if (nnz) {
            // per-unit exact counts: the count pass (exact planner) or the
            // units' emitted outputs of the last numeric pass; under mirroring
            // the band counts taken when the plan was made
            Tmp<unsigned long long> tn(ctx->pool, T);
            if (k.tile_band.get()) {
                FCC_HIP(hipMemcpyAsync(tn.get(), k.tile_band.get(), 8ull * T, hipMemcpyDeviceToDevice, s));
            } else {
                fill_u64(tn, T, 0, s);
                const unsigned long long *uc =
                    plan->c->o.planner == FCC_PLAN_LEARNED ? k.unit_cur.get() : k.unit_count.get();
                k_unit_to_tile<<<grid256(k.U.n), 256, 0, s>>>(uc, k.U.tile, k.U.kind, k.U.n, tn);
                k_direct_tile_count<<<grid256(T), 256, 0, s>>>(k.ix.st_direct, (const uint64_t *)k.ix.st_outd, T,
                                                               k.plan_force, tn);
            }
            FCC_HIP(hipMemcpyAsync(nnz, tn.get(), 8ull * T, hipMemcpyDeviceToHost, s));
            FCC_HIP(hipStreamSynchronize(s));
        }
        if (products) FCC_HIP(hipMemcpyAsync(products, k.ix.st_products, 8ull * T, hipMemcpyDeviceToHost, s));
        if (predicted) {
            if (k.pred.get()) FCC_HIP(hipMemcpyAsync(predicted, k.pred.get(), 4ull * T, hipMemcpyDeviceToHost, s));
            else
                for (uint32_t t = 0; t < T; t++) predicted[t] = -1.0f;
        }
        FCC_HIP(hipStreamSynchronize(s));
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where);
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    }
    return FCC_OK;
}

int fcc_plan_execute(fcc_plan *plan, const double *a_vals, const double *b_vals, int32_t on_device,
                     int32_t use_graph, fcc_result *out) {
    if (!plan) return FCC_ERR_INVALID;
    fcc_ctx *ctx = plan->ctx;
    if (!out) return fail(ctx, FCC_ERR_INVALID, "out is NULL");
    memset(out, 0, sizeof(*out));
    const bool self = plan->self_values;
    if ((plan->nA && !a_vals) || (!self && plan->nB && !b_vals)) return fail(ctx, FCC_ERR_INVALID, "NULL values");
    Contraction::Kept &k = plan->c->k;
    out->order = plan->order;
    memcpy(out->extents, plan->extents, sizeof(out->extents));
    out->stats = plan->stats;
    if (k.empty) return FCC_OK;
    if (plan->pairs) return fail(ctx, FCC_ERR_INVALID, "a batch plan runs with fcc_plan_execute_batch");
    try {
        std::vector<unsigned long long> sk, gd;
        run_plan(plan, a_vals, b_vals, on_device, use_graph, sk, gd, out->stats);
        out->nnz = (int64_t)k.nnz_out;
        out->coord_stride = (int64_t)k.stride;
        if (k.sink) {
            out->nnz = (int64_t)sk[0];
            out->coord_hash = sk[0] ? sk[1] : 0;
            memcpy(&out->val_sum, &sk[2], 8);
            memcpy(&out->val_wsum, &sk[3], 8);
        } else {
            out->coords = k.out_c.get();
            out->vals = k.out_v.get();
        }
        ctx->err.clear();
        return FCC_OK;
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where + ": " + hipGetErrorString(e.code));
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    }
}

int fcc_plan_create_batch(fcc_ctx *ctx, int32_t n, const fcc_tensor *a, const int32_t *a_modes,
                          const fcc_tensor *b, const int32_t *b_modes, int32_t n_contr, const fcc_options *options,
                          fcc_plan **plan) {
    if (!ctx) return FCC_ERR_INVALID;
    if (!plan) return fail(ctx, FCC_ERR_INVALID, "plan is NULL");
    *plan = nullptr;
    if (n <= 0 || !a || !b) return fail(ctx, FCC_ERR_INVALID, "bad batch arguments");
    std::string msg;
    std::vector<int32_t> ka, kb;
    int rc = FCC_OK;
    for (int32_t i = 0; i < n && !rc; i++) {
        rc = validate(&a[i], "a[i]", msg);
        if (!rc) rc = validate(&b[i], "b[i]", msg);
        if (!rc && (a[i].order != a[0].order || b[i].order != b[0].order)) {
            msg = "batch pairs differ in order";
            rc = FCC_ERR_INVALID;
        }
    }
    if (!rc) rc = check_modes(&a[0], &b[0], a_modes, b_modes, n_contr, nullptr, nullptr, 0, ka, kb, msg);
    if (rc) return fail(ctx, rc, msg);
    const int dA = a[0].order, dB = b[0].order, dC = dA + dB - 2 * n_contr;
    if (dA + 1 > FCC_MAX_ORDER || dB + 1 > FCC_MAX_ORDER || dC + 1 > FCC_MAX_ORDER)
        return fail(ctx, FCC_ERR_INVALID, "batch plan: no room for the pair-index mode (orders must stay below "
                                          "FCC_MAX_ORDER)");
    std::unique_ptr<fcc_plan> P(new fcc_plan);
    P->ctx = ctx;
    P->c.reset(new Contraction);
    Contraction &c = *P->c;
    c.ctx = ctx;
    c.s = ctx->stream;
    fcc_options_default(&c.o);
    if (options) c.o = *options;
    if (c.o.left_tile_stride < 1) c.o.left_tile_stride = 1;
    c.keep = true;
    c.groups = (uint32_t)n;
    fcc_result first;
    memset(&first, 0, sizeof(first));
    c.out = &first;
    hipStream_t s = ctx->stream;
    DevPool &pool = ctx->pool;
    try {
        FCC_HIP(hipSetDevice(ctx->device));
        if (chunk_pairs(ctx, n, a, b, a_modes, b_modes, n_contr).size() != 1)
            throw LibError{FCC_ERR_INVALID, "batch plan: the pairs do not fit one fused launch (split the batch)"};
        FusedSide side[2];
        fuse_side(ctx, (uint32_t)n, a, dA, side[0]);
        fuse_side(ctx, (uint32_t)n, b, dB, side[1]);
        fcc_tensor ta{dA + 1, (int64_t)side[0].off[n], side[0].c.get(), side[0].v.get(), side[0].E, 1};
        fcc_tensor tb{dB + 1, (int64_t)side[1].off[n], side[1].c.get(), side[1].v.get(), side[1].E, 1};
        std::vector<int32_t> fa{dA}, fb{dB};
        fa.insert(fa.end(), a_modes, a_modes + n_contr);
        fb.insert(fb.end(), b_modes, b_modes + n_contr);
        P->nA = ta.nnz;
        P->nB = tb.nnz;
        P->self_values = false;
        c.run(&ta, fa.data(), &tb, fb.data(), n_contr + 1, 1);
        side[0].c.reset();
        side[0].v.reset();
        side[1].c.reset();
        side[1].v.reset();
        finish_plan(P.get(), first);
        P->pairs = (uint32_t)n;
        P->pair_order = dC;
        Contraction::Kept &k = c.k;
        // completion clocks: stamped at the start of each execution instead
        k.X.clock_start = nullptr;
        P->pair_nnz.assign(n, 0);
        P->pair_off.assign(n, 0);
        if (!k.empty && !k.sink && first.nnz > 0) {
            // the fused output's per-pair counts are fixed by the coordinates
            const uint64_t fn = (uint64_t)first.nnz;
            P->split_counts = Tmp<unsigned long long>(pool, n);
            P->split_offs = Tmp<unsigned long long>(pool, n);
            P->split_cursor = Tmp<unsigned long long>(pool, n);
            fill_u64(P->split_counts, n, 0, s);
            split_count(first.coords, fn, (uint32_t)n, P->split_counts, s);
            FCC_HIP(hipMemcpyAsync(P->pair_nnz.data(), P->split_counts.get(), 8ull * n, hipMemcpyDeviceToHost, s));
            FCC_HIP(hipStreamSynchronize(s));
            unsigned long long acc = 0;
            for (int32_t i = 0; i < n; i++) {
                P->pair_off[i] = acc;
                acc += P->pair_nnz[i];
            }
            FCC_HIP(hipMemcpyAsync(P->split_offs.get(), P->pair_off.data(), 8ull * n, hipMemcpyHostToDevice, s));
            if (dC) P->split_c = Tmp<uint32_t>(pool, (size_t)dC * fn);
            P->split_v = Tmp<double>(pool, (size_t)fn);
            FCC_HIP(hipStreamSynchronize(s));
        }
        *plan = P.release();
        ctx->err.clear();
        return FCC_OK;
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where + ": " + hipGetErrorString(e.code));
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    } catch (std::bad_alloc &) {
        return fail(ctx, FCC_ERR_OOM, "host allocation failed");
    }
}

int fcc_plan_execute_batch(fcc_plan *plan, const double *a_vals, const double *b_vals, int32_t on_device,
                           int32_t use_graph, fcc_result *outs) {
    if (!plan) return FCC_ERR_INVALID;
    fcc_ctx *ctx = plan->ctx;
    if (!plan->pairs) return fail(ctx, FCC_ERR_INVALID, "not a batch plan (fcc_plan_create_batch)");
    if (!outs) return fail(ctx, FCC_ERR_INVALID, "outs is NULL");
    const uint32_t n = plan->pairs;
    for (uint32_t i = 0; i < n; i++) memset(&outs[i], 0, sizeof(fcc_result));
    if ((plan->nA && !a_vals) || (plan->nB && !b_vals)) return fail(ctx, FCC_ERR_INVALID, "NULL values");
    Contraction::Kept &k = plan->c->k;
    const int dC = plan->pair_order;
    for (uint32_t i = 0; i < n; i++) {
        outs[i].order = dC;
        for (int m = 0; m < dC; m++) outs[i].extents[m] = plan->extents[m + 1];
        outs[i].stats = plan->stats;
    }
    if (k.empty) return FCC_OK;
    try {
        std::vector<unsigned long long> sk, gd;
        fcc_stats st;
        run_plan(plan, a_vals, b_vals, on_device, use_graph, sk, gd, st);
        // a pair's output is complete when its last unit ends (entries that
        // went through the global hash only after the drain)
        const double t0 = ev_ms(ctx->ev[0], ctx->ev[4]);
        const unsigned long long c0 = gd[n];
        for (uint32_t i = 0; i < n; i++) {
            fcc_result &r = outs[i];
            r.stats = st;
            if (!st.ghash_entries && c0 && gd[i] >= c0)
                r.stats.ms_done = t0 + (double)(gd[i] - c0) / (double)ctx->clock_khz;
            if (k.sink) {
                const unsigned long long *g4 = &sk[4 * (size_t)i];
                r.nnz = (int64_t)g4[0];
                r.coord_hash = g4[0] ? g4[1] : 0;
                memcpy(&r.val_sum, &g4[2], 8);
                memcpy(&r.val_wsum, &g4[3], 8);
            } else {
                r.nnz = (int64_t)plan->pair_nnz[i];
                r.coord_stride = r.nnz;
                if (r.nnz) {
                    r.coords = dC ? plan->split_c.get() + (size_t)dC * plan->pair_off[i] : nullptr;
                    r.vals = plan->split_v.get() + plan->pair_off[i];
                }
            }
        }
        ctx->err.clear();
        return FCC_OK;
    } catch (HipError &e) {
        (void)hipGetLastError();
        return fail(ctx, FCC_ERR_HIP, e.where + ": " + hipGetErrorString(e.code));
    } catch (LibError &e) {
        return fail(ctx, e.code, e.msg);
    }
}

void fcc_plan_destroy(fcc_plan *plan) {
    if (!plan) return;
    (void)hipSetDevice(plan->ctx->device);
    delete plan;
}

}  // extern "C"
