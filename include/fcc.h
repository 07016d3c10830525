/* fcc.h -- C ABI of the MI355X (gfx950) FaSTCC sparse-tensor contraction.
 *
 * Plain pointers and sizes only. Every entry point returns 0 on success or a
 * negative FCC_ERR_* code and never exits the process; fcc_last_error()
 * describes the most recent failure on a context.
 *
 * The contraction computed is the reference's fastcc_multiply
 * (contract.hpp:824-912):
 *     C(x, y) = sum_k A(x, k) * B(k, y)
 * where k runs over the paired contraction coordinates a_modes[i] <-> b_modes[i],
 * x over A's non-contracted modes in ascending order and y over B's. The
 * output keeps every coordinate reached by at least one product (structural
 * support, index_tensors.hpp:1028-1037), exact-zero sums included, and holds
 * A's external coordinates first, then B's. Output order is unspecified (the
 * reference's order depends on its OpenMP schedule).
 */
#ifndef FCC_H
#define FCC_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FCC_MAX_ORDER 8

enum {
    FCC_OK = 0,
    FCC_ERR_INVALID = -1,  /* bad argument, shape or mode list */
    FCC_ERR_OOM = -2,      /* device or host allocation failed */
    FCC_ERR_OVERFLOW = -3, /* a bounded pool was exhausted and could not be grown */
    FCC_ERR_HIP = -4       /* a HIP runtime call failed */
};

enum { FCC_PLAN_EXACT = 0, FCC_PLAN_LEARNED = 1 };
enum { FCC_SINK_MATERIALIZE = 0, FCC_SINK_COUNT_CHECKSUM = 1 };
/* accumulator kinds (per unit of work); FCC_ACC_AUTO lets the planner choose */
enum { FCC_ACC_AUTO = 0, FCC_ACC_DENSE = 1, FCC_ACC_LHASH = 2, FCC_ACC_GHASH = 3 };

/* A sparse tensor in coordinate form. Stands in for the reference's
 * Tensor<double> storage (contract.hpp:25-35, vector<NNZ<double>>). */
typedef struct {
    int32_t order;
    int64_t nnz;
    const uint32_t *coords;   /* SoA: coords[m * nnz + n], m < order */
    const double *vals;       /* nnz values */
    const uint32_t *extents;  /* host array of order extents, or NULL: max+1 per mode,
                                 as Tensor::_infer_shape (contract.hpp:124-144) */
    int32_t on_device;        /* 1: coords/vals are device pointers, 0: host */
} fcc_tensor;

typedef struct {
    int32_t planner;     /* FCC_PLAN_EXACT (count-then-fill) or FCC_PLAN_LEARNED (MLP) */
    int32_t tile_hint;   /* reference tile_size (rows per left tile); 0 = planner/autotuner */
    int32_t sink;        /* FCC_SINK_MATERIALIZE or FCC_SINK_COUNT_CHECKSUM */
    int32_t force_acc;   /* FCC_ACC_AUTO, or force one kind for every unit (testing) */
    int32_t col_window;  /* columns per dense unit; 0 = planner */
    int32_t workgroups;  /* persistent workgroups for accumulate kernels; 0 = planner */
    int32_t collect_stats; /* 1: record per-tile features and exact labels (MLP training) */
    int32_t left_tile_stride; /* 1 (or 0) = all left tiles; s > 1 = only left rows x with
                                 (x / T) % s == 0, T = tile_hint rounded up to a power of
                                 two (128 when 0): the sampled baseline of the largest shapes */
    int32_t summarize;   /* 1: also fill coord_hash / val_sum / val_wsum for a materialised result */
    float dense_keys;    /* contraction keys with nA(k) * nB(k) >= dense_keys * spanA * spanB are
                            contracted as one dense FP64 GEMM (matrix cores) instead of through the
                            tile accumulators; same outputs. 0 = default threshold (0.02, and only
                            when their products average four or more per block cell), > 0 = this
                            threshold unconditionally, < 0 = never */
    int32_t mirror;      /* self-contractions (a == b, same mode lists): the output is symmetric,
                            C(x, y) = C(y, x), so 0 (default) computes each unordered pair once and
                            writes it at both coordinates; -1 computes every output. Same outputs
                            either way. Off with left_tile_stride > 1 or collect_stats. */
    int32_t judge;       /* learned planner with mirroring: 1 = judge the predictions with an extra
                            exact count pass over a sample of about 256 tiles (stats.mlp_mape,
                            mape_tiles); 0 = not (mlp_mape = -1). Without mirroring the tile-level
                            MAPE is always taken from the pass's own per-tile outputs. */
} fcc_options;

typedef struct {
    double ms_total;      /* device time of the whole contraction (hipEvents) */
    double ms_index;      /* extents, linearise, sort, CSR, tiles */
    double ms_plan;       /* tile features, MLP or bounds, unit formation */
    double ms_symbolic;   /* exact count pass + scan (0 for the learned planner) */
    double ms_numeric;    /* accumulate + drain */
    double ms_overflow;   /* global-hash drain and any re-run */
    uint64_t products;    /* multiply-adds: sum_k nnzA(k) * nnzB(k) */
    uint64_t bytes_alg;   /* nnzA(4dA+8) + nnzB(4dB+8) + nnzC(4dC+8) */
    uint64_t left_tiles;  /* active left tiles */
    uint64_t units;       /* (tile, column window) units */
    uint64_t units_dense, units_lhash, units_ghash;
    uint64_t ghash_entries; /* outputs accumulated in the global hash (planned + spilled) */
    uint64_t reruns;      /* learned-planner capacity re-runs */
    int32_t row_tile;     /* rows per left tile used */
    int32_t col_window;   /* columns per dense unit used */
    double mlp_mape;      /* learned planner: tile-level MAPE of predicted nnz vs actual (-1 if n/a) */
    double ms_numeric_kernel; /* average duration of the numeric accumulate kernel */
    double ms_done;       /* time from the call's start until this contraction's output was
                             complete: ms_total for fcc_contract; per pair for fcc_contract_batch
                             (device clock at the pair's last accumulate unit, plus the drain) */
    uint64_t units_direct; /* units of tiles whose entries sit on distinct rows: products written
                              straight out, no accumulator */
    uint64_t dense_keys;   /* contraction keys taken by the dense GEMM (0: none) */
    uint64_t dense_key_products; /* their products (part of products) */
    double ms_dense;       /* the dense-key block: bitsets, presence map, value panels, GEMM
                              (on a second stream, beside the left-tile build) */
    double ms_dense_gemm;  /* the dense GEMM kernel alone */
    uint64_t dense_flops;  /* flops the dense GEMM executes (2 * padded rows * columns * keys; a self-contraction computes the tiles on and above the diagonal) */
    uint64_t dense_block_bytes; /* bytes of the block and presence map the accumulate kernels read */
    uint64_t units_block_counted; /* exact planner: dense units the dense-key block covers
                                     entirely, counted from it without a traversal */
    int32_t mirrored;      /* 1: the symmetric output was computed once per unordered pair */
    uint64_t mape_tiles;   /* tiles mlp_mape is taken over: all, or with mirroring (whose per-tile
                              outputs are not whole-tile counts) a sample of about 256 tiles whose
                              outputs an extra exact count pass counts */
} fcc_stats;

typedef struct {
    int32_t order;      /* dA + dB - 2r (dA + dB - 2r - nb with nb batch pairs) */
    int64_t nnz;        /* output nonzeros */
    uint32_t *coords;   /* device SoA: coords[m * coord_stride + n], library-owned (NULL in sink mode) */
    int64_t coord_stride; /* >= nnz; equals nnz for the exact planner */
    double *vals;       /* device [nnz], library-owned (NULL in sink mode) */
    uint32_t extents[FCC_MAX_ORDER]; /* output extents */
    uint64_t coord_hash; /* sum of mix64(linear output index), mod 2^64 */
    double val_sum;      /* sum of output values */
    double val_wsum;     /* sum of value * u(mix64(linear index)), u in [0,1) */
    fcc_stats stats;
    void *priv;
} fcc_result;

typedef struct fcc_ctx fcc_ctx;

/* Context: one HIP device, one stream, a caching device pool, MLP weights.
 * Stands in for the reference's per-call heaps/thread setup
 * (contract.hpp:847-848, timer.hpp:66-125). One context per host thread. */
int fcc_create(fcc_ctx **ctx, int device);
void fcc_destroy(fcc_ctx *ctx);
const char *fcc_last_error(const fcc_ctx *ctx);
/* the HIP stream (hipStream_t) all work of this context is issued on */
void *fcc_stream(fcc_ctx *ctx);

void fcc_options_default(fcc_options *opt);

/* Device memory of a context: all device temporaries and results come from a
 * caching pool. fcc_set_memory_budget caps what the pool may hold (cached and
 * in use; 0 = no cap, the default): a call that would need more fails with
 * FCC_ERR_OOM and leaves nothing allocated beyond what was held before it.
 * Stands in for the reference's fixed-size per-thread heaps, which exit(1)
 * when exhausted (timer.hpp:76-101). fcc_memory_stats reports the bytes the
 * pool holds and how many of them are in use (results not yet freed, plans). */
int fcc_set_memory_budget(fcc_ctx *ctx, uint64_t bytes);
int fcc_memory_stats(const fcc_ctx *ctx, uint64_t *held, uint64_t *in_use);

/* Replaces Tensor<double>::fastcc_multiply<TileAccumulator<double>,double>(
 *     Tensor<double>& other, CoOrdinate left_contr, CoOrdinate right_contr,
 *     int tile_size) -- contract.hpp:824-827. Here a = *this, b = other,
 * a_modes = left_contr, b_modes = right_contr, options->tile_hint = tile_size. */
int fcc_contract(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const fcc_tensor *b,
                 const int32_t *b_modes, int32_t n_contr, const fcc_options *options, fcc_result *out);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: fcc_contract that allocates nothing. Every device
 * temporary comes from a fixed scratch array of FCC_NOALLOC_SCRATCH_BYTES
 * inside the library (one per device, held until the process ends); a
 * contraction that needs more is rejected with FCC_ERR_INVALID and a message
 * naming the limit. Outputs go to the caller's device buffers: out_coords
 * holds (output order) rows of out_capacity entries (SoA, coord_stride ==
 * out_capacity), out_vals out_capacity values; more outputs than that give
 * FCC_ERR_OVERFLOW with the needed count in out->nnz. In count + checksum
 * mode they may be NULL. The dense-key GEMM and collect_stats are not used
 * here; outputs equal fcc_contract's. out views the caller's buffers
 * (fcc_result_free on it releases nothing). Calls that overlap on one device
 * (from any context) share the scratch and race; the scratch keeps the last
 * call's contents and every call initialises what it reads. */
#define FCC_NOALLOC_SCRATCH_BYTES (64u << 20)
int fcc_contract_noalloc(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const fcc_tensor *b,
                         const int32_t *b_modes, int32_t n_contr, const fcc_options *options, uint32_t *out_coords,
                         double *out_vals, int64_t out_capacity, fcc_result *out);

/* Replaces Tensor<double>::multiply<double,double>(Tensor<double>& other,
 *     CoOrdinate left_contr, CoOrdinate left_batch, CoOrdinate right_contr,
 *     CoOrdinate right_batch) -- contract.hpp:285-367 -- on the FaSTCC tiled
 * path. a_batch[i] <-> b_batch[i] are batch pairs: their coordinates must
 * agree like a contraction pair's, and they lead the output coordinate, which
 * is (batch coords in list order, A's other externals ascending, B's other
 * externals ascending), index_tensors.hpp:318-353. Output order is
 * n_batch + (a->order - n_contr - n_batch) + (b->order - n_contr - n_batch).
 * With n_batch = 0 this is fcc_contract. */
int fcc_contract_ex(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                    const fcc_tensor *b, const int32_t *b_modes, const int32_t *b_batch, int32_t n_contr,
                    int32_t n_batch, const fcc_options *options, fcc_result *out);

/* Many independent contractions (same orders and mode lists) in one call.
 * No reference counterpart: the reference contracts one pair per call
 * (benchmark.cpp:68-73). The pairs run as ONE batch-mode contraction whose
 * batch pair is a pair index appended to both operands, so every pair's tiles
 * share one persistent work-stealing accumulate launch (up to 4096 pairs per
 * launch, fewer when the fused index would exceed the limits of one
 * contraction: the pair index multiplies the left span and the key span;
 * consecutive pairs are taken as many at a time as fit). outs[i] is pair i's result, with the
 * same coordinates and values as fcc_contract on that pair, except that:
 *   - extents are the per-mode maximum over all pairs (the summaries
 *     coord_hash / val_wsum linearise with these extents);
 *   - stats describe the fused launch, and stats.ms_done is the time from the
 *     call's start until pair i's output was complete;
 *   - materialised outputs of one call share one device block, released when
 *     the last of them is freed (fcc_result_free on each, as usual). */
int fcc_contract_batch(fcc_ctx *ctx, int32_t n, const fcc_tensor *a, const int32_t *a_modes,
                       const fcc_tensor *b, const int32_t *b_modes, int32_t n_contr,
                       const fcc_options *options, fcc_result *outs);

/* Copies a materialised result to host SoA arrays (coords: order*nnz, vals: nnz).
 * Stands in for walking the reference's ListTensor nodes (index_tensors.hpp:80-126). */
int fcc_result_to_host(fcc_ctx *ctx, const fcc_result *r, uint32_t *coords, double *vals);
void fcc_result_free(fcc_ctx *ctx, fcc_result *r);
/* Chaining: a materialised result as the operand of a later contraction,
 * without leaving the device (the reference converts through a host Tensor,
 * CompactTensor::to_tensor, contract.hpp:1824-1832). Fills *t with the
 * result's device arrays (on_device = 1, extents = r->extents). A result whose
 * coord_stride exceeds its nnz (learned planner) is first compacted: its
 * coordinates move to a stride of nnz. *t views the result's memory and the
 * r->extents array: valid until fcc_result_free(r). Plan outputs (results of
 * fcc_plan_execute) are views of the plan and must already have
 * coord_stride == nnz. */
int fcc_result_as_tensor(fcc_ctx *ctx, fcc_result *r, fcc_tensor *t);

/* ---- plans: one coordinate pattern, many value sets --------------------- */
/* A plan fixes everything that depends only on the coordinates of a and b:
 * the right operand's index, the left tiles, each unit's accumulator kind and
 * output offsets. fcc_plan_execute then reruns only the numeric pass on new
 * values for the same coordinates, optionally as a captured hipGraph. The
 * reference redoes all of fastcc_multiply per call (contract.hpp:824-912);
 * its symbolic queries (SymbolicTensor::output_shape / count_ops,
 * index_tensors.hpp:287-353; bound_output_nnz_*, contract.hpp:180-281) are
 * answered by fcc_plan_info_get and fcc_plan_tiles. */
typedef struct fcc_plan fcc_plan;

typedef struct {
    int64_t out_nnz;          /* exact output nonzeros */
    uint64_t products;        /* multiply-adds per execution */
    uint64_t left_tiles, units, units_dense, units_lhash, units_ghash;
    int32_t row_tile, col_window;
    int32_t order;
    uint32_t extents[FCC_MAX_ORDER];
    uint64_t nnz_lower_bound; /* sum over tiles of the largest single-key product m*|fiber| */
    uint64_t nnz_upper_bound; /* sum over tiles of min(products, rows * right span) */
    double ms_plan;           /* device time of the planning call (one full contraction) */
} fcc_plan_info;

/* Plans (and performs once) the contraction of fcc_contract_ex. options as
 * there; the learned planner's first run supplies the exact per-unit counts. */
int fcc_plan_create(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                    const fcc_tensor *b, const int32_t *b_modes, const int32_t *b_batch, int32_t n_contr,
                    int32_t n_batch, const fcc_options *options, fcc_plan **plan);
int fcc_plan_info_get(const fcc_plan *plan, fcc_plan_info *info);
/* Per active left tile: exact output nonzeros, products, the MLP's predicted
 * outputs (-1 when the plan used the exact planner). n = fcc_plan_tile_count. */
int64_t fcc_plan_tile_count(const fcc_plan *plan);
int fcc_plan_tiles(fcc_plan *plan, uint64_t *nnz, uint64_t *products, float *predicted, int64_t n);
/* The numeric pass on new values: a_vals holds a->nnz values and b_vals
 * b->nnz, in the entry order given to fcc_plan_create (b_vals is ignored when
 * a and b were the same tensor). on_device: 1 device pointers, 0 host.
 * use_graph: 1 = the pass (value gather, counter resets, accumulate, drain) is
 * captured into a hipGraph on the first such call and replayed afterwards.
 * out is a view of the plan's own output buffers, valid until the next
 * execute or fcc_plan_destroy (fcc_result_free on it releases nothing).
 * Executions of one plan must not overlap. */
int fcc_plan_execute(fcc_plan *plan, const double *a_vals, const double *b_vals, int32_t on_device,
                     int32_t use_graph, fcc_result *out);
/* Batch plans: fcc_contract_batch's n pairs, planned (and performed once) as
 * its one fused launch; every pair's coordinates are then fixed.
 * fcc_plan_execute_batch reruns the numeric pass on new values of all pairs,
 * eagerly or as ONE replayed hipGraph (use_graph: value staging, the
 * accumulate launch, the drain and the per-pair split captured together, no
 * host read-back inside): a_vals holds every pair's A values in pair order
 * (pair i's a[i].nnz values after those of pairs 0..i-1), b_vals likewise.
 * outs[i] (n of them) as fcc_contract_batch's; materialised outputs are views
 * of the plan's buffers, valid until the next execution or fcc_plan_destroy;
 * stats.ms_done is pair i's completion time. All pairs must fit one fused
 * launch (FCC_ERR_INVALID otherwise: split the batch). The reference runs one
 * fastcc_multiply per pair (benchmark.cpp:68-73). fcc_plan_info_get and
 * fcc_plan_tiles describe the fused contraction (the pair index leads the
 * output coordinates); fcc_plan_execute rejects a batch plan. */
int fcc_plan_create_batch(fcc_ctx *ctx, int32_t n, const fcc_tensor *a, const int32_t *a_modes,
                          const fcc_tensor *b, const int32_t *b_modes, int32_t n_contr,
                          const fcc_options *options, fcc_plan **plan);
int fcc_plan_execute_batch(fcc_plan *plan, const double *a_vals, const double *b_vals, int32_t on_device,
                           int32_t use_graph, fcc_result *outs);
/* Destroy plans before their context. */
void fcc_plan_destroy(fcc_plan *plan);

/* ---- learned planner (MLP 6 -> 64 -> 64 -> 2, FP32) -------------------- */
#define FCC_MLP_IN 6
#define FCC_MLP_HID 64
#define FCC_MLP_OUT 2
/* number of floats in the flat parameter vector: W1,b1,W2,b2,W3,b3 plus the
 * feature/label normalisation (mean, std per input, mean, std per output) */
int64_t fcc_mlp_param_count(void);
/* set / get the context's MLP parameters (host arrays of fcc_mlp_param_count floats) */
int fcc_mlp_set_params(fcc_ctx *ctx, const float *params);
int fcc_mlp_get_params(fcc_ctx *ctx, float *params);
/* Tile statistics recorded by the last fcc_contract with collect_stats=1:
 * n rows of FCC_MLP_IN features and FCC_MLP_OUT labels (host arrays). */
int64_t fcc_tile_stats_count(fcc_ctx *ctx);
int fcc_tile_stats_copy(fcc_ctx *ctx, float *features, float *labels, int64_t n);
/* Train the context's MLP on the GPU (hand-written forward/backward kernels,
 * Adam) on host feature/label rows; returns the final loss in *loss. */
int fcc_mlp_train(fcc_ctx *ctx, const float *features, const float *labels, int64_t n, int32_t epochs,
                  int32_t batch, float lr, uint64_t seed, float *loss);
/* One forward + backward on the GPU for a batch: outputs predictions
 * (n*FCC_MLP_OUT), loss, and the gradient of the mean squared loss with
 * respect to the flat parameter vector (fcc_mlp_param_count floats). */
int fcc_mlp_forward_backward(fcc_ctx *ctx, const float *features, const float *labels, int64_t n,
                             float *pred, float *loss, float *grad);

/* ---- autotuner ---------------------------------------------------------- */
/* Entries choose (row_tile, col_window, workgroups) per density bucket;
 * fcc_contract uses them when the options leave those at 0. */
typedef struct {
    int32_t bucket;      /* density bucket index, see fcc_density_bucket */
    int32_t row_tile;
    int32_t col_window;
    int32_t workgroups;
    double ms;           /* measured time of the winning candidate */
} fcc_tune_entry;
int32_t fcc_density_bucket(double products_per_output_cell);
int fcc_tune_set(fcc_ctx *ctx, const fcc_tune_entry *entries, int32_t n);
int fcc_tune_get(fcc_ctx *ctx, fcc_tune_entry *entries, int32_t cap, int32_t *n);
/* The autotuner: times this contraction (the fastest of `reps` runs each)
 * over rows per tile 1, 2, 4, ..., 1024, then narrower dense windows and
 * 1, 3, 4 workgroups per CU for the winner, and records the fastest in the
 * context's table under the contraction's density bucket (replacing an
 * entry of that bucket); *best receives it. Calls that leave tile_hint,
 * col_window and workgroups at 0 then use the entry. No reference
 * counterpart: the reference's scripts sweep tile_size by hand
 * (scripts/chicago01.sh and its siblings). */
int fcc_autotune(fcc_ctx *ctx, const fcc_tensor *a, const int32_t *a_modes, const fcc_tensor *b,
                 const int32_t *b_modes, int32_t n_contr, const fcc_options *options, int32_t reps,
                 fcc_tune_entry *best);

#ifdef __cplusplus
}
#endif
#endif
