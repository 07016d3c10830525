/* CPU oracle for the FaSTCC contraction -- TEST INFRASTRUCTURE ONLY.
 *
 * Only the tests, the smoke check and the benchmark's CPU-baseline leg load
 * this library, and only as the reference result / the timed host-core
 * baseline.
 * The product path (the HIP library) never calls it.
 */
#ifndef FCC_ORACLE_H
#define FCC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t order;
    int64_t nnz;
    const uint32_t *coords; /* SoA: coords[m * nnz + n] */
    const double *vals;
    const uint32_t *extents; /* order extents, or NULL: max+1 per mode */
} oc_tensor;

typedef struct {
    int32_t order;     /* output order dA + dB - 2r - nb (nb batch pairs) */
    int64_t nnz;       /* output nonzeros (structural support) */
    uint32_t *coords;  /* SoA [order][nnz], malloc'd; NULL in sink mode */
    double *vals;      /* [nnz], malloc'd; NULL in sink mode */
    double *absum;     /* [nnz] sum of |a*b| over the output's terms (materialised mode) */
    uint32_t *nterms;  /* [nnz] number of terms (products) of each output */
    /* order-independent summaries, filled in both modes */
    uint64_t coord_hash;  /* sum over outputs of mix64(linear output index) mod 2^64 */
    double val_sum;       /* sum of output values */
    double val_wsum;      /* sum of value * u(linear index), u in [0,1) from mix64 */
    uint64_t products;    /* multiply-adds performed */
    double seconds;       /* wall time of the contraction proper (index + tile loop + lists) */
    double index_seconds; /* of which: building both tile indices (not sampled) */
    int32_t threads;      /* OpenMP threads used */
} oc_result;

/* Restatement of Tensor<double>::fastcc_multiply<TileAccumulator<double>,double>
 * (contract.hpp:824-912) with the reference's decode defects corrected.
 *   tile_size      reference tile size; rounded up to a power of two like
 *                  contract.hpp:844-846
 *   nthreads       0 = all hardware threads (contract.hpp:847)
 *   high_precision 1 = accumulate in long double (test reference); 0 = double (baseline)
 *   sample_stride  1 = every left tile; s > 1 = only left tiles i with i % s == 0
 *   sink           0 = materialise SoA output; 1 = count + checksums only
 * Returns 0, or a negative error code. */
int oc_contract_tiled(const oc_tensor *a, const int32_t *a_modes, const oc_tensor *b,
                      const int32_t *b_modes, int32_t n_contr, int32_t tile_size,
                      int32_t nthreads, int32_t high_precision, int32_t sample_stride,
                      int32_t sink, oc_result *out);

/* Batch-mode form of oc_contract_tiled: Tensor::multiply(other, left_contr,
 * left_batch, right_contr, right_batch) (contract.hpp:285-367). A batch pair
 * a_batch[i] <-> b_batch[i] must agree like a contraction pair but is kept in
 * the output: coordinates are (batch coordinates in list order, A's other
 * externals ascending, B's other externals ascending)
 * (index_tensors.hpp:318-353). Restated as the tiled algorithm with the batch
 * coordinates leading both the contraction key and the left external index. */
int oc_contract_tiled_ex(const oc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                         const oc_tensor *b, const int32_t *b_modes, const int32_t *b_batch,
                         int32_t n_contr, int32_t n_batch, int32_t tile_size, int32_t nthreads,
                         int32_t high_precision, int32_t sample_stride, int32_t sink, oc_result *out);

/* Direct statement of the contraction semantics: every (p in A, q in B) pair
 * whose contraction coordinates are equal contributes a_p * b_q at output
 * coordinate (A's external coords, B's external coords). Output sorted
 * lexicographically by coordinate, sums in long double. O(nnzA * nnzB):
 * tiny inputs only. */
int oc_contract_bruteforce(const oc_tensor *a, const int32_t *a_modes, const oc_tensor *b,
                           const int32_t *b_modes, int32_t n_contr, oc_result *out);
/* ... with batch pairs: a pair (p, q) contributes only when its batch
 * coordinates are equal too, at (batch coords, A's other externals, B's). */
int oc_contract_bruteforce_ex(const oc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                              const oc_tensor *b, const int32_t *b_modes, const int32_t *b_batch,
                              int32_t n_contr, int32_t n_batch, oc_result *out);

/* Sort a materialised result lexicographically by coordinate (in place). */
int oc_result_sort(oc_result *r);

/* The mixer used by coord_hash / val_wsum (SplitMix64 finaliser). */
uint64_t oc_mix64(uint64_t x);

void oc_result_free(oc_result *r);

#ifdef __cplusplus
}
#endif
#endif
