/* Deterministic synthetic sparse-tensor generator (host C). */
#ifndef FCC_GEN_H
#define FCC_GEN_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define FCCGEN_MAX_ORDER 8

enum {
    FCCGEN_UNIFORM_PM1 = 0, /* uniform [-1, 1) */
    FCCGEN_UNIFORM_01 = 1,  /* uniform (0, 1] */
    FCCGEN_LOGNORMAL = 2,   /* exp(N(0, 1)) */
    FCCGEN_NORMAL = 3,      /* N(0, 1) */
    FCCGEN_ONES = 4,        /* every value 1.0 */
    FCCGEN_SMALLINT = 5     /* uniform over {-3,-2,-1,1,2,3}: sums stay exact */
};

typedef struct {
    int32_t order;
    uint32_t dims[FCCGEN_MAX_ORDER];
    double zipf_s[FCCGEN_MAX_ORDER]; /* 0 = uniform mode, s > 0 = Zipf(s) */
    int64_t nnz;
    int32_t value_dist;
    uint64_t seed;
} fccgen_spec;

/* Fills coords (SoA, coords[m*nnz + n], caller-allocated order*nnz) and vals
 * (nnz). Returns 0, -1 on an invalid spec, -2 on host allocation failure. */
int fccgen_generate(const fccgen_spec *spec, uint32_t *coords, double *vals);

#ifdef __cplusplus
}
#endif
#endif
