/* Deterministic synthetic sparse-tensor generator (host C).
 *
 * Produces the seeded inputs every contraction workload in this project is
 * measured on: unique coordinates drawn per mode from a uniform or a Zipf(s)
 * law, FP64 values from one of a few fixed laws, final order shuffled.
 * The reference has no generator of its own (it reads FROSTT files,
 * read.hpp:58-106); this fills that role for the synthetic shapes.
 *
 * PRNG: xoshiro256** (Blackman & Vigna, 2018) seeded through SplitMix64.
 * Zipf(s): inverse CDF over P(c) proportional to (c+1)^-s, c in [0, dim).
 * Uniqueness: draw, reject a coordinate whose row-major linear index was
 * already drawn, keep drawing until nnz unique coordinates exist.
 * Normal/lognormal values use Box-Muller on two uniforms.
 *
 * Coordinates are written SoA: coords[m * nnz + n].
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "fcc_gen.h"

typedef struct { uint64_t s[4]; } xrng;

static uint64_t splitmix64(uint64_t *x) {
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void xrng_seed(xrng *r, uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) r->s[i] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t xrng_next(xrng *r) {
    uint64_t *s = r->s;
    uint64_t result = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
}

/* uniform double in [0, 1) with 53 random mantissa bits */
static inline double u01(xrng *r) { return (double)(xrng_next(r) >> 11) * 0x1.0p-53; }
/* uniform double in (0, 1] */
static inline double u01_open0(xrng *r) { return (double)((xrng_next(r) >> 11) + 1) * 0x1.0p-53; }

static inline uint32_t uniform_below(xrng *r, uint32_t n) {
    return (uint32_t)(((unsigned __int128)xrng_next(r) * n) >> 64);
}

static uint32_t zipf_draw(const double *cdf, uint32_t n, double u) {
    double target = u * cdf[n - 1];
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (cdf[mid] > target) hi = mid; else lo = mid + 1;
    }
    return lo;
}

static double normal01(xrng *r) {
    double u1 = u01_open0(r), u2 = u01(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586476925 * u2);
}

static double draw_value(xrng *r, int dist) {
    switch (dist) {
    case FCCGEN_UNIFORM_PM1: return 2.0 * u01(r) - 1.0;
    case FCCGEN_UNIFORM_01: return u01_open0(r);
    case FCCGEN_LOGNORMAL: return exp(normal01(r));
    case FCCGEN_NORMAL: return normal01(r);
    case FCCGEN_ONES: return 1.0;
    case FCCGEN_SMALLINT: {
        static const double v[6] = {-3.0, -2.0, -1.0, 1.0, 2.0, 3.0};
        return v[uniform_below(r, 6)];
    }
    default: return 0.0;
    }
}

int fccgen_generate(const fccgen_spec *sp, uint32_t *coords, double *vals) {
    if (!sp || sp->order < 1 || sp->order > FCCGEN_MAX_ORDER || sp->nnz < 0) return -1;
    const int d = sp->order;
    const int64_t nnz = sp->nnz;
    /* the coordinate space must hold nnz unique points and fit 63 bits */
    long double space = 1.0L;
    for (int m = 0; m < d; m++) {
        if (sp->dims[m] == 0) return -1;
        space *= (long double)sp->dims[m];
    }
    if (space >= 9.2e18L || (long double)nnz > space) return -1;
    if (nnz == 0) return 0;

    double *cdf[FCCGEN_MAX_ORDER] = {0};
    for (int m = 0; m < d; m++) {
        if (sp->zipf_s[m] > 0.0) {
            cdf[m] = (double *)malloc(sizeof(double) * sp->dims[m]);
            if (!cdf[m]) goto oom;
            double acc = 0.0;
            for (uint32_t c = 0; c < sp->dims[m]; c++) {
                acc += pow((double)c + 1.0, -sp->zipf_s[m]);
                cdf[m][c] = acc;
            }
        }
    }

    uint64_t cap = 16;
    while (cap < (uint64_t)nnz * 2) cap <<= 1;
    uint64_t *set = (uint64_t *)calloc(cap, sizeof(uint64_t));
    if (!set) goto oom;

    xrng r;
    xrng_seed(&r, sp->seed);
    uint32_t c[FCCGEN_MAX_ORDER];
    int64_t have = 0;
    while (have < nnz) {
        uint64_t lin = 0;
        for (int m = 0; m < d; m++) {
            c[m] = cdf[m] ? zipf_draw(cdf[m], sp->dims[m], u01(&r)) : uniform_below(&r, sp->dims[m]);
            lin = lin * sp->dims[m] + c[m];
        }
        uint64_t key = lin + 1;
        uint64_t h = (key * 0x9E3779B97F4A7C15ULL) >> 17;
        int fresh = 1;
        for (;;) {
            h &= cap - 1;
            if (set[h] == 0) { set[h] = key; break; }
            if (set[h] == key) { fresh = 0; break; }
            h++;
        }
        if (!fresh) continue;
        for (int m = 0; m < d; m++) coords[(int64_t)m * nnz + have] = c[m];
        have++;
    }
    free(set);
    for (int m = 0; m < d; m++) free(cdf[m]);

    for (int64_t i = 0; i < nnz; i++) vals[i] = draw_value(&r, sp->value_dist);

    /* Fisher-Yates over the joint (coords, value) records */
    for (int64_t i = nnz - 1; i > 0; i--) {
        int64_t j = (int64_t)(((unsigned __int128)xrng_next(&r) * (uint64_t)(i + 1)) >> 64);
        if (j == i) continue;
        for (int m = 0; m < d; m++) {
            uint32_t t = coords[(int64_t)m * nnz + i];
            coords[(int64_t)m * nnz + i] = coords[(int64_t)m * nnz + j];
            coords[(int64_t)m * nnz + j] = t;
        }
        double tv = vals[i]; vals[i] = vals[j]; vals[j] = tv;
    }
    return 0;
oom:
    for (int m = 0; m < d; m++) free(cdf[m]);
    return -2;
}
