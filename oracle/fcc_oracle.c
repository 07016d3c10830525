/* CPU oracle for the FaSTCC contraction -- TEST INFRASTRUCTURE ONLY.
 *
 * Loaded only by tests/, the smoke check and bench.py's CPU-baseline leg, as
 * the reference result the tests compare against and as the timed host-core
 * baseline; never by the product path.
 *
 * oc_contract_tiled restates the reference's hot path,
 * Tensor<double>::fastcc_multiply<TileAccumulator<double>,double>
 * (contract.hpp:824-912), step by step:
 *   - shapes = max coordinate per mode, radix = max+1   (contract.hpp:124-144)
 *   - contraction key: row-major over the contraction list in list order
 *                                                      (coordinate.hpp:723-733)
 *   - external index: row-major over the non-contracted modes, ascending
 *                                                      (coordinate.hpp:767-781)
 *   - tile size rounded up to a power of two           (contract.hpp:844-846,
 *                                                       index_tensors.hpp:476-489)
 *   - TileIndexedTensor: ntiles = ceil(span/T), per tile a map key -> list of
 *     (ext % T, value) in first-appearance order; min(ntiles, hw/2) threads,
 *     every thread scans every nonzero and keeps the tiles it owns
 *                                                      (index_tensors.hpp:506-547)
 *   - both operands indexed concurrently                (contract.hpp:852-862)
 *   - tile-pair loop, collapse(2) schedule(dynamic), per-thread TileAccumulator
 *     with bitmask + active list, probe the right tile per left key
 *                                                      (contract.hpp:876-904,
 *                                                       index_tensors.hpp:994-1065)
 *   - drain: decode (left ext, right ext) to output coordinates and push into a
 *     per-thread bump-arena linked list; lists concatenated in thread order
 *                                                      (index_tensors.hpp:80-126,
 *                                                       coordinate.hpp:402-427,
 *                                                       timer.hpp:66-125)
 * oc_contract_tiled_ex adds the batch pairs of Tensor::multiply
 * (contract.hpp:285-367): each batch pair joins the contraction key ahead of
 * the contraction pairs (the reference indexes both operands by
 * (left_batch, left_contr), contract.hpp:291-295) and leads the output
 * coordinate (index_tensors.hpp:318-353, output_coords built from
 * batch_coords, left_external, right_external at contract.hpp:330-331).
 * Three reference defects are corrected, not reproduced: the drain shift uses
 * log2(T) (the reference uses natural log, index_tensors.hpp:1042); decode
 * radices are the external modes' extents (the reference takes the leading
 * modes of the full shape, contract.hpp:836-843); and a contraction pair's key
 * radix is the larger of the two paired extents so both operands agree
 * (coordinate.hpp:728 uses each operand's own maximum).
 */
#include "fcc_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define OC_MAXD 8

uint64_t oc_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------ geometry */

typedef struct {
    int dA, dB, r, nb, dC;
    uint32_t EA[OC_MAXD], EB[OC_MAXD];
    int ka[OC_MAXD], kb[OC_MAXD]; /* key modes: batch pairs first, then contraction pairs */
    uint64_t R[OC_MAXD];          /* key radix per pair */
    int nxa, nxb;
    int xa[OC_MAXD], xb[OC_MAXD];  /* A: batch modes, then other externals ascending; B: externals ascending */
    uint64_t xaE[OC_MAXD], xbE[OC_MAXD];
    uint64_t spanA, spanB;
} oc_geom;

static int in_list(const int *l, int n, int m) {
    for (int i = 0; i < n; i++)
        if (l[i] == m) return 1;
    return 0;
}

static int extents_of(const oc_tensor *t, uint32_t *E) {
    for (int m = 0; m < t->order; m++) {
        uint32_t mx = 0;
        for (int64_t n = 0; n < t->nnz; n++) { uint32_t c = t->coords[(int64_t)m * t->nnz + n]; if (c > mx) mx = c; }
        if (t->extents) {
            if (t->nnz && mx >= t->extents[m]) return -1;
            E[m] = t->extents[m];
        } else {
            E[m] = t->nnz ? mx + 1 : 0;
        }
    }
    return 0;
}

/* Key = (batch pairs, contraction pairs) row-major, radix per pair the larger
 * paired extent; A's external index = (batch coords, other A modes ascending). */
static int geom_init(oc_geom *g, const oc_tensor *a, const int32_t *am, const int32_t *ab, const oc_tensor *b,
                     const int32_t *bm, const int32_t *bb, int r, int nb) {
    memset(g, 0, sizeof(*g));
    if (!a || !b || a->order < 1 || b->order < 1 || a->order > OC_MAXD || b->order > OC_MAXD) return -1;
    if (r < 0 || nb < 0 || r + nb > a->order || r + nb > b->order) return -1;
    g->dA = a->order; g->dB = b->order; g->r = r + nb; g->nb = nb;
    g->dC = a->order + b->order - 2 * (r + nb) + nb;
    if (g->dC > OC_MAXD) return -1;
    for (int i = 0; i < nb; i++) { g->ka[i] = ab[i]; g->kb[i] = bb[i]; }
    for (int i = 0; i < r; i++) { g->ka[nb + i] = am[i]; g->kb[nb + i] = bm[i]; }
    for (int i = 0; i < g->r; i++) {
        if (g->ka[i] < 0 || g->ka[i] >= a->order || g->kb[i] < 0 || g->kb[i] >= b->order) return -1;
        for (int k = 0; k < i; k++)
            if (g->ka[k] == g->ka[i] || g->kb[k] == g->kb[i]) return -1;
    }
    if (extents_of(a, g->EA) || extents_of(b, g->EB)) return -1;
    for (int i = 0; i < g->r; i++) {
        uint32_t ea = g->EA[g->ka[i]], eb = g->EB[g->kb[i]];
        g->R[i] = ea > eb ? ea : eb;
    }
    g->spanA = 1; g->spanB = 1;
    for (int i = 0; i < nb; i++) { g->xa[g->nxa] = g->ka[i]; g->xaE[g->nxa] = g->R[i]; g->spanA *= g->R[i]; g->nxa++; }
    for (int m = 0; m < a->order; m++)
        if (!in_list(g->ka, g->r, m)) { g->xa[g->nxa] = m; g->xaE[g->nxa] = g->EA[m]; g->spanA *= g->EA[m]; g->nxa++; }
    for (int m = 0; m < b->order; m++)
        if (!in_list(g->kb, g->r, m)) { g->xb[g->nxb] = m; g->xbE[g->nxb] = g->EB[m]; g->spanB *= g->EB[m]; g->nxb++; }
    return 0;
}

static inline uint64_t key_of(const oc_tensor *t, const int *modes, const oc_geom *g, int64_t n) {
    uint64_t k = 0;
    for (int i = 0; i < g->r; i++) k = k * g->R[i] + t->coords[(int64_t)modes[i] * t->nnz + n];
    return k;
}

static inline uint64_t ext_of(const oc_tensor *t, const int *xm, const uint64_t *xE, int nx, int64_t n) {
    uint64_t e = 0;
    for (int i = 0; i < nx; i++) e = e * xE[i] + t->coords[(int64_t)xm[i] * t->nnz + n];
    return e;
}

static inline void decode(uint64_t idx, const uint64_t *E, int nx, uint32_t *out) {
    for (int i = nx - 1; i >= 0; i--) { out[i] = (uint32_t)(idx % E[i]); idx /= E[i]; }
}

static inline void summarise(uint64_t lin, double v, uint64_t *h, double *s, double *ws) {
    uint64_t z = oc_mix64(lin);
    *h += z;
    *s += v;
    *ws += v * ((double)(z >> 11) * 0x1.0p-53);
}

/* --------------------------------------------------------- tile index (P2) */

typedef struct { uint64_t inner; double val; } oc_ev; /* (ext % T, value) pair */

typedef struct {
    uint64_t key;
    uint32_t n, cap;
    oc_ev *ev;
} oc_group;

typedef struct {
    uint32_t ngroups, gcap;
    oc_group *groups;    /* insertion (first-appearance) order */
    uint32_t hcap;
    uint32_t *slots;     /* open addressing: group index + 1, 0 = empty */
} oc_tile;

static inline uint32_t hkey(uint64_t k) { return (uint32_t)(oc_mix64(k) >> 32); }

static oc_group *tile_find(const oc_tile *t, uint64_t key) {
    if (!t->hcap) return NULL;
    uint32_t h = hkey(key) & (t->hcap - 1);
    for (;;) {
        uint32_t s = t->slots[h];
        if (!s) return NULL;
        if (t->groups[s - 1].key == key) return &t->groups[s - 1];
        h = (h + 1) & (t->hcap - 1);
    }
}

static int tile_insert(oc_tile *t, uint64_t key, uint64_t inner, double val) {
    if ((uint64_t)(t->ngroups + 1) * 2 > t->hcap) {
        uint32_t ncap = t->hcap ? t->hcap * 2 : 16;
        uint32_t *ns = (uint32_t *)calloc(ncap, sizeof(uint32_t));
        if (!ns) return -2;
        for (uint32_t gi = 0; gi < t->ngroups; gi++) {
            uint32_t h = hkey(t->groups[gi].key) & (ncap - 1);
            while (ns[h]) h = (h + 1) & (ncap - 1);
            ns[h] = gi + 1;
        }
        free(t->slots);
        t->slots = ns;
        t->hcap = ncap;
    }
    uint32_t h = hkey(key) & (t->hcap - 1);
    oc_group *g = NULL;
    for (;;) {
        uint32_t s = t->slots[h];
        if (!s) break;
        if (t->groups[s - 1].key == key) { g = &t->groups[s - 1]; break; }
        h = (h + 1) & (t->hcap - 1);
    }
    if (!g) {
        if (t->ngroups == t->gcap) {
            uint32_t nc = t->gcap ? t->gcap * 2 : 4;
            oc_group *ng = (oc_group *)realloc(t->groups, sizeof(oc_group) * nc);
            if (!ng) return -2;
            t->groups = ng;
            t->gcap = nc;
        }
        g = &t->groups[t->ngroups];
        g->key = key; g->n = 0; g->cap = 0; g->ev = NULL;
        t->slots[h] = ++t->ngroups;
    }
    if (g->n == g->cap) {
        uint32_t nc = g->cap ? g->cap * 2 : 1;
        oc_ev *ne = (oc_ev *)realloc(g->ev, sizeof(oc_ev) * nc);
        if (!ne) return -2;
        g->ev = ne;
        g->cap = nc;
    }
    g->ev[g->n].inner = inner;
    g->ev[g->n].val = val;
    g->n++;
    return 0;
}

typedef struct {
    uint64_t ntiles;
    uint64_t T;
    oc_tile *tiles;
    int err;
} oc_tindex;

/* index_tensors.hpp:506-547 */
static void tindex_build(oc_tindex *ix, const oc_tensor *t, const oc_geom *g, int left, uint64_t T, int hw) {
    const int *modes = left ? g->ka : g->kb;
    const int *xm = left ? g->xa : g->xb;
    const uint64_t *xE = left ? g->xaE : g->xbE;
    int nx = left ? g->nxa : g->nxb;
    uint64_t span = left ? g->spanA : g->spanB;
    ix->T = T;
    ix->ntiles = span / T + (span % T != 0);
    ix->tiles = (oc_tile *)calloc(ix->ntiles ? ix->ntiles : 1, sizeof(oc_tile));
    ix->err = ix->tiles ? 0 : -2;
    if (ix->err) return;
    uint64_t half = (uint64_t)(hw / 2 > 0 ? hw / 2 : 1);
    int nth = (int)(ix->ntiles < half ? ix->ntiles : half);
    if (nth < 1) nth = 1;
#pragma omp parallel num_threads(nth)
    {
        int tid = omp_get_thread_num();
        int nt = omp_get_num_threads();
        for (int64_t n = 0; n < t->nnz; n++) {
            uint64_t rem = ext_of(t, xm, xE, nx, n);
            uint64_t tile = rem / T;
            if (tile % (uint64_t)nt != (uint64_t)tid) continue;
            uint64_t key = key_of(t, modes, g, n);
            if (tile_insert(&ix->tiles[tile], key, rem % T, t->vals[n]) != 0) ix->err = -2;
        }
    }
}

static void tindex_free(oc_tindex *ix) {
    if (!ix->tiles) return;
    for (uint64_t i = 0; i < ix->ntiles; i++) {
        oc_tile *t = &ix->tiles[i];
        for (uint32_t k = 0; k < t->ngroups; k++) free(t->groups[k].ev);
        free(t->groups);
        free(t->slots);
    }
    free(ix->tiles);
    ix->tiles = NULL;
}

/* ----------------------------------------------- output lists (arena, P3) */

typedef struct oc_node {
    double val;
    double absum;
    uint32_t nterms;
    uint32_t *coords;
    struct oc_node *next;
} oc_node;

typedef struct {
    char *chunk;
    size_t used, size;
    char **old;
    size_t nold, capold;
    oc_node *head, *tail;
    uint64_t count;
    /* sink summaries */
    uint64_t hsum;
    double vsum, wsum;
    uint64_t products;
} oc_tls;

#define OC_CHUNK ((size_t)64 << 20)

static void *arena_alloc(oc_tls *t, size_t n) {
    n = (n + 7) & ~(size_t)7;
    if (!t->chunk || t->used + n > t->size) {
        if (t->chunk) {
            if (t->nold == t->capold) {
                size_t nc = t->capold ? t->capold * 2 : 16;
                char **no = (char **)realloc(t->old, sizeof(char *) * nc);
                if (!no) return NULL;
                t->old = no;
                t->capold = nc;
            }
            t->old[t->nold++] = t->chunk;
        }
        t->chunk = (char *)malloc(OC_CHUNK);
        if (!t->chunk) return NULL;
        t->size = OC_CHUNK;
        t->used = 0;
    }
    void *p = t->chunk + t->used;
    t->used += n;
    return p;
}

static void tls_free(oc_tls *t) {
    for (size_t i = 0; i < t->nold; i++) free(t->old[i]);
    free(t->old);
    free(t->chunk);
    memset(t, 0, sizeof(*t));
}

/* ------------------------------------------------ accumulator (TileAccumulator) */

typedef struct {
    double *d;        /* double accumulator, or NULL */
    long double *q;   /* long double accumulator, or NULL */
    double *ab;       /* sum of |terms| (long-double runs only) */
    uint32_t *nt;     /* number of terms (long-double runs only) */
    uint8_t *bits;
    uint64_t *active;
    uint64_t nact;
} oc_acc;

/* ---------------------------------------------------------------- driver */

int oc_contract_tiled(const oc_tensor *a, const int32_t *a_modes, const oc_tensor *b,
                      const int32_t *b_modes, int32_t n_contr, int32_t tile_size,
                      int32_t nthreads, int32_t high_precision, int32_t sample_stride,
                      int32_t sink, oc_result *out) {
    return oc_contract_tiled_ex(a, a_modes, NULL, b, b_modes, NULL, n_contr, 0, tile_size, nthreads,
                                high_precision, sample_stride, sink, out);
}

int oc_contract_tiled_ex(const oc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                         const oc_tensor *b, const int32_t *b_modes, const int32_t *b_batch,
                         int32_t n_contr, int32_t n_batch, int32_t tile_size, int32_t nthreads,
                         int32_t high_precision, int32_t sample_stride, int32_t sink, oc_result *out) {
    if (!out) return -1;
    memset(out, 0, sizeof(*out));
    oc_geom g;
    if (geom_init(&g, a, a_modes, a_batch, b, b_modes, b_batch, n_contr, n_batch) != 0) return -1;
    out->order = g.dC;
    if (sample_stride < 1) sample_stride = 1;
    int hw = nthreads > 0 ? nthreads : omp_get_num_procs();
    out->threads = hw;

    double t0 = now_s();
    uint64_t T = tile_size < 1 ? 1 : (uint64_t)tile_size;
    if (T & (T - 1)) { uint64_t p = 1; while (p < T) p <<= 1; T = p; }
    unsigned logT = 0;
    while ((1ULL << logT) < T) logT++;

    oc_tindex L = {0}, Rx = {0};
    int nested_prev = omp_get_max_active_levels();
    omp_set_max_active_levels(2);
#pragma omp parallel num_threads(2)
    {
        if (omp_get_thread_num() == 0) tindex_build(&L, a, &g, 1, T, hw);
        else tindex_build(&Rx, b, &g, 0, T, hw);
    }
    omp_set_max_active_levels(nested_prev);
    out->index_seconds = now_s() - t0;
    if (L.err || Rx.err || !L.tiles || !Rx.tiles) { tindex_free(&L); tindex_free(&Rx); return -2; }

    oc_tls *tls = (oc_tls *)calloc(hw, sizeof(oc_tls));
    oc_acc *accs = (oc_acc *)calloc(hw, sizeof(oc_acc));
    int err = (!tls || !accs) ? -2 : 0;
    for (int i = 0; i < hw && !err; i++) {
        if (high_precision) {
            accs[i].q = (long double *)calloc(T * T, sizeof(long double));
            accs[i].ab = (double *)calloc(T * T, sizeof(double));
            accs[i].nt = (uint32_t *)calloc(T * T, sizeof(uint32_t));
            if (!accs[i].ab || !accs[i].nt) err = -2;
        } else {
            accs[i].d = (double *)calloc(T * T, sizeof(double));
        }
        accs[i].bits = (uint8_t *)calloc(T * T / 8 + 1, 1);
        accs[i].active = (uint64_t *)calloc(T * T, sizeof(uint64_t));
        if ((!accs[i].q && !accs[i].d) || !accs[i].bits || !accs[i].active) err = -2;
    }
    const int64_t nsl = (int64_t)((L.ntiles + sample_stride - 1) / sample_stride);
    const int64_t ntr = (int64_t)Rx.ntiles;
    const int dC = g.dC;
    if (!err) {
#pragma omp parallel for collapse(2) schedule(dynamic) num_threads(hw)
        for (int64_t ii = 0; ii < nsl; ii++) {
            for (int64_t j = 0; j < ntr; j++) {
                int64_t i = ii * sample_stride;
                int tid = omp_get_thread_num();
                oc_acc *acc = &accs[tid];
                oc_tls *tl = &tls[tid];
                const oc_tile *lt = &L.tiles[i];
                const oc_tile *rt = &Rx.tiles[j];
                for (uint32_t gi = 0; gi < lt->ngroups; gi++) {
                    const oc_group *lg = &lt->groups[gi];
                    const oc_group *rg = tile_find(rt, lg->key);
                    if (!rg) continue;
                    for (uint32_t x = 0; x < lg->n; x++) {
                        for (uint32_t y = 0; y < rg->n; y++) {
                            uint64_t pos = lg->ev[x].inner * T + rg->ev[y].inner;
                            uint8_t bit = (uint8_t)(1u << (7 - (pos % 8)));
                            if (!(acc->bits[pos / 8] & bit)) {
                                acc->active[acc->nact++] = pos;
                                acc->bits[pos / 8] |= bit;
                            }
                            double p = lg->ev[x].val * rg->ev[y].val;
                            if (acc->q) {
                                acc->q[pos] += (long double)lg->ev[x].val * (long double)rg->ev[y].val;
                                acc->ab[pos] += fabs(p);
                                acc->nt[pos]++;
                            } else {
                                acc->d[pos] += p;
                            }
                        }
                    }
                    tl->products += (uint64_t)lg->n * rg->n;
                }
                /* drain_into (index_tensors.hpp:1039-1060), decode corrected */
                for (uint64_t k = 0; k < acc->nact; k++) {
                    uint64_t pos = acc->active[k];
                    uint64_t li = pos >> logT, ri = pos & (T - 1);
                    uint64_t left = (uint64_t)i * T + li, right = (uint64_t)j * T + ri;
                    double v = acc->q ? (double)acc->q[pos] : acc->d[pos];
                    if (sink) {
                        summarise(left * g.spanB + right, v, &tl->hsum, &tl->vsum, &tl->wsum);
                        tl->count++;
                    } else {
                        oc_node *nd = (oc_node *)arena_alloc(tl, sizeof(oc_node));
                        uint32_t *cc = (uint32_t *)arena_alloc(tl, sizeof(uint32_t) * (dC > 0 ? dC : 1));
                        if (!nd || !cc) { err = -2; }
                        else {
                            decode(left, g.xaE, g.nxa, cc);
                            decode(right, g.xbE, g.nxb, cc + g.nxa);
                            nd->val = v; nd->coords = cc; nd->next = NULL;
                            nd->absum = acc->ab ? acc->ab[pos] : 0.0;
                            nd->nterms = acc->nt ? acc->nt[pos] : 0;
                            if (tl->tail) tl->tail->next = nd; else tl->head = nd;
                            tl->tail = nd;
                            tl->count++;
                        }
                    }
                    if (acc->q) { acc->q[pos] = 0.0L; acc->ab[pos] = 0.0; acc->nt[pos] = 0; }
                    else acc->d[pos] = 0.0;
                    acc->bits[pos / 8] = 0;
                    acc->active[k] = 0;
                }
                acc->nact = 0;
            }
        }
    }
    /* concatenate per-thread lists in thread order (contract.hpp:906-909) */
    oc_node *head = NULL, *tail = NULL;
    uint64_t total = 0;
    if (!err) {
        for (int t = 0; t < hw; t++) {
            if (tls[t].head) {
                if (tail) tail->next = tls[t].head; else head = tls[t].head;
                tail = tls[t].tail;
            }
            total += tls[t].count;
        }
    }
    out->seconds = now_s() - t0;

    if (!err) {
        out->nnz = (int64_t)total;
        for (int t = 0; t < hw; t++) {
            out->coord_hash += tls[t].hsum;
            out->val_sum += tls[t].vsum;
            out->val_wsum += tls[t].wsum;
            out->products += tls[t].products;
        }
        if (!sink) {
            out->coords = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(dC > 0 ? dC : 1) * (total ? total : 1));
            out->vals = (double *)malloc(sizeof(double) * (total ? total : 1));
            out->absum = (double *)malloc(sizeof(double) * (total ? total : 1));
            out->nterms = (uint32_t *)malloc(sizeof(uint32_t) * (total ? total : 1));
            if (!out->coords || !out->vals || !out->absum || !out->nterms) err = -2;
            uint64_t n = 0;
            for (oc_node *nd = head; nd && !err; nd = nd->next, n++) {
                uint64_t l = 0;
                for (int m = 0; m < dC; m++) {
                    out->coords[(uint64_t)m * total + n] = nd->coords[m];
                }
                for (int m = 0; m < g.nxa; m++) l = l * g.xaE[m] + nd->coords[m];
                for (int m = 0; m < g.nxb; m++) l = l * g.xbE[m] + nd->coords[g.nxa + m];
                out->vals[n] = nd->val;
                out->absum[n] = nd->absum;
                out->nterms[n] = nd->nterms;
                summarise(l, nd->val, &out->coord_hash, &out->val_sum, &out->val_wsum);
            }
        }
    }
    for (int i = 0; i < hw && accs; i++) {
        free(accs[i].d); free(accs[i].q); free(accs[i].ab); free(accs[i].nt);
        free(accs[i].bits); free(accs[i].active);
    }
    free(accs);
    for (int i = 0; i < hw && tls; i++) tls_free(&tls[i]);
    free(tls);
    tindex_free(&L);
    tindex_free(&Rx);
    if (err) { oc_result_free(out); return err; }
    return 0;
}

/* ----------------------------------------------------------- brute force */

typedef struct { uint64_t lin; uint64_t seq; double p; } oc_term;

static int term_cmp(const void *x, const void *y) {
    const oc_term *a = (const oc_term *)x, *b = (const oc_term *)y;
    if (a->lin != b->lin) return a->lin < b->lin ? -1 : 1;
    return a->seq < b->seq ? -1 : (a->seq > b->seq);
}

int oc_contract_bruteforce(const oc_tensor *a, const int32_t *a_modes, const oc_tensor *b,
                           const int32_t *b_modes, int32_t n_contr, oc_result *out) {
    return oc_contract_bruteforce_ex(a, a_modes, NULL, b, b_modes, NULL, n_contr, 0, out);
}

int oc_contract_bruteforce_ex(const oc_tensor *a, const int32_t *a_modes, const int32_t *a_batch,
                              const oc_tensor *b, const int32_t *b_modes, const int32_t *b_batch,
                              int32_t n_contr, int32_t n_batch, oc_result *out) {
    if (!out) return -1;
    memset(out, 0, sizeof(*out));
    oc_geom g;
    if (geom_init(&g, a, a_modes, a_batch, b, b_modes, b_batch, n_contr, n_batch) != 0) return -1;
    out->order = g.dC;
    double t0 = now_s();
    size_t cap = 1024, n = 0;
    oc_term *terms = (oc_term *)malloc(sizeof(oc_term) * cap);
    if (!terms) return -2;
    for (int64_t p = 0; p < a->nnz; p++) {
        for (int64_t q = 0; q < b->nnz; q++) {
            int match = 1;
            for (int i = 0; i < g.r && match; i++)
                match = a->coords[(int64_t)g.ka[i] * a->nnz + p] == b->coords[(int64_t)g.kb[i] * b->nnz + q];
            if (!match) continue;
            /* output coordinate: batch coords, A's other externals (ascending),
             * then B's externals; its lexicographic order equals the order of
             * this mixed-radix index */
            uint64_t lin = 0;
            for (int i = 0; i < g.nxa; i++) lin = lin * g.xaE[i] + a->coords[(int64_t)g.xa[i] * a->nnz + p];
            for (int i = 0; i < g.nxb; i++) lin = lin * g.xbE[i] + b->coords[(int64_t)g.xb[i] * b->nnz + q];
            if (n == cap) {
                cap *= 2;
                oc_term *nt = (oc_term *)realloc(terms, sizeof(oc_term) * cap);
                if (!nt) { free(terms); return -2; }
                terms = nt;
            }
            terms[n].lin = lin; terms[n].seq = n; terms[n].p = a->vals[p] * b->vals[q];
            n++;
        }
    }
    out->products = n;
    qsort(terms, n, sizeof(oc_term), term_cmp);
    size_t uniq = 0;
    for (size_t k = 0; k < n; k++)
        if (k == 0 || terms[k].lin != terms[k - 1].lin) uniq++;
    int dC = g.dC;
    out->coords = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(dC > 0 ? dC : 1) * (uniq ? uniq : 1));
    out->vals = (double *)malloc(sizeof(double) * (uniq ? uniq : 1));
    out->absum = (double *)malloc(sizeof(double) * (uniq ? uniq : 1));
    out->nterms = (uint32_t *)malloc(sizeof(uint32_t) * (uniq ? uniq : 1));
    if (!out->coords || !out->vals || !out->absum || !out->nterms) { free(terms); oc_result_free(out); return -2; }
    uint64_t E[2 * OC_MAXD];
    for (int i = 0; i < g.nxa; i++) E[i] = g.xaE[i];
    for (int i = 0; i < g.nxb; i++) E[g.nxa + i] = g.xbE[i];
    size_t o = 0;
    for (size_t k = 0; k < n;) {
        size_t e = k;
        long double s = 0.0L;
        double ab = 0.0;
        while (e < n && terms[e].lin == terms[k].lin) { s += (long double)terms[e].p; ab += fabs(terms[e].p); e++; }
        out->absum[o] = ab;
        out->nterms[o] = (uint32_t)(e - k);
        uint32_t c[2 * OC_MAXD];
        decode(terms[k].lin, E, dC, c);
        for (int m = 0; m < dC; m++) out->coords[(size_t)m * uniq + o] = c[m];
        out->vals[o] = (double)s;
        summarise(terms[k].lin, (double)s, &out->coord_hash, &out->val_sum, &out->val_wsum);
        o++;
        k = e;
    }
    out->nnz = (int64_t)uniq;
    out->seconds = now_s() - t0;
    out->threads = 1;
    free(terms);
    return 0;
}

/* ------------------------------------------------------------- utilities */

static const oc_result *g_sort_r;
static int idx_cmp(const void *x, const void *y) {
    uint64_t i = *(const uint64_t *)x, j = *(const uint64_t *)y;
    const oc_result *r = g_sort_r;
    for (int m = 0; m < r->order; m++) {
        uint32_t a = r->coords[(uint64_t)m * r->nnz + i], b = r->coords[(uint64_t)m * r->nnz + j];
        if (a != b) return a < b ? -1 : 1;
    }
    return 0;
}

int oc_result_sort(oc_result *r) {
    if (!r || !r->coords || r->nnz <= 1) return 0;
    uint64_t n = (uint64_t)r->nnz;
    uint64_t *idx = (uint64_t *)malloc(sizeof(uint64_t) * n);
    uint32_t *nc = (uint32_t *)malloc(sizeof(uint32_t) * n * (r->order > 0 ? r->order : 1));
    double *nv = (double *)malloc(sizeof(double) * n);
    double *na = (double *)malloc(sizeof(double) * n);
    uint32_t *nn = (uint32_t *)malloc(sizeof(uint32_t) * n);
    if (!idx || !nc || !nv || !na || !nn) { free(idx); free(nc); free(nv); free(na); free(nn); return -2; }
    for (uint64_t i = 0; i < n; i++) idx[i] = i;
    g_sort_r = r;
    qsort(idx, n, sizeof(uint64_t), idx_cmp);
    for (uint64_t k = 0; k < n; k++) {
        for (int m = 0; m < r->order; m++) nc[(uint64_t)m * n + k] = r->coords[(uint64_t)m * n + idx[k]];
        nv[k] = r->vals[idx[k]];
        na[k] = r->absum ? r->absum[idx[k]] : 0.0;
        nn[k] = r->nterms ? r->nterms[idx[k]] : 0;
    }
    free(r->coords); free(r->vals); free(r->absum); free(r->nterms); free(idx);
    r->coords = nc;
    r->vals = nv;
    r->absum = na;
    r->nterms = nn;
    return 0;
}

void oc_result_free(oc_result *r) {
    if (!r) return;
    free(r->coords);
    free(r->vals);
    free(r->absum);
    free(r->nterms);
    r->coords = NULL;
    r->vals = NULL;
    r->absum = NULL;
    r->nterms = NULL;
}
