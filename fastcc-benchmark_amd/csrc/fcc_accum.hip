// Planning, exact count pass and numeric accumulation.
//
// A unit of work is (left tile, column window). Its accumulator is one of
//   DENSE  an f64 LDS tile TL x W: the reference's TileAccumulator
//          (index_tensors.hpp:994-1065) with the bitmask replaced by the sign
//          of zero (slots start at -0.0; a slot is touched iff its bits are no
//          longer those of -0.0) or, when inputs hold zeros or tiny values, by
//          an explicit LDS bitmap;
//   LHASH  an open-addressing LDS hash (u32 local key, f64 value), the
//          TileAccumulatorMap kind (index_tensors.hpp:843-883); when it fills,
//          its slots are sealed and new keys go to the global hash;
//   GHASH  the shared global-memory hash only.
// The per-product loop restates the reference's join (contract.hpp:886-896):
// for every key of the left tile, every (left, right) pair of the two fibers
// adds lv*rv at (left row, right column).
// Work distribution: a persistent grid pulls units from an atomic counter;
// inside a unit, heavy groups (fiber >= HEAVY_MIN) are taken one per
// wavefront (lanes across the fiber, the group's rows looped), light entries
// one per lane.
#include "fcc_accum.h"

namespace fcc {

constexpr uint32_t LEMPTY = 0xFFFFFFFFu, LSEALED = 0xFFFFFFFEu;
constexpr unsigned long long GEMPTY = ~0ull;
constexpr int NWAVES = ACC_THREADS / 64;
constexpr uint64_t NEG_ZERO_BITS = 0x8000000000000000ull;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint32_t lds_load(const uint32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_add_f64(double *p, double v) {
    (void)__hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ uint32_t uniform_u32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

// First index in [lo, hi) with arr[i] >= x; all 64 lanes call it together.
__device__ uint32_t wave_lower_bound(const uint32_t *arr, uint32_t lo, uint32_t hi, uint32_t x) {
    const uint32_t lane = threadIdx.x & 63;
    while (hi - lo > 64) {
        uint32_t step = (hi - lo + 63) >> 6;
        uint32_t p = lo + lane * step;
        bool lt = p < hi && arr[p] < x;
        uint32_t c = (uint32_t)__popcll(__ballot(lt));
        if (c == 0) return lo;
        uint32_t nlo = lo + (c - 1) * step + 1;
        uint32_t nhi = lo + c * step;
        lo = nlo;
        hi = nhi < hi ? nhi : hi;
    }
    uint32_t p = lo + lane;
    bool lt = p < hi && arr[p] < x;
    return lo + (uint32_t)__popcll(__ballot(lt));
}

// ---------------------------------------------------------- global hash

// A full table fails fast: after a long probe the fill level is checked and
// the insert gives up (flag set) once the table is 7/8 full; the host then
// grows the table or re-plans.
__device__ __forceinline__ bool gtable_full(const AccArgs &A, int probe) {
    if (probe < 16 || (probe & 15)) return false;
    unsigned long long c = __hip_atomic_load(A.gcount, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return c > ((A.gmask + 1) >> 3) * 7;
}

// New keys are counted once per wavefront (ballot after the probe loop, where
// the lanes that entered reconverge): one global atomic per wavefront instead
// of one per key on a single counter. Every lane of a wavefront works on the
// same unit. The unit's count takes a key twice when it stands for two
// outputs (mirror mode, off the diagonal).
__device__ __forceinline__ void count_new_keys(const AccArgs &A, bool inserted, bool twice,
                                               unsigned long long *unit_counter) {
    const uint64_t bal = __ballot(inserted);
    const uint64_t bal2 = __ballot(inserted && twice);
    const uint32_t lane = threadIdx.x & 63;
    if (bal && lane == (uint32_t)(__ffsll((unsigned long long)bal) - 1)) {
        const unsigned long long n = (unsigned long long)__popcll(bal);
        atomicAdd(A.gcount, n);
        if (unit_counter) atomicAdd(unit_counter, n + (unsigned long long)__popcll(bal2));
    }
}

__device__ bool ghash_add(const AccArgs &A, uint64_t key, double v, uint32_t unit) {
    uint64_t h = mix64(key) & A.gmask;
    bool done = false, inserted = false;
    for (int probe = 0; probe < 8192; probe++) {
        if (gtable_full(A, probe)) break;
        unsigned long long old = atomicCAS(&A.gkeys[h], GEMPTY, (unsigned long long)key);
        if (old == GEMPTY) {
            A.gunit[h] = unit;
            unsafeAtomicAdd(&A.gvals[h], v);
            done = inserted = true;
            break;
        }
        if (old == key) {
            unsafeAtomicAdd(&A.gvals[h], v);
            done = true;
            break;
        }
        h = (h + 1) & A.gmask;
    }
    count_new_keys(A, inserted, false, nullptr);
    if (!done) atomicOr(A.gflag, 1u);
    return done;
}

// count pass: set semantics, a new key adds one to its unit's count (two when
// `twice`: a mirrored cell off the diagonal)
__device__ void gset_insert(const AccArgs &A, uint64_t key, uint32_t unit, bool twice) {
    uint64_t h = mix64(key) & A.gmask;
    bool done = false, inserted = false;
    for (int probe = 0; probe < 8192; probe++) {
        if (gtable_full(A, probe)) break;
        unsigned long long old = atomicCAS(&A.gkeys[h], GEMPTY, (unsigned long long)key);
        if (old == GEMPTY) {
            done = inserted = true;
            break;
        }
        if (old == key) {
            done = true;
            break;
        }
        h = (h + 1) & A.gmask;
    }
    count_new_keys(A, inserted, twice, &A.unit_count[unit]);
    if (!done) atomicOr(A.gflag, 1u);
}

// ------------------------------------------------------------ LDS hash
// Linear probing with sealing: once `limit` keys are held, an empty slot met
// by a new key is sealed instead of claimed, and any key whose probe meets a
// sealed slot lives in the global hash. A key claimed earlier always sits
// before any slot sealed on its probe path, so it is never duplicated.
// The key count is added per wavefront after the probe (lds_probe_counted):
// one LDS atomic per wavefront instead of 64 on one address when most
// products are new keys. The count may lag by one probe round of every
// wavefront of the block (at most ACC_THREADS keys), so a table holds at most
// limit + ACC_THREADS keys, which stays below its slot count.
template <int BITS>
__device__ __forceinline__ int lds_probe(uint32_t *keys, const uint32_t *cnt, uint32_t limit, uint32_t key,
                                         bool &claimed) {
    // returns slot index, or -1 when the key must go to the global hash
    uint32_t h = (key * 2654435761u) >> (32 - BITS);
    for (uint32_t probe = 0; probe < (1u << BITS); probe++) {
        uint32_t cur = lds_load(&keys[h]);
        if (cur == key) return (int)h;
        if (cur == LSEALED) return -1;
        if (cur == LEMPTY) {
            uint32_t want = lds_load(cnt) >= limit ? LSEALED : key;
            uint32_t old = atomicCAS(&keys[h], LEMPTY, want);
            if (old == LEMPTY) {
                if (want == LSEALED) return -1;
                claimed = true;
                return (int)h;
            }
            if (old == key) return (int)h;
            if (old == LSEALED) return -1;
        }
        h = (h + 1) & ((1u << BITS) - 1);
    }
    return -1;
}

static_assert(LH_SLOTS >= LH_MAX + ACC_THREADS && SYM_SLOTS >= SYM_MAX + ACC_THREADS,
              "an LDS table must hold its limit plus one probe round of every wavefront");

// Every active lane calls this together (divergent code is fine: the ballot
// covers the lanes that are here).
// dcnt (optional, mirror-mode count pass): claimed keys on the diagonal
// (`diag`) are also counted there.
template <int BITS>
__device__ __forceinline__ int lds_probe_counted(uint32_t *keys, uint32_t *cnt, uint32_t limit, uint32_t key,
                                                 uint32_t *dcnt = nullptr, bool diag = false) {
    bool claimed = false;
    const int slot = lds_probe<BITS>(keys, cnt, limit, key, claimed);
    const uint64_t bal = __ballot(claimed);
    if (bal) {
        const uint32_t lane = threadIdx.x & 63;
        const bool first = lane == (uint32_t)(__ffsll((unsigned long long)bal) - 1);
        if (first) atomicAdd(cnt, (uint32_t)__popcll(bal));
        if (dcnt) {
            const uint64_t bd = __ballot(claimed && diag);
            if (bd && first) atomicAdd(dcnt, (uint32_t)__popcll(bd));
        }
    }
    return slot;
}

// ------------------------------------------------------------- per unit

struct SinkAcc {
    uint64_t count = 0, hash = 0;
    double vsum = 0.0, wsum = 0.0;
    uint64_t group = 0;  // row group this accumulator belongs to
    uint64_t glo = 0;    // its first row (group * group_rows)
};

__device__ __forceinline__ double hash_weight(uint64_t z) { return (double)(z >> 11) * 0x1.0p-53; }

__device__ __forceinline__ void sink_atomic(const AccArgs &A, uint64_t g, unsigned long long c, unsigned long long h,
                                            double vs, double ws) {
    atomicAdd(&A.sink[4 * g], c);
    atomicAdd(&A.sink[4 * g + 1], h);
    unsafeAtomicAdd(&A.sinkf[4 * g], vs);
    unsafeAtomicAdd(&A.sinkf[4 * g + 1], ws);
}

__device__ __forceinline__ void emit_output(const AccArgs &A, uint64_t pos, uint64_t row, uint64_t col, double v,
                                            bool sink, SinkAcc &sk) {
    if (sink) {
        if (A.group_rows) {
            // rows and group sizes are below 2^32 (left spans are)
            if (row - sk.glo < A.group_rows) {
                row -= sk.glo;
            } else {  // a unit straddling two groups: add this output directly
                const uint32_t g = (uint32_t)row / (uint32_t)A.group_rows;
                row -= (uint64_t)g * A.group_rows;
                uint64_t z = mix64(row * A.spanB + col);
                sink_atomic(A, g, 1ull, z, v, v * hash_weight(z));
                return;
            }
        }
        uint64_t z = mix64(row * A.spanB + col);
        sk.count++;
        sk.hash += z;
        sk.vsum += v;
        sk.wsum += v * hash_weight(z);
        return;
    }
    if (pos >= A.capacity) {  // a count/fill disagreement: flag it, never write past the output
        atomicOr(A.fail_flag, 2u);
        return;
    }
    // Outputs are written once and not read again by this pass: non-temporal
    // stores keep them from evicting the right fibers the units re-read from
    // the L2.
    uint32_t r = (uint32_t)row, c = (uint32_t)col;
    for (int m = A.nxa - 1; m >= 0; m--) {
        uint32_t E = A.xaE[m];
        uint32_t q = r / E;
        __builtin_nontemporal_store(r - q * E, &A.ocoords[(uint64_t)m * A.ostride + pos]);
        r = q;
    }
    for (int m = A.nxb - 1; m >= 0; m--) {
        uint32_t E = A.xbE[m];
        uint32_t q = c / E;
        __builtin_nontemporal_store(c - q * E, &A.ocoords[(uint64_t)(A.nxa + m) * A.ostride + pos]);
        c = q;
    }
    __builtin_nontemporal_store(v, &A.ovals[pos]);
}


struct UnitCtx {
    uint32_t u, a, tile, c0, W, kind, multi;
    uint32_t Wp;  // dense row stride: W rounded up to a multiple of 32 (hash kinds: W)
    uint64_t row0;
    uint64_t obase;  // DIRECT units: first output position of the unit
};

template <bool NUM, bool BITMAP>
struct Acc {
    // NUM: DENSE f64 slots / LHASH vals+keys; SYM: DENSE byte map / HASH key set
    double *fv;       // numeric values
    uint32_t *keys;   // LHASH keys (numeric) or key set (count)
    uint32_t *bits;   // DENSE bitmap (numeric BITMAP mode)
    uint8_t *seen;    // DENSE count pass: one byte per slot, set to 1 by plain
                      // stores (idempotent: no atomics, no per-word merging)
    uint32_t *cnt;
    uint32_t *dcnt;   // mirror-mode count pass: LDS-hash keys on the diagonal (else null)
    SinkAcc *sk;      // DIRECT units: count + checksum outputs go here
    bool sink;
};

// A distinct output written straight out: count + checksum, or a position
// reserved per wavefront from the unit's LDS counter (order within the unit
// is free).
// In mirror mode the caller passes cells with col >= row only; off the
// diagonal the transposed cell is written too (after the wavefront's direct
// cells, so each store instruction writes consecutive positions).
template <bool NUM, bool BITMAP>
__device__ __forceinline__ void direct_emit(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S,
                                            uint64_t row, uint64_t col, double v) {
    const bool twice = A.mirror && col > row;
    if (S.sink) {
        emit_output(A, 0, row, col, v, true, *S.sk);
        if (twice) emit_output(A, 0, col, row, v, true, *S.sk);
        return;
    }
    const uint32_t lane = threadIdx.x & 63;
    const uint64_t act = __ballot(1);
    const uint64_t act2 = __ballot(twice);
    const int first = __ffsll((unsigned long long)act) - 1;
    uint32_t b = 0;
    if (lane == (uint32_t)first) b = atomicAdd(S.cnt, (uint32_t)(__popcll(act) + __popcll(act2)));
    b = __shfl(b, first, 64);
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    emit_output(A, U.obase + b + (uint32_t)__popcll(act & lt), row, col, v, false, *S.sk);
    if (twice)
        emit_output(A, U.obase + b + (uint32_t)__popcll(act) + (uint32_t)__popcll(act2 & lt), col, row, v, false,
                    *S.sk);
}

template <bool NUM, bool BITMAP>
__device__ __forceinline__ void accum(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S, uint32_t li,
                                      uint32_t col, double v) {
    if (!NUM && U.kind == KIND_CBITS) {
        const uint32_t slot = li * U.Wp + col;
        atomicOr(&S.keys[slot >> 5], 1u << (slot & 31));
        return;
    }
    if (U.kind == KIND_DENSE) {
        uint32_t slot = li * U.Wp + col;
        if (!NUM) S.seen[slot] = 1;
        if (NUM && BITMAP) atomicOr(&S.bits[slot >> 5], 1u << (slot & 31));
        if (NUM) lds_add_f64(&S.fv[slot], v);
        return;
    }
    // hash kinds: in mirror mode cells below the diagonal are not kept (a
    // dense slot exists anyway and is dropped when the tile is written out)
    const uint32_t w = A.mirror ? mirror_weight(U.row0 + li, (uint64_t)U.c0 + col) : 1u;
    if (w == 0) return;
    if (U.kind == KIND_LHASH) {
        uint32_t key = li * U.W + col;
        int slot = NUM ? lds_probe_counted<12>(S.keys, S.cnt, LH_MAX, key)
                       : lds_probe_counted<14>(S.keys, S.cnt, SYM_MAX, key, S.dcnt, w == 1);
        if (slot >= 0) {
            if (NUM) lds_add_f64(&S.fv[slot], v);
        } else {
            uint64_t gk = (U.row0 + li) * A.spanB + (U.c0 + col);
            if (NUM) ghash_add(A, gk, v, U.u);
            else gset_insert(A, gk, U.u, w == 2);
        }
    } else if (U.kind == KIND_DIRECT) {
        if (NUM) direct_emit<NUM, BITMAP>(A, U, S, U.row0 + li, (uint64_t)U.c0 + col, v);
    } else {
        uint64_t gk = (U.row0 + li) * A.spanB + (U.c0 + col);
        if (NUM) ghash_add(A, gk, v, U.u);
        else gset_insert(A, gk, U.u, w == 2);
    }
}

__device__ __forceinline__ uint32_t rlane(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
__device__ __forceinline__ double rlane_f64(double v, uint32_t l) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    uint32_t lo = rlane((uint32_t)b, l), hi = rlane((uint32_t)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// per-lane lower bound: first index in [lo, hi) with arr[i] >= x
__device__ __forceinline__ uint32_t lane_lower_bound(const uint32_t *arr, uint32_t lo, uint32_t hi, uint32_t x) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (arr[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Heavy groups of a dense unit with row ownership: every wavefront walks
// every heavy group but only applies rows li with li % NWAVES == its index, so
// no two wavefronts ever touch the same slot and the accumulator (and bitmap)
// is updated with plain LDS read-modify-write instead of atomics. Within a
// wavefront the 64 lanes hold distinct columns of one row at a time.
template <bool NUM, bool BITMAP>
__device__ __forceinline__ void dense_heavy_owned(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t g0 = A.atile_g0[U.a], gl = A.atile_lg[U.a];
    const uint32_t cend = U.c0 + U.W;
    const uint32_t wpr = U.Wp >> 5;
    constexpr bool BITS = NUM && BITMAP;
    for (uint32_t bg = g0; bg < gl; bg += 64) {
        const uint32_t nb = min(64u, gl - bg);
        uint4 d = make_uint4(0, 0, 0, 0);
        if (lane < nb) d = A.gdesc[bg + lane];
        uint32_t s_l = d.x, e_l = d.x + d.y;
        if (U.multi && lane < nb) {
            s_l = lane_lower_bound(A.Bcol, s_l, e_l, U.c0);
            e_l = lane_lower_bound(A.Bcol, s_l, e_l, cend);
        }
        uint32_t s = rlane(s_l, 0), e = rlane(e_l, 0), ep0 = rlane(d.z, 0), m = rlane(d.w, 0);
        uint32_t nli = 0, ncol = 0;
        double nav = 0.0, nbv = 0.0;
        if (lane < m) {
            nli = A.Ali[ep0 + lane];
            if (NUM) nav = A.Aval[ep0 + lane];
        }
        if (s + lane < e) {
            ncol = A.Bcol[s + lane];
            if (NUM) nbv = A.Bval[s + lane];
        }
        for (uint32_t j = 0; j < nb; j++) {
            const uint32_t cs = s, ce = e, cep0 = ep0, cm = m;
            const uint32_t cli = nli, ccol = ncol;
            const double cav = nav, cbv = nbv;
            if (j + 1 < nb) {
                s = rlane(s_l, j + 1);
                e = rlane(e_l, j + 1);
                ep0 = rlane(d.z, j + 1);
                m = rlane(d.w, j + 1);
                nli = 0;
                ncol = 0;
                nav = 0.0;
                nbv = 0.0;
                if (lane < m) {
                    nli = A.Ali[ep0 + lane];
                    if (NUM) nav = A.Aval[ep0 + lane];
                }
                if (s + lane < e) {
                    ncol = A.Bcol[s + lane];
                    if (NUM) nbv = A.Bval[s + lane];
                }
            }
            // rows of this group owned by this wavefront (first 64 rows)
            uint64_t own0 = __ballot(lane < cm && (cli & (NWAVES - 1)) == wave);
            if (own0 == 0 && cm <= 64) continue;
            for (uint32_t cb = cs; cb < ce; cb += 64) {
                uint32_t col = ccol;
                double b = cbv;
                const bool ok = cb + lane < ce;
                if (cb != cs) {
                    col = 0;
                    b = 0.0;
                    if (ok) {
                        col = A.Bcol[cb + lane];
                        if (NUM) b = A.Bval[cb + lane];
                    }
                }
                col -= U.c0;
                uint32_t wc = ok ? (col >> 5) : 0xFFFFFFFFu, bm = ok ? (1u << (col & 31)) : 0u;
                bool leader = false;
                if (BITS) {
#pragma unroll
                    for (int dd = 1; dd < 64; dd <<= 1) {
                        uint32_t ob = __shfl_down(bm, dd, 64), ow = __shfl_down(wc, dd, 64);
                        if (lane + dd < 64 && ow == wc) bm |= ob;
                    }
                    uint32_t pw = __shfl_up(wc, 1, 64);
                    leader = ok && (lane == 0 || pw != wc);
                }
                for (uint32_t kb = 0; kb < cm; kb += 64) {
                    uint32_t li_l = cli;
                    double av_l = cav;
                    uint64_t own = own0;
                    if (kb != 0) {
                        li_l = 0;
                        av_l = 0.0;
                        if (kb + lane < cm) {
                            li_l = A.Ali[cep0 + kb + lane];
                            if (NUM) av_l = A.Aval[cep0 + kb + lane];
                        }
                        own = __ballot(kb + lane < cm && (li_l & (NWAVES - 1)) == wave);
                    }
                    while (own) {
                        const uint32_t kk = (uint32_t)__builtin_ctzll(own);
                        own &= own - 1;
                        const uint32_t li = rlane(li_l, kk);
                        if (BITS && leader) S.bits[li * wpr + wc] |= bm;
                        if (!NUM && ok) S.seen[li * U.Wp + col] = 1;
                        if (NUM && ok) {
                            double *slot = &S.fv[li * U.Wp + col];
                            *slot = *slot + rlane_f64(av_l, kk) * b;
                        }
                    }
                }
            }
        }
    }
}

// Heavy groups claimed 64 at a time per wavefront; lane l holds group l's
// descriptor (fiber range, entry range), read back as scalars with readlane.
// For group j the wavefront's lanes cover 64 columns of the fiber segment and
// loop over the group's rows. Groups are pipelined in blocks of PIPE: the
// first row and column chunks of the next PIPE groups are loaded while the
// current PIPE groups accumulate, so 2*PIPE groups' loads are in flight per
// wavefront. Updates are atomic.
constexpr int PIPE = 4;

template <bool NUM>
__device__ __forceinline__ void load_first_chunks(const AccArgs &A, uint32_t lane, uint32_t s, uint32_t e, uint32_t ep0,
                                                  uint32_t m, uint32_t &li, double &av, uint32_t &col, double &b) {
    li = 0;
    col = 0;
    av = 0.0;
    b = 0.0;
    if (lane < m) {
        li = A.Ali[ep0 + lane];
        if (NUM) av = A.Aval[ep0 + lane];
    }
    if (s + lane < e) {
        col = A.Bcol[s + lane];
        if (NUM) b = A.Bval[s + lane];
    }
}

// One 64-column chunk of a group's fiber (lane = column) against the group's
// rows li_l/av_l (lane k holds row k, k < mk).
template <bool NUM, bool BITMAP, bool DONLY = false>
__device__ __forceinline__ void heavy_chunk(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S,
                                            uint32_t lane, uint32_t col, double b, bool ok, uint32_t li_l,
                                            double av_l, uint32_t mk) {
    col -= U.c0;
    if (DONLY || U.kind == KIND_DENSE) {
        uint32_t wc = 0, bm = 0;
        bool leader = false;
        if (NUM && BITMAP) {
            // The chunk's columns are ascending, so lanes sharing a 32-column
            // bitmap word are contiguous; OR their bits together and let one
            // lane per word touch the bitmap.
            wc = ok ? (col >> 5) : 0xFFFFFFFFu;
            bm = ok ? (1u << (col & 31)) : 0u;
#pragma unroll
            for (int dd = 1; dd < 64; dd <<= 1) {
                uint32_t ob = __shfl_down(bm, dd, 64), ow = __shfl_down(wc, dd, 64);
                if (lane + dd < 64 && ow == wc) bm |= ob;
            }
            uint32_t pw = __shfl_up(wc, 1, 64);
            leader = ok && (lane == 0 || pw != wc);
        }
        const uint32_t wpr = U.Wp >> 5;
        for (uint32_t kk = 0; kk < mk; kk++) {
            const uint32_t li = rlane(li_l, kk);
            if (NUM && BITMAP && leader) atomicOr(&S.bits[li * wpr + wc], bm);
            if (!NUM && ok) S.seen[li * U.Wp + col] = 1;
            if (NUM && ok) lds_add_f64(&S.fv[li * U.Wp + col], rlane_f64(av_l, kk) * b);
        }
    } else {
        for (uint32_t kk = 0; kk < mk; kk++) {
            const uint32_t li = rlane(li_l, kk);
            const double av = NUM ? rlane_f64(av_l, kk) : 0.0;
            if (ok) accum<NUM, BITMAP>(A, U, S, li, col, av * b);
        }
    }
}

template <bool NUM, bool BITMAP>
__device__ __forceinline__ void heavy_group(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S,
                                            uint32_t lane, uint32_t cs, uint32_t ce, uint32_t cep0, uint32_t cm,
                                            uint32_t cli, double cav, uint32_t ccol, double cbv) {
    if (cm <= 64) {
        // All rows already in registers: nothing is loaded between the next
        // chunk's prefetch and this chunk's use, so the prefetch stays in flight.
        uint32_t col = ccol;
        double b = cbv;
        for (uint32_t cb = cs; cb < ce; cb += 64) {
            uint32_t ncol = 0;
            double nbv = 0.0;
            if (cb + 64 + lane < ce) {
                ncol = A.Bcol[cb + 64 + lane];
                if (NUM) nbv = A.Bval[cb + 64 + lane];
            }
            heavy_chunk<NUM, BITMAP>(A, U, S, lane, col, b, cb + lane < ce, cli, cav, cm);
            col = ncol;
            b = nbv;
        }
        return;
    }
    // groups with more than 64 rows: rows reloaded 64 at a time per chunk
    for (uint32_t cb = cs; cb < ce; cb += 64) {
        uint32_t col = ccol;
        double b = cbv;
        if (cb != cs) {
            col = 0;
            b = 0.0;
            if (cb + lane < ce) {
                col = A.Bcol[cb + lane];
                if (NUM) b = A.Bval[cb + lane];
            }
        }
        for (uint32_t kb = 0; kb < cm; kb += 64) {
            uint32_t li_l = cli;
            double av_l = cav;
            if (kb != 0) {
                li_l = 0;
                av_l = 0.0;
                if (kb + lane < cm) {
                    li_l = A.Ali[cep0 + kb + lane];
                    if (NUM) av_l = A.Aval[cep0 + kb + lane];
                }
            }
            heavy_chunk<NUM, BITMAP>(A, U, S, lane, col, b, cb + lane < ce, li_l, av_l, min(64u, cm - kb));
        }
    }
}

// Heavy groups (or segments) are claimed in batches per wavefront; lane l
// holds descriptor l (fiber range clipped to the unit's window, first entry,
// entries), read back as scalars with readlane.
//
// The batch is walked as one stream of (group, 64-column chunk) items with a
// ring of HRING items in flight: the fiber chunk (lane = column) and, at a
// group's first chunk, its rows (lane = row) are loaded HRING items before they
// accumulate. Loads are unconditional on clamped addresses so the compiler's
// memory counters stay exact across the ring. Batches holding a group with
// more than 64 rows take the per-group path (heavy_group).
constexpr int HRING_DENSE = 8;  // dense units: a small chunk body, a deep ring
constexpr int HRING_HASH = 4;   // hash units: the chunk body holds the hash probes


// The (group, chunk) stream of one claimed batch, RING items in flight.
template <bool NUM, bool BITMAP, int RING, bool DONLY>
__device__ __forceinline__ void heavy_ring(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S,
                                           uint32_t lane, uint32_t nb, uint32_t s_l, uint32_t e_l, uint4 d) {
    const uint32_t nc_l = (lane < nb && e_l > s_l) ? (e_l - s_l + 63) / 64 : 0u;
    uint32_t ij = 0, ic = 0;  // issue cursor: group, chunk (wave-uniform)
    while (ij < nb && rlane(nc_l, ij) == 0) ij++;
    uint32_t rcol[RING], rli[RING], rj[RING], rc[RING];
    double rb[RING], rav[RING];
    bool rv[RING];
    auto issue = [&](int k) {
        rv[k] = ij < nb;
        rcol[k] = 0;
        rb[k] = 0.0;
        rli[k] = 0;
        rav[k] = 0.0;
        if (!rv[k]) return;
        const uint32_t s0 = rlane(s_l, ij) + ic * 64, e = rlane(e_l, ij);
        const uint32_t p = min(s0 + lane, e - 1);
        rcol[k] = A.Bcol[p];
        if (NUM) rb[k] = A.Bval[p];
        if (ic == 0) {
            const uint32_t ep0 = rlane(d.z, ij), m = rlane(d.w, ij);
            const uint32_t q = ep0 + min(lane, m - 1);
            rli[k] = A.Ali[q];
            if (NUM) rav[k] = A.Aval[q];
        }
        rj[k] = ij;
        rc[k] = ic;
        if (++ic == rlane(nc_l, ij)) {
            ic = 0;
            ij++;
            while (ij < nb && rlane(nc_l, ij) == 0) ij++;
        }
    };
#pragma unroll
    for (int k = 0; k < RING; k++) issue(k);
    uint32_t gli = 0, gm = 0;
    double gav = 0.0;
    for (bool more = rv[0]; more;) {
#pragma unroll
        for (int k = 0; k < RING; k++) {
            if (rv[k]) {
                const uint32_t j = rj[k], c = rc[k];
                if (c == 0) {
                    gli = rli[k];
                    gav = rav[k];
                    gm = rlane(d.w, j);
                }
                const uint32_t s0 = rlane(s_l, j) + c * 64, e = rlane(e_l, j);
                const uint32_t col = rcol[k];
                const double b = rb[k];
                issue(k);  // refill: the item RING ahead
                heavy_chunk<NUM, BITMAP, DONLY>(A, U, S, lane, col, b, s0 + lane < e, gli, gav, gm);
            }
        }
        more = rv[0];
    }
}

// Mirror mode, a row's last column window when the window starts at or left
// of the tile's first row: every range (an entry's, a group's or a segment's)
// starts at or right of its row's own column, inside the window, and ends
// with its fiber at the last cut, so the ranges need no window cut.
__device__ __forceinline__ bool mirror_whole_fiber(const AccArgs &A, const UnitCtx &U) {
    return A.mirror && (uint64_t)U.c0 <= U.row0 && U.c0 / A.fcut_W + 2 == A.fcut_n;
}

// DEEP: dense units keep HRING_DENSE items in flight; without it (kernels
// whose other paths already need many registers) they keep HRING_HASH, so
// the kernel stays within 128 VGPRs and two workgroups fit per CU.
template <bool NUM, bool BITMAP, bool DEEP = true>
__device__ __forceinline__ void heavy_claimed(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S, uint32_t *s_gnext) {
    const uint32_t lane = threadIdx.x & 63;
    // A tile with few heavy groups works on fiber segments (long fibers spread
    // over the block's wavefronts, claims sized so every wavefront gets some);
    // a tile with many works on whole groups, 64 per claim (fewer descriptors
    // to clip against column windows).
    const uint32_t ngh = A.atile_lg[U.a] - A.atile_g0[U.a];
    const bool seg = ngh < A.seg_below;
    const uint4 *desc = seg ? A.hdesc : A.gdesc;
    const uint32_t g0 = seg ? A.atile_h0[U.a] : A.atile_g0[U.a];
    const uint32_t gl = seg ? A.atile_h0[U.a + 1] : A.atile_lg[U.a];
    const uint32_t cend = U.c0 + U.W;
    const uint32_t claim =
        (seg && A.claim_div) ? max(1u, min(64u, (gl - g0 + A.claim_div - 1) / A.claim_div)) : 64u;
    for (;;) {
        uint32_t bs = 0;
        if (lane == 0) bs = atomicAdd(s_gnext, claim);
        bs = uniform_u32(__shfl(bs, 0, 64));
        if (bs >= gl - g0) break;
        const uint32_t bg = g0 + bs;
        const uint32_t nb = min(claim, gl - bg);
        uint4 d = make_uint4(0, 0, 0, 0);
        if (lane < nb) d = desc[bg + lane];
        uint32_t s_l = d.x, e_l = d.x + d.y;
        if (A.unclipped && lane < nb && d.y) s_l = A.fiber_ptr[A.group_fiber[bg + lane]];
        const bool cuts = U.multi && A.fcut && U.kind == KIND_DENSE;
        if (cuts && mirror_whole_fiber(A, U)) {
            // the descriptor's range as it is
        } else if (cuts && lane < nb && d.y) {
            // precomputed window cuts of the group's fiber (in mirror mode the
            // descriptor may start later: no cell left of the tile's first
            // row); a heavy segment, a piece of its group's fiber, keeps the
            // part of the piece inside the cuts
            const uint32_t w = U.c0 / A.fcut_W;
            const uint32_t f = seg ? A.Afib[d.z] : A.group_fiber[bg + lane];
            const uint32_t *fc = A.fcut + (size_t)f * A.fcut_n + w;
            if (seg) {
                e_l = min(e_l, fc[1]);
                s_l = min(max(s_l, fc[0]), e_l);
            } else {
                e_l = fc[1];
                s_l = A.mirror ? min(max(fc[0], d.x), e_l) : fc[0];
            }
        } else if (U.multi && lane < nb && s_l < e_l) {
            // segments wholly outside the window are dropped on their end columns
            if (A.Bcol[s_l] >= cend || A.Bcol[e_l - 1] < U.c0) {
                e_l = s_l;
            } else {
                s_l = lane_lower_bound(A.Bcol, s_l, e_l, U.c0);
                e_l = lane_lower_bound(A.Bcol, s_l, e_l, cend);
            }
        }

        if (__ballot(lane < nb && d.w > 64) != 0) {
            // rows beyond one wavefront: group by group
            for (uint32_t j = 0; j < nb; j++) {
                uint32_t li, col;
                double av, bv;
                load_first_chunks<NUM>(A, lane, rlane(s_l, j), rlane(e_l, j), rlane(d.z, j), rlane(d.w, j), li, av,
                                       col, bv);
                heavy_group<NUM, BITMAP>(A, U, S, lane, rlane(s_l, j), rlane(e_l, j), rlane(d.z, j), rlane(d.w, j),
                                         li, av, col, bv);
            }
            continue;
        }
        if (U.kind == KIND_DENSE)
            heavy_ring<NUM, BITMAP, DEEP ? HRING_DENSE : HRING_HASH, true>(A, U, S, lane, nb, s_l, e_l, d);
        else
            heavy_ring<NUM, BITMAP, HRING_HASH, false>(A, U, S, lane, nb, s_l, e_l, d);
    }
}

// Smallest i >= from with inc[i] > q, where inc is nondecreasing and
// inc[n - 1] > q (inclusive prefix sums of products per entry, in LDS). A
// lane's product index only grows, so it searches forward from its last
// entry: one probe while it stays inside a long fiber, a galloping search
// (log of the distance) when it skips short ones -- instead of a full binary
// search over the chunk per product.
__device__ __forceinline__ uint32_t gallop_upper(const uint32_t *inc, uint32_t from, uint32_t n, uint32_t q) {
    if (inc[from] > q) return from;
    uint32_t lo = from, step = 1, hi = from + 1;
    while (hi < n - 1 && inc[hi] <= q) {
        lo = hi;
        step <<= 1;
        hi = min(lo + step, n - 1);
    }
    uint32_t a = lo + 1, b = hi;  // inc[lo] <= q < inc[hi]
    while (a < b) {
        const uint32_t mid = (a + b) >> 1;
        if (inc[mid] > q) b = mid;
        else a = mid + 1;
    }
    return a;
}

// Light entries [le0, le1) of a unit: the products of every entry (its right
// fiber, clipped to the unit's column window) are flattened and spread over
// all lanes, so a tile of a few entries with long fibers keeps the block busy.
// With `stage` (2 * ACC_THREADS words of LDS) the flattening is block-wide:
// an inclusive scan of fiber lengths over 512 entries at a time, and each
// product finds its entry by binary search in LDS. Without it each wavefront
// flattens its own 64 entries in registers (owner found by shuffles).
template <bool NUM, bool BITMAP>
__device__ __forceinline__ void light_entries(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S, uint32_t le0,
                              uint32_t le1, uint32_t *stage, uint32_t *s_w) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t cend = U.c0 + U.W;
    const bool cuts = U.multi && A.fcut && U.kind == KIND_DENSE;
    const bool whole = cuts && mirror_whole_fiber(A, U);
    auto span_of = [&](uint32_t e, uint32_t &s0, uint32_t &len) {
        uint32_t a, b;
        if (whole) {
            const uint2 fp = A.Afp[e];
            a = fp.x;
            b = fp.y;
        } else if (cuts) {
            const uint32_t *fc = A.fcut + (size_t)A.Afib[e] * A.fcut_n + U.c0 / A.fcut_W;
            a = fc[0];
            b = fc[1];
            if (A.dk_tile || A.mirror) {
                // entries of dense keys carry an empty fiber range; in mirror
                // mode the range starts at the entry's own column
                const uint2 fp = A.Afp[e];
                if (fp.x == fp.y) b = a;
                if (A.mirror) a = min(max(a, fp.x), b);
            }
        } else {
            const uint2 fp = A.Afp[e];
            a = (A.unclipped && fp.x != fp.y) ? A.fiber_ptr[A.Afib[e]] : fp.x;
            b = fp.y;
            if (U.multi) {
                a = lane_lower_bound(A.Bcol, a, b, U.c0);
                b = lane_lower_bound(A.Bcol, a, b, cend);
            }
        }
        s0 = a;
        len = b - a;
    };
    if (stage) {
        uint32_t *s_inc = stage, *s_s0 = stage + ACC_THREADS;
        for (uint32_t eb = le0; eb < le1; eb += ACC_THREADS) {
            const uint32_t e = eb + threadIdx.x;
            uint32_t s0 = 0, len = 0;
            if (e < le1) span_of(e, s0, len);
            uint32_t inc = len;
            for (int d = 1; d < 64; d <<= 1) {
                uint32_t t = __shfl_up(inc, d, 64);
                if (lane >= (uint32_t)d) inc += t;
            }
            if (lane == 63) s_w[wave] = inc;
            __syncthreads();
            uint32_t pre = 0, P = 0;
#pragma unroll
            for (int w = 0; w < NWAVES; w++) {
                const uint32_t c = s_w[w];
                pre += (w < (int)wave) ? c : 0;
                P += c;
            }
            s_inc[threadIdx.x] = inc + pre;
            s_s0[threadIdx.x] = s0;
            __syncthreads();
            const uint32_t nent = min(le1 - eb, (uint32_t)ACC_THREADS);
            // product q's entry and fiber position; the loads for the next
            // product of this lane are issued before the current one accumulates
            uint32_t cur = 0;  // this lane's last entry (its products come in increasing order)
            auto locate = [&](uint32_t q, uint32_t &p, uint32_t &ee) {
                const uint32_t lo = gallop_upper(s_inc, cur, nent, q);  // s_inc[nent - 1] == P > q
                cur = lo;
                p = s_s0[lo] + (q - (lo ? s_inc[lo - 1] : 0u));
                ee = eb + lo;
            };
            uint32_t q = threadIdx.x, p = 0, ee = 0, col = 0, li = 0;
            double av = 0.0, b = 0.0;
            if (q < P) {
                locate(q, p, ee);
                col = A.Bcol[p];
                li = A.Ali[ee];
                if (NUM) {
                    b = A.Bval[p];
                    av = A.Aval[ee];
                }
            }
            for (; q < P; q += ACC_THREADS) {
                uint32_t ncol = 0, nli = 0;
                double nav = 0.0, nb = 0.0;
                if (q + ACC_THREADS < P) {
                    uint32_t np, nee;
                    locate(q + ACC_THREADS, np, nee);
                    ncol = A.Bcol[np];
                    nli = A.Ali[nee];
                    if (NUM) {
                        nb = A.Bval[np];
                        nav = A.Aval[nee];
                    }
                }
                accum<NUM, BITMAP>(A, U, S, li, col - U.c0, NUM ? av * b : 0.0);
                col = ncol;
                li = nli;
                av = nav;
                b = nb;
            }
            __syncthreads();  // s_inc / s_s0 / s_w are rewritten by the next chunk
        }
        return;
    }
    for (uint32_t eb = le0 + wave * 64; eb < le1; eb += ACC_THREADS) {
        const uint32_t e = eb + lane;
        uint32_t s0 = 0, len = 0, li = 0;
        double av = 0.0;
        if (e < le1) {
            span_of(e, s0, len);
            li = A.Ali[e];
            if (NUM) av = A.Aval[e];
        }
        uint32_t inc = len;
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t t = __shfl_up(inc, d, 64);
            if (lane >= (uint32_t)d) inc += t;
        }
        const uint32_t P = __shfl(inc, 63, 64);
        for (uint32_t q0 = 0; q0 < P; q0 += 64) {
            const uint32_t q = q0 + lane;
            uint32_t j = 0;  // owner: number of lanes whose inclusive end is <= q
#pragma unroll
            for (int b = 32; b >= 1; b >>= 1) {
                const uint32_t v = __shfl(inc, (int)(j + b - 1), 64);
                if (v <= q) j += b;
            }
            const uint32_t oinc = __shfl(inc, (int)(j & 63), 64), olen = __shfl(len, (int)(j & 63), 64);
            const uint32_t os0 = __shfl(s0, (int)(j & 63), 64), oli = __shfl(li, (int)(j & 63), 64);
            const double oav = NUM ? __shfl(av, (int)(j & 63), 64) : 0.0;
            if (q < P) {
                const uint32_t p = os0 + (q - (oinc - olen));
                const uint32_t col = A.Bcol[p];
                accum<NUM, BITMAP>(A, U, S, oli, col - U.c0, NUM ? oav * A.Bval[p] : 0.0);
            }
        }
    }
}

static_assert(LH_SLOTS + LH_SLOTS / 2 + ACC_THREADS <= DENSE_SLOTS, "LHASH flattening stage must fit in the LDS tile");

// A run of DIRECT tiles: left entries [e0, e1) (whole tiles, one window),
// every product a distinct output. Products are flattened over the block like
// light entries (scan of fiber lengths staged in LDS, owner by binary search).
template <bool BITMAP>
__device__ __forceinline__ void direct_run(const AccArgs &A, const UnitCtx &U, const Acc<true, BITMAP> &S, uint32_t e0, uint32_t e1,
                           uint32_t *stage, uint32_t *s_w) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t *s_inc = stage, *s_s0 = stage + ACC_THREADS;
    for (uint32_t eb = e0; eb < e1; eb += ACC_THREADS) {
        const uint32_t e = eb + threadIdx.x;
        uint32_t s0 = 0, len = 0;
        if (e < e1) {
            const uint2 fp = A.Afp[e];
            s0 = fp.x;  // mirror mode: the range starts at the entry's own column
            len = fp.y - s0;
        }
        uint32_t inc = len;
        for (int d = 1; d < 64; d <<= 1) {
            uint32_t t = __shfl_up(inc, d, 64);
            if (lane >= (uint32_t)d) inc += t;
        }
        if (lane == 63) s_w[wave] = inc;
        __syncthreads();
        uint32_t pre = 0, P = 0;
#pragma unroll
        for (int w = 0; w < NWAVES; w++) {
            const uint32_t c = s_w[w];
            pre += (w < (int)wave) ? c : 0;
            P += c;
        }
        s_inc[threadIdx.x] = inc + pre;
        s_s0[threadIdx.x] = s0;
        __syncthreads();
        const uint32_t nent = min(e1 - eb, (uint32_t)ACC_THREADS);
        uint32_t cur = 0;  // this lane's last entry (its products come in increasing order)
        auto locate = [&](uint32_t q, uint32_t &p, uint32_t &ee) {
            const uint32_t lo = gallop_upper(s_inc, cur, nent, q);  // s_inc[nent - 1] == P > q
            cur = lo;
            p = s_s0[lo] + (q - (lo ? s_inc[lo - 1] : 0u));
            ee = eb + lo;
        };
        // the next product's loads are in flight while the current one is written
        uint32_t q = threadIdx.x, row = 0, col = 0;
        double av = 0.0, b = 0.0;
        if (q < P) {
            uint32_t p, ee;
            locate(q, p, ee);
            row = A.Arow[ee];
            col = A.Bcol[p];
            av = A.Aval[ee];
            b = A.Bval[p];
        }
        for (; q < P; q += ACC_THREADS) {
            uint32_t nrow = 0, ncol = 0;
            double nav = 0.0, nb = 0.0;
            if (q + ACC_THREADS < P) {
                uint32_t np, nee;
                locate(q + ACC_THREADS, np, nee);
                nrow = A.Arow[nee];
                ncol = A.Bcol[np];
                nav = A.Aval[nee];
                nb = A.Bval[np];
            }
            direct_emit<true, BITMAP>(A, U, S, row, col, av * b);
            row = nrow;
            col = ncol;
            av = nav;
            b = nb;
        }
        __syncthreads();  // s_inc / s_s0 / s_w are rewritten by the next chunk
    }
}

// ------------------------------------------------------- dense-key block
// Presence bits of the dense-key block on `row`, columns [col, col + 32); col
// need not be word-aligned (the map carries a spare word at its end).
__device__ __forceinline__ uint32_t dk_word(const AccArgs &A, uint64_t row, uint64_t col) {
    const uint32_t *pr = A.dk_P + row * A.dk_pw + (col >> 5);
    const uint32_t sh = (uint32_t)(col & 31);
    uint32_t w = pr[0];
    if (sh) w = (uint32_t)(((((uint64_t)pr[1]) << 32) | w) >> sh);
    return w;
}

// Block value at (row, col): the sum of the GEMM's k-split partial blocks.
__device__ __forceinline__ double dk_value(const AccArgs &A, uint64_t row, uint64_t col) {
    const double *p = A.dk_C + row * A.dk_rb + col;
    double v = p[0];
    for (uint32_t z = 1; z < A.dk_splits; z++) v += p[(uint64_t)z * A.dk_slab];
    return v;
}

// Bitmap word w of a dense unit (rows of U.Wp slots) from the block.
__device__ __forceinline__ uint32_t dk_dense_word(const AccArgs &A, const UnitCtx &U, uint32_t w) {
    const uint32_t wpr = U.Wp >> 5;
    const uint32_t li = w / wpr, cl0 = (w - li * wpr) * 32;
    const uint64_t row = U.row0 + li;
    if (row >= A.spanA || cl0 >= U.W) return 0u;
    uint32_t b = dk_word(A, row, (uint64_t)U.c0 + cl0);
    if (U.W - cl0 < 32) b &= (1u << (U.W - cl0)) - 1u;
    return b;
}

// Dense unit accumulator start: block values where present, else the empty
// marker (+0 with a bitmap, -0 without; a present -0 is stored as +0).
template <bool BITMAP>
__device__ __forceinline__ void dk_seed_dense(const AccArgs &A, const UnitCtx &U, double *fv, uint32_t *bits, uint32_t nslots) {
    const double init = BITMAP ? 0.0 : -0.0;
    const uint32_t TL = 1u << A.logTL;
    if (U.Wp >= 2 * ACC_THREADS && !(U.c0 & 1u)) {
        // wide rows: row by row, two cells per thread (one 16-byte load of
        // each slab, no division); in mirror mode the cells left of a row's
        // diagonal are never emitted and are not read
        const uint32_t np = U.Wp / 2;
        for (uint32_t li = 0; li < TL; li++) {
            const uint64_t row = U.row0 + li;
            double2 *out = reinterpret_cast<double2 *>(fv + (size_t)li * U.Wp);
            const uint32_t *prow = A.dk_P + row * A.dk_pw;
            const double *crow = A.dk_C + row * A.dk_rb + U.c0;
            const uint32_t cfirst = (A.mirror && row > U.c0) ? (uint32_t)min<uint64_t>(row - U.c0, U.W) : 0u;
#pragma unroll 4
            for (uint32_t q = threadIdx.x; q < np; q += ACC_THREADS) {
                const uint32_t cl = 2 * q;
                double2 v = make_double2(init, init);
                if (row < A.spanA && cl + 1 >= cfirst && cl < U.W) {
                    const uint64_t col = (uint64_t)U.c0 + cl;
                    const uint32_t pw = prow[col >> 5];
                    double2 c = *reinterpret_cast<const double2 *>(crow + cl);
                    for (uint32_t z = 1; z < A.dk_splits; z++) {
                        const double2 c2 = *reinterpret_cast<const double2 *>(crow + cl + (uint64_t)z * A.dk_slab);
                        c.x += c2.x;
                        c.y += c2.y;
                    }
                    if ((pw >> (col & 31)) & 1u) v.x = BITMAP ? c.x : c.x + 0.0;
                    if (cl + 1 < U.W && ((pw >> ((col + 1) & 31)) & 1u)) v.y = BITMAP ? c.y : c.y + 0.0;
                }
                out[q] = v;
            }
        }
        if (BITMAP)
            for (uint32_t w = threadIdx.x; w < nslots / 32; w += ACC_THREADS) bits[w] = dk_dense_word(A, U, w);
        return;
    }
    // the presence word and the value are loaded together (every cell of the
    // unit's rows x window lies inside the padded block), several cells per
    // thread in flight: the loop is bound by load latency, not bandwidth
#pragma unroll 4
    for (uint32_t s = threadIdx.x; s < nslots; s += ACC_THREADS) {
        const uint32_t li = s / U.Wp, cl = s - li * U.Wp;
        const uint64_t row = U.row0 + li, col = (uint64_t)U.c0 + cl;
        double v = init;
        if (cl < U.W && row < A.spanA) {
            const uint32_t pw = A.dk_P[row * A.dk_pw + (col >> 5)];
            const double c = dk_value(A, row, col);
            if ((pw >> (col & 31)) & 1u) v = BITMAP ? c : c + 0.0;
        }
        fv[s] = v;
    }
    if (BITMAP)
        for (uint32_t w = threadIdx.x; w < nslots / 32; w += ACC_THREADS) bits[w] = dk_dense_word(A, U, w);
}

// Hash unit start: every present cell of the unit's rows x window goes in
// like a product (count pass: the key; numeric pass: the block value).
template <bool NUM, bool BITMAP>
__device__ __forceinline__ void dk_seed_cells(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S) {
    const uint32_t wpr = (U.W + 31) / 32;
    const uint32_t n = (1u << A.logTL) * wpr;
    for (uint32_t i = threadIdx.x; i < n; i += ACC_THREADS) {
        const uint32_t li = i / wpr, cl0 = (i - li * wpr) * 32;
        const uint64_t row = U.row0 + li;
        if (row >= A.spanA) continue;
        uint32_t w = dk_word(A, row, (uint64_t)U.c0 + cl0);
        if (U.W - cl0 < 32) w &= (1u << (U.W - cl0)) - 1u;
        if (!NUM && U.kind == KIND_CBITS) {
            // rows are padded to 32 cells: the word lands whole
            if (w) atomicOr(&S.keys[(li * U.Wp + cl0) >> 5], w);
            continue;
        }
        while (w) {
            const uint32_t cl = cl0 + (uint32_t)__builtin_ctz(w);
            w &= w - 1;
            const uint32_t mw = A.mirror ? mirror_weight(row, (uint64_t)U.c0 + cl) : 1u;
            if (mw == 0) continue;  // mirror mode: below the diagonal
            const uint64_t gk = row * A.spanB + (U.c0 + cl);
            if (U.kind == KIND_LHASH) {
                const uint32_t key = li * U.W + cl;
                const int slot =
                    NUM ? lds_probe_counted<12>(S.keys, S.cnt, LH_MAX, key)
                        : lds_probe_counted<14>(S.keys, S.cnt, SYM_MAX, key, S.dcnt, mw == 1);
                if (slot >= 0) {
                    if (NUM) lds_add_f64(&S.fv[slot], dk_value(A, row, U.c0 + cl));
                    continue;
                }
            }
            if (NUM) ghash_add(A, gk, dk_value(A, row, U.c0 + cl), U.u);
            else gset_insert(A, gk, U.u, mw == 2);
        }
    }
}

// One unit: heavy groups (a wavefront per group), then the light entries.
template <bool NUM, bool BITMAP, bool DEEP = true>
__device__ __forceinline__ void process_unit(const AccArgs &A, const UnitCtx &U, const Acc<NUM, BITMAP> &S, uint32_t *s_gnext,
                             uint32_t *stage, uint32_t *s_w) {
    const uint32_t gl = A.atile_lg[U.a], g1 = A.atile_g0[U.a + 1];
    if (U.kind != KIND_DENSE && A.dk_tile && A.dk_tile[U.a]) dk_seed_cells<NUM, BITMAP>(A, U, S);
    if (A.row_own && U.kind == KIND_DENSE && (1u << A.logTL) >= (uint32_t)NWAVES) {
        dense_heavy_owned<NUM, BITMAP>(A, U, S);
        __syncthreads();  // plain updates above, atomic updates below
    } else {
        heavy_claimed<NUM, BITMAP, DEEP>(A, U, S, s_gnext);
    }
    light_entries<NUM, BITMAP>(A, U, S, A.group_ptr[gl], A.group_ptr[g1], stage, s_w);
}

// Block-wide ordered compaction over slots [0, n): emit(slot, rank) for every
// slot with pred(slot); returns the number emitted.
template <class Pred, class Emit>
__device__ uint32_t block_compact(uint32_t n, uint32_t *s_w, Pred pred, Emit emit) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t running = 0;
    for (uint32_t base = 0; base < n; base += ACC_THREADS) {
        uint32_t s = base + threadIdx.x;
        bool t = s < n && pred(s);
        uint64_t bal = __ballot(t);
        if (lane == 0) s_w[wave] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t woff = 0, tot = 0;
#pragma unroll
        for (int w = 0; w < NWAVES; w++) {
            uint32_t c = s_w[w];
            woff += (w < (int)wave) ? c : 0;
            tot += c;
        }
        if (t) emit(s, running + woff + (uint32_t)__popcll(bal & lt));
        running += tot;
        __syncthreads();
    }
    return running;
}

template <class Pred>
__device__ uint32_t block_count(uint32_t n, uint32_t *s_w, Pred pred) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t c = 0;
    for (uint32_t s = threadIdx.x; s < n; s += ACC_THREADS) c += pred(s) ? 1 : 0;
    for (int d = 32; d > 0; d >>= 1) c += __shfl_xor(c, d, 64);
    if (lane == 0) s_w[wave] = c;
    __syncthreads();
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < NWAVES; w++) tot += s_w[w];
    __syncthreads();
    return tot;
}

__device__ __forceinline__ void sink_flush(const AccArgs &A, SinkAcc &sk, unsigned long long *s_u, double *s_d) {
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int d = 32; d > 0; d >>= 1) {
        sk.count += __shfl_xor(sk.count, d, 64);
        sk.hash += __shfl_xor(sk.hash, d, 64);
        sk.vsum += __shfl_xor(sk.vsum, d, 64);
        sk.wsum += __shfl_xor(sk.wsum, d, 64);
    }
    if (lane == 0) {
        s_u[2 * wave] = sk.count;
        s_u[2 * wave + 1] = sk.hash;
        s_d[2 * wave] = sk.vsum;
        s_d[2 * wave + 1] = sk.wsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0, h = 0;
        double vs = 0, ws = 0;
        for (int w = 0; w < NWAVES; w++) {
            c += s_u[2 * w];
            h += s_u[2 * w + 1];
            vs += s_d[2 * w];
            ws += s_d[2 * w + 1];
        }
        if (c) sink_atomic(A, sk.group, c, h, vs, ws);
    }
}

// MODE: 0 exact placement, 1 learned placement (cursor)
template <bool BITMAP, bool SINK, int MODE>
__global__ __launch_bounds__(ACC_THREADS, 4) void k_numeric(AccArgs A) {
    __shared__ __align__(16) double s_f64[DENSE_SLOTS];
    __shared__ uint32_t s_bits[BITMAP ? DENSE_SLOTS / 32 : 1];
    __shared__ uint32_t s_cnt, s_gnext, s_unit, s_w[NWAVES];
    __shared__ unsigned long long s_base;
    __shared__ unsigned long long s_su[2 * NWAVES];
    __shared__ double s_sd[2 * NWAVES];
    SinkAcc sk;
    Acc<true, BITMAP> S;
    S.fv = s_f64;
    S.keys = reinterpret_cast<uint32_t *>(s_f64 + LH_SLOTS);
    S.bits = s_bits;
    S.seen = nullptr;
    S.cnt = &s_cnt;
    S.dcnt = nullptr;
    S.sk = &sk;
    S.sink = SINK;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(A.work, 1u);
        __syncthreads();
        const uint32_t ui = s_unit;
        if (ui >= A.nunits) break;
        UnitCtx U;
        U.u = A.unit_order ? A.unit_order[ui] : ui;
        U.a = A.unit_tile[U.u];
        U.tile = A.atile_id[U.a];
        U.c0 = A.unit_c0[U.u];
        U.W = A.unit_W[U.u];
        U.kind = A.unit_kind[U.u];
        U.multi = A.unit_multi[U.u];
        U.Wp = U.kind == KIND_DENSE ? ((U.W + 31) & ~31u) : U.W;
        U.row0 = (uint64_t)U.tile << A.logTL;
        const uint32_t TL = 1u << A.logTL;
        const uint32_t nslots = U.kind == KIND_DENSE ? TL * U.Wp : (U.kind == KIND_LHASH ? LH_SLOTS : 0);
        if (SINK && A.group_rows && U.row0 / A.group_rows != sk.group) {
            // summaries are kept per group: flush when this block moves to another group
            sink_flush(A, sk, s_su, s_sd);
            sk = SinkAcc();
            sk.group = U.row0 / A.group_rows;
            sk.glo = sk.group * A.group_rows;
        }
        if (A.clock_start && ui == 0 && threadIdx.x == 0) *A.clock_start = wall_clock64();
        if (U.kind == KIND_DENSE && A.dk_tile && A.dk_tile[U.a]) {
            dk_seed_dense<BITMAP>(A, U, s_f64, s_bits, nslots);
        } else if (U.kind == KIND_DENSE) {
            const double init = BITMAP ? 0.0 : -0.0;
            for (uint32_t s = threadIdx.x; s < nslots; s += ACC_THREADS) s_f64[s] = init;
            if (BITMAP)
                for (uint32_t s = threadIdx.x; s < (nslots + 31) / 32; s += ACC_THREADS) s_bits[s] = 0;
        } else if (U.kind == KIND_LHASH) {
            for (uint32_t s = threadIdx.x; s < LH_SLOTS; s += ACC_THREADS) {
                s_f64[s] = 0.0;
                S.keys[s] = LEMPTY;
            }
        }
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_gnext = 0;
            if (U.kind == KIND_DIRECT && SINK) {
                const uint32_t t1 = U.u + 1 < A.nunits ? A.unit_tile[U.u + 1] : A.natiles;
                A.unit_cursor[U.u] = A.st_pprefix[t1] - A.st_pprefix[U.a];  // the unit's outputs
            }
            if (U.kind == KIND_DIRECT && !SINK) {
                // the unit's outputs are exactly the products of its run of tiles
                const uint32_t t1 = U.u + 1 < A.nunits ? A.unit_tile[U.u + 1] : A.natiles;
                const unsigned long long n = A.st_pprefix[t1] - A.st_pprefix[U.a];
                unsigned long long b;
                if (MODE == 1) {
                    b = atomicAdd(A.cursor, n);
                    if (b + n > A.capacity) {
                        atomicOr(A.fail_flag, 1u);
                        b = ~0ull;
                    }
                } else {
                    b = A.unit_off[U.u];
                }
                s_base = b;
                if (b != ~0ull) A.unit_cursor[U.u] = n;
            }
        }
        __syncthreads();
        if (U.kind == KIND_DIRECT && !SINK && s_base == ~0ull) continue;  // over capacity: re-run
        U.obase = s_base;
        // flattening stage: LDS the unit's accumulator leaves free (LHASH uses
        // 4096 values + 4096 keys = 6144 doubles' worth, GHASH none)
        uint32_t *stage = nullptr;
        if (U.kind == KIND_LHASH) stage = reinterpret_cast<uint32_t *>(s_f64 + 6144);
        else if (U.kind == KIND_GHASH || U.kind == KIND_DIRECT) stage = reinterpret_cast<uint32_t *>(s_f64);
        else if (nslots + ACC_THREADS <= DENSE_SLOTS) stage = reinterpret_cast<uint32_t *>(s_f64 + DENSE_SLOTS - ACC_THREADS);
        if (U.kind == KIND_DIRECT) {
            const uint32_t t1 = U.u + 1 < A.nunits ? A.unit_tile[U.u + 1] : A.natiles;
            direct_run<BITMAP>(A, U, S, A.group_ptr[A.atile_g0[U.a]], A.group_ptr[A.atile_g0[t1]], stage, s_w);
        } else {
            // count + checksum and bitmap variants carry more state: the
            // shallow dense ring keeps them at two workgroups per CU
            process_unit<true, BITMAP, !SINK && !BITMAP>(A, U, S, &s_gnext, stage, s_w);
        }
        __syncthreads();
        if (U.kind == KIND_DENSE || U.kind == KIND_LHASH) {
            auto pred = [&](uint32_t s) -> bool {
                if (U.kind == KIND_DENSE) {
                    if (BITMAP) return (s_bits[s >> 5] >> (s & 31)) & 1u;
                    return __double_as_longlong(s_f64[s]) != (long long)NEG_ZERO_BITS;
                }
                return S.keys[s] < LSEALED;
            };
            auto cell = [&](uint32_t s, uint64_t &r, uint64_t &c) {
                const uint32_t key = U.kind == KIND_DENSE ? s : S.keys[s];
                const uint32_t li = key / U.Wp, cl = key - li * U.Wp;
                r = U.row0 + li;
                c = (uint64_t)U.c0 + cl;
            };
            // outputs slot s stands for: 0 untouched; in mirror mode 0 below
            // the diagonal, 1 on it, 2 (the cell and its transpose) above
            auto weight = [&](uint32_t s) -> uint32_t {
                if (!pred(s)) return 0u;
                if (!A.mirror) return 1u;
                uint64_t r, c;
                cell(s, r, c);
                return mirror_weight(r, c);
            };
            // Each wavefront owns a contiguous run of slots (whole 64-slot
            // chunks), so the unit's outputs keep slot order: count per
            // wavefront, one barrier for the offsets, then emit without
            // further barriers. Count + checksum outputs need no offsets.
            // A 64-slot chunk writes its cells, then their transposes.
            const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            const uint32_t per = ((nslots + NWAVES * 64 - 1) / (NWAVES * 64)) * 64;
            const uint32_t w0 = min(nslots, wave * per), w1 = min(nslots, w0 + per);
            if (SINK) {
                uint32_t n = 0;
                for (uint32_t s = w0 + lane; s < w1; s += 64) {
                    const uint32_t w = weight(s);
                    if (!w) continue;
                    n += w;
                    uint64_t r, c;
                    cell(s, r, c);
                    emit_output(A, 0, r, c, s_f64[s], true, sk);
                    if (w == 2) emit_output(A, 0, c, r, s_f64[s], true, sk);
                }
                // the unit's outputs (plan queries, planner error)
                for (int d = 32; d > 0; d >>= 1) n += __shfl_xor(n, d, 64);
                if (lane == 0) s_w[wave] = n;
                __syncthreads();
                if (threadIdx.x == 0) {
                    uint32_t tot = 0;
#pragma unroll
                    for (int w = 0; w < NWAVES; w++) tot += s_w[w];
                    A.unit_cursor[U.u] = tot;
                }
            } else {
                const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
                uint32_t cnt = 0;
                for (uint32_t b = w0; b < w1; b += 64) {
                    const uint32_t w = b + lane < w1 ? weight(b + lane) : 0u;
                    cnt += (uint32_t)__popcll(__ballot(w != 0)) + (uint32_t)__popcll(__ballot(w == 2));
                }
                if (lane == 0) s_w[wave] = cnt;
                __syncthreads();
                uint32_t pre = 0, tot = 0;
#pragma unroll
                for (int w = 0; w < NWAVES; w++) {
                    const uint32_t c = s_w[w];
                    pre += (w < (int)wave) ? c : 0;
                    tot += c;
                }
                if (threadIdx.x == 0) {
                    unsigned long long b;
                    if (MODE == 1) {
                        b = atomicAdd(A.cursor, (unsigned long long)tot);
                        if (b + tot > A.capacity) {
                            atomicOr(A.fail_flag, 1u);
                            b = ~0ull;
                        }
                    } else {
                        b = A.unit_off[U.u];
                    }
                    s_base = b;
                    if (b != ~0ull) A.unit_cursor[U.u] = tot;
                }
                __syncthreads();
                const uint64_t base = s_base;
                if (base != ~0ull) {
                    uint32_t run = pre;
                    for (uint32_t b = w0; b < w1; b += 64) {
                        const uint32_t w = b + lane < w1 ? weight(b + lane) : 0u;
                        const uint64_t bal = __ballot(w != 0), bal2 = __ballot(w == 2);
                        const uint32_t n1 = (uint32_t)__popcll(bal);
                        if (w) {
                            uint64_t r, c;
                            cell(b + lane, r, c);
                            const double v = s_f64[b + lane];
                            emit_output(A, base + run + (uint32_t)__popcll(bal & lt), r, c, v, false, sk);
                            if (w == 2)
                                emit_output(A, base + run + n1 + (uint32_t)__popcll(bal2 & lt), c, r, v, false, sk);
                        }
                        run += n1 + (uint32_t)__popcll(bal2);
                    }
                }
            }
        } else if (U.kind == KIND_GHASH && threadIdx.x == 0) {
            A.unit_cursor[U.u] = 0;  // the global-hash drain counts the unit's outputs
        }
        if (A.group_left && threadIdx.x == 0) {
            const uint64_t g = U.row0 / A.group_rows;
            __threadfence();
            if (atomicSub(&A.group_left[g], 1u) == 1u) A.group_done[g] = wall_clock64();
        }
        __syncthreads();
    }
    if (SINK) sink_flush(A, sk, s_su, s_sd);
}

static_assert(DENSE_SLOTS <= 4 * SYM_SLOTS, "the count pass's dense byte map must fit its key set");

// Count pass: structural support only. The key set of hash units (or the
// byte map of dense units) is dynamic LDS of A.set_words words: a launch with
// no LDS-hash units sizes it for the dense byte map alone (DENSE_SLOTS bytes),
// so twice as many workgroups share a CU.
__global__ __launch_bounds__(ACC_THREADS) void k_count(AccArgs A) {
    extern __shared__ __align__(16) uint32_t s_set[];
    __shared__ uint32_t s_stage[2 * ACC_THREADS];
    __shared__ uint32_t s_cnt, s_dcnt, s_gnext, s_unit, s_w[NWAVES];
    Acc<false, false> S;
    S.fv = nullptr;
    S.keys = s_set;
    S.bits = nullptr;
    S.seen = reinterpret_cast<uint8_t *>(s_set);
    S.cnt = &s_cnt;
    S.dcnt = A.mirror ? &s_dcnt : nullptr;
    S.sk = nullptr;
    S.sink = false;
    for (;;) {
        if (threadIdx.x == 0) s_unit = atomicAdd(A.work, 1u);
        __syncthreads();
        const uint32_t ui = s_unit;
        if (ui >= A.nunits) break;
        UnitCtx U;
        U.u = A.unit_order ? A.unit_order[ui] : ui;
        U.a = A.unit_tile[U.u];
        U.tile = A.atile_id[U.a];
        U.c0 = A.unit_c0[U.u];
        U.W = A.unit_W[U.u];
        U.kind = A.unit_kind[U.u];
        // an LDS-hash unit in a launch without room for its key set counts
        // through the global set (the same count)
        if (U.kind == KIND_LHASH && A.set_words < SYM_SLOTS) U.kind = KIND_GHASH;
        U.multi = A.unit_multi[U.u];
        // A one-window (all columns) hash unit whose rows x padded width fit
        // the launch's key-set LDS as one bit per cell counts through that
        // bit set: the number of distinct cells does not depend on the
        // accumulator the numeric pass will use, and an LDS bit beats a probe
        // (for the global set, a compare-and-swap in HBM).
        const uint64_t cells = ((uint64_t)1 << A.logTL) * (((uint64_t)U.W + 31) & ~31ull);
        if ((U.kind == KIND_GHASH || U.kind == KIND_LHASH) && !U.multi && cells <= (uint64_t)A.set_words * 32)
            U.kind = KIND_CBITS;
        U.Wp = (U.kind == KIND_DENSE || U.kind == KIND_CBITS) ? ((U.W + 31) & ~31u) : U.W;
        U.row0 = (uint64_t)U.tile << A.logTL;
        if (U.kind == KIND_DIRECT) {  // every product is a distinct output
            if (threadIdx.x == 0) {
                const uint32_t t1 = U.u + 1 < A.nunits ? A.unit_tile[U.u + 1] : A.natiles;
                A.unit_count[U.u] = A.st_pprefix[t1] - A.st_pprefix[U.a];
            }
            // every thread has read s_unit before thread 0 claims the next unit
            __syncthreads();
            continue;
        }
        const uint32_t TL = 1u << A.logTL;
        // dense units: a byte per slot, read back four slots per word
        const uint32_t nwords = U.kind == KIND_DENSE ? TL * U.Wp / 4 : 0;
        if (U.kind == KIND_DENSE) {
            const bool seeded = A.dk_tile && A.dk_tile[U.a];
            if (seeded) {
                // presence bits of the block, one nibble per byte word
                uint32_t have = 0;
                for (uint32_t w = threadIdx.x; w < nwords / 8; w += ACC_THREADS) {
                    const uint32_t b = dk_dense_word(A, U, w);
                    have += (uint32_t)__popc(b);
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const uint32_t n = (b >> (4 * j)) & 0xFu;
                        s_set[8 * w + j] = (n & 1u) | ((n & 2u) << 7) | ((n & 4u) << 14) | ((n & 8u) << 21);
                    }
                }
                // A unit whose every cell the block already holds has nothing
                // left to count: its sparse products can only land on present
                // cells. Its count is the cell count, without the traversal.
                for (int d = 32; d > 0; d >>= 1) have += __shfl_xor(have, d, 64);
                if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = have;
                __syncthreads();
                uint32_t tot = 0;
#pragma unroll
                for (int w = 0; w < NWAVES; w++) tot += s_w[w];
                __syncthreads();  // s_w is written again below
                const uint64_t rows_in = min((uint64_t)TL, A.spanA - U.row0);
                if ((uint64_t)tot == rows_in * U.W) {
                    if (threadIdx.x == 0) {
                        unsigned long long n = tot;
                        if (A.mirror) {
                            // every cell present: per row, the cells at or right
                            // of the diagonal, twice off it
                            n = 0;
                            const uint64_t ce = (uint64_t)U.c0 + U.W;
                            for (uint64_t r = U.row0; r < U.row0 + rows_in; r++) {
                                const uint64_t lo = r > U.c0 ? r : U.c0;
                                if (lo < ce) n += 2 * (ce - lo) - (r >= U.c0 ? 1 : 0);
                            }
                        }
                        atomicAdd(&A.unit_count[U.u], n);
                        atomicAdd(A.block_counted, 1ull);
                    }
                    continue;  // every thread has read s_unit (barriers above)
                }
            } else {
                for (uint32_t s = threadIdx.x; s < nwords; s += ACC_THREADS) s_set[s] = 0u;
            }
        } else if (U.kind == KIND_CBITS) {
            for (uint32_t s = threadIdx.x; s < TL * U.Wp / 32; s += ACC_THREADS) s_set[s] = 0u;
        } else if (U.kind == KIND_LHASH) {
            for (uint32_t s = threadIdx.x; s < SYM_SLOTS; s += ACC_THREADS) s_set[s] = LEMPTY;
        }
        if (threadIdx.x == 0) {
            s_cnt = 0;
            s_dcnt = 0;
            s_gnext = 0;
        }
        __syncthreads();
        process_unit<false, false>(A, U, S, &s_gnext, s_stage, s_w);
        __syncthreads();
        uint32_t c;
        if (U.kind == KIND_DENSE) {
            uint32_t part = 0;
            if (!A.mirror) {
                for (uint32_t s = threadIdx.x; s < nwords; s += ACC_THREADS) part += __popc(s_set[s]);
            } else {
                // the four slots of a word share a row (rows hold a multiple of
                // 32 slots); each touched slot counts its mirror weight
                for (uint32_t s = threadIdx.x; s < nwords; s += ACC_THREADS) {
                    const uint32_t wd = s_set[s];
                    if (!wd) continue;
                    const uint32_t li = (4 * s) / U.Wp, cl = 4 * s - li * U.Wp;
                    const uint64_t r = U.row0 + li, c0 = (uint64_t)U.c0 + cl;
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if ((wd >> (8 * j)) & 1u) part += mirror_weight(r, c0 + j);
                }
            }
            for (int d = 32; d > 0; d >>= 1) part += __shfl_xor(part, d, 64);
            if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = part;
            __syncthreads();
            c = 0;
            for (int w = 0; w < NWAVES; w++) c += s_w[w];
        } else if (U.kind == KIND_CBITS) {
            uint32_t part = 0;
            const uint32_t nb = TL * U.Wp / 32;
            for (uint32_t s = threadIdx.x; s < nb; s += ACC_THREADS) {
                const uint32_t wd = s_set[s];
                if (!wd) continue;
                if (!A.mirror) {
                    part += __popc(wd);
                    continue;
                }
                // a word's 32 cells share a row: twice right of the diagonal,
                // once on it
                const uint32_t li = (32 * s) / U.Wp, cl = 32 * s - li * U.Wp;
                const uint64_t r = U.row0 + li, c0 = (uint64_t)U.c0 + cl;
                if (c0 > r) {
                    part += 2 * __popc(wd);
                } else if (r - c0 < 32) {
                    const uint32_t d = (uint32_t)(r - c0);
                    const uint32_t gt = d == 31 ? 0u : (~0u << (d + 1));
                    part += 2 * __popc(wd & gt) + ((wd >> d) & 1u);
                }
            }
            for (int d = 32; d > 0; d >>= 1) part += __shfl_xor(part, d, 64);
            if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = part;
            __syncthreads();
            c = 0;
            for (int w = 0; w < NWAVES; w++) c += s_w[w];
        } else {
            // LDS-hash keys (mirror mode: all at or right of the diagonal,
            // twice except the diagonal ones)
            c = A.mirror ? 2 * s_cnt - s_dcnt : s_cnt;
        }
        if (threadIdx.x == 0 && c) atomicAdd(&A.unit_count[U.u], (unsigned long long)c);
        __syncthreads();
    }
}

// Drains the global hash: GHASH units and sealed-LHASH overflow.
template <bool SINK, int MODE>
__global__ __launch_bounds__(256) void k_gdrain(AccArgs A) {
    __shared__ unsigned long long s_su[2 * 4];
    __shared__ double s_sd[2 * 4];
    __shared__ unsigned long long s_base;
    __shared__ uint32_t s_w[4];
    SinkAcc sk;
    const uint64_t cap = A.gmask + 1;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint64_t base = (uint64_t)blockIdx.x * 256; base < cap; base += (uint64_t)gridDim.x * 256) {
        uint64_t h = base + threadIdx.x;
        unsigned long long key = h < cap ? A.gkeys[h] : GEMPTY;
        bool t = key != GEMPTY;
        uint64_t row = t ? key / A.spanB : 0, col = t ? key - row * A.spanB : 0;
        double v = t ? A.gvals[h] : 0.0;
        // mirror mode: every key is at or right of the diagonal; off it the
        // transposed cell is written too
        const bool twice = t && A.mirror && col > row;
        if (SINK) {
            if (A.group_rows) {
                // per-group summaries: one set of atomics per wave when the
                // wave's entries share a group, else per entry
                const uint64_t g = t ? row / A.group_rows : 0;
                const uint64_t lrow = row - g * A.group_rows;
                const uint64_t z = t ? mix64(lrow * A.spanB + col) : 0;
                const uint64_t act = __ballot(t);
                if (!act) continue;
                const uint64_t g0 = __shfl(g, __ffsll((unsigned long long)act) - 1, 64);
                if (__ballot(t && g != g0) == 0) {
                    unsigned long long c = t ? 1ull : 0ull, hz = z;
                    double vs = v, ws = t ? v * hash_weight(z) : 0.0;
                    for (int d = 32; d > 0; d >>= 1) {
                        c += __shfl_xor(c, d, 64);
                        hz += __shfl_xor(hz, d, 64);
                        vs += __shfl_xor(vs, d, 64);
                        ws += __shfl_xor(ws, d, 64);
                    }
                    if (lane == 0) sink_atomic(A, g0, c, hz, vs, ws);
                } else if (t) {
                    sink_atomic(A, g, 1ull, z, v, v * hash_weight(z));
                }
                if (t) atomicAdd(&A.unit_cursor[A.gunit[h]], 1ull);  // per-unit outputs
                continue;
            }
            if (t) emit_output(A, 0, row, col, v, true, sk);
            if (twice) emit_output(A, 0, col, row, v, true, sk);
            if (t) atomicAdd(&A.unit_cursor[A.gunit[h]], twice ? 2ull : 1ull);  // per-unit outputs
            continue;
        }
        if (MODE == 0) {
            if (t) {
                uint32_t u = A.gunit[h];
                unsigned long long r = atomicAdd(&A.unit_cursor[u], twice ? 2ull : 1ull);
                emit_output(A, A.unit_off[u] + r, row, col, v, false, sk);
                if (twice) emit_output(A, A.unit_off[u] + r + 1, col, row, v, false, sk);
            }
        } else {
            const uint64_t bal = __ballot(t), bal2 = __ballot(twice);
            if (lane == 0) s_w[wave] = (uint32_t)(__popcll(bal) + __popcll(bal2));
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
                unsigned long long b = tot ? atomicAdd(A.cursor, (unsigned long long)tot) : 0;
                if (tot && b + tot > A.capacity) {
                    atomicOr(A.fail_flag, 1u);
                    b = ~0ull;
                }
                s_base = b;
            }
            __syncthreads();
            uint32_t woff = 0;
            for (uint32_t w = 0; w < wave; w++) woff += s_w[w];
            if (t && s_base != ~0ull) {
                emit_output(A, s_base + woff + (uint32_t)__popcll(bal & lt), row, col, v, false, sk);
                if (twice)
                    emit_output(A, s_base + woff + (uint32_t)__popcll(bal) + (uint32_t)__popcll(bal2 & lt), col, row,
                                v, false, sk);
            }
            // per-unit actual count (planner error)
            if (t) atomicAdd(&A.unit_cursor[A.gunit[h]], twice ? 2ull : 1ull);
            __syncthreads();
        }
    }
    if (SINK) {
        for (int d = 32; d > 0; d >>= 1) {
            sk.count += __shfl_xor(sk.count, d, 64);
            sk.hash += __shfl_xor(sk.hash, d, 64);
            sk.vsum += __shfl_xor(sk.vsum, d, 64);
            sk.wsum += __shfl_xor(sk.wsum, d, 64);
        }
        if (lane == 0 && sk.count) sink_atomic(A, 0, sk.count, sk.hash, sk.vsum, sk.wsum);
    }
}

// --------------------------------------------------------------- planner

// Per active tile: accumulator kind and number of column windows.
// est = exact planner: the products bound; learned planner: predicted outputs.
// In mirror mode the windows wholly left of the tile's first row hold no kept
// cell and are skipped: t_w0 is the first window kept.
__global__ void k_plan_tiles(const uint64_t *st_products, const uint64_t *st_maxgp, const uint8_t *st_direct,
                             const float *pred, uint32_t natiles, PlanParams P, uint32_t *t_nwin, uint32_t *t_W,
                             uint8_t *t_kind, uint32_t *t_w0) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natiles) return;
    t_w0[a] = 0;
    // distinct outputs: one window, no accumulator (the unit's LDS output
    // counter is 32-bit, so a DIRECT tile holds fewer than 2^31 products)
    if (!P.force && st_direct && st_direct[a] && st_products[a] < (1ull << 31)) {
        t_nwin[a] = 1;
        t_W[a] = (uint32_t)P.spanB;
        t_kind[a] = KIND_DIRECT;
        return;
    }
    const double prod = (double)st_products[a];
    const double lb = (double)st_maxgp[a];  // one group's products are distinct outputs
    const double S = (double)P.spanB;
    const double area = (double)P.TL * S;
    // dense rows are padded to a multiple of 32 columns
    const double parea = (double)P.TL * (double)((P.spanB + 31) & ~31ull);
    const bool dense_ok = (uint64_t)P.TL * 32 <= DENSE_SLOTS;
    double est = pred ? (double)pred[a] * (1.0 + P.margin) + 8.0 : prod;
    if (est < lb) est = lb;
    if (est > prod) est = prod;
    if (est > area) est = area;
    uint32_t kind, nwin = 1;
    uint64_t W = P.spanB;
    const int force = P.force;
    bool dense;
    if (!dense_ok) dense = false;
    else if (force) dense = force == KIND_DENSE;
    else if (est <= (double)HASH_TARGET) dense = false;
    else dense = parea <= (double)DENSE_SLOTS || est >= P.dense_ratio * area;
    if (dense) {
        kind = KIND_DENSE;
        if (parea > (double)DENSE_SLOTS) {
            W = P.Wd;
            nwin = (uint32_t)((P.spanB + W - 1) / W);
        }
    } else {
        kind = force == KIND_GHASH ? KIND_GHASH : KIND_LHASH;
        double nw = kind == KIND_GHASH ? 1.0 : ceil(est / (double)HASH_TARGET);
        if (nw < 1) nw = 1;
        if (nw > S) nw = S;
        nwin = (uint32_t)nw;
        W = (P.spanB + nwin - 1) / nwin;
        // local keys li*W + col must stay below the LDS sentinels
        while ((uint64_t)P.TL * W >= 0xFFFFFFF0ull) {
            nwin *= 2;
            W = (P.spanB + nwin - 1) / nwin;
        }
        nwin = (uint32_t)((P.spanB + W - 1) / W);
    }
    if (P.mirror && nwin > 1) {
        const uint64_t row0 = (uint64_t)P.atile_id[a] << P.logTL;
        const uint32_t w0 = (uint32_t)(row0 / W);  // window holding column row0 (row0 < spanB)
        t_w0[a] = w0;
        nwin -= w0;
    }
    t_nwin[a] = nwin;
    t_W[a] = (uint32_t)W;
    t_kind[a] = (uint8_t)kind;
}

// Consecutive DIRECT tiles form one unit of about DIRECT_RUN products (no LDS
// is involved, so units can be large); a run breaks at a non-DIRECT tile, at
// each multiple of `run` cumulative products and at a row-group boundary.
__global__ void k_direct_runs(const uint8_t *t_kind, const uint64_t *pp, const uint32_t *atile_id, uint32_t natiles,
                              int logTL, uint64_t group_rows, uint64_t run, uint32_t *t_nwin) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natiles || t_kind[a] != KIND_DIRECT) return;
    bool start = true;
    if (run && a > 0 && t_kind[a - 1] == KIND_DIRECT) {
        start = pp[a] / run != pp[a - 1] / run;
        if (!start && group_rows)
            start = ((uint64_t)atile_id[a] << logTL) / group_rows != ((uint64_t)atile_id[a - 1] << logTL) / group_rows;
    }
    t_nwin[a] = start ? 1u : 0u;
}

__global__ void k_emit_units(const uint32_t *t_nwin, const uint32_t *t_W, const uint8_t *t_kind,
                             const uint32_t *t_ubase, const uint32_t *t_w0, uint32_t natiles, uint64_t spanB,
                             const float *pred, uint32_t *u_tile, uint32_t *u_c0, uint32_t *u_W, uint8_t *u_kind,
                             uint8_t *u_multi, float *u_pred) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natiles) return;
    uint32_t nw = t_nwin[a], W = t_W[a], base = t_ubase[a], w0 = t_w0[a];
    for (uint32_t w = 0; w < nw; w++) {
        uint64_t c0 = (uint64_t)(w0 + w) * W;
        uint64_t wid = spanB - c0 < W ? spanB - c0 : W;
        u_tile[base + w] = a;
        u_c0[base + w] = (uint32_t)c0;
        u_W[base + w] = (uint32_t)wid;
        u_kind[base + w] = t_kind[a];
        u_multi[base + w] = wid < spanB;
        if (u_pred) u_pred[base + w] = pred ? pred[a] / (float)nw : 0.0f;
    }
}

// Exact planner, after the count pass: hash units whose exact count exceeds
// the LDS hash go to the global hash; returns their total through *gtotal.
// In mirror mode a unit's count takes each kept cell twice except those on
// the diagonal (at most one per row), so its distinct cells are at most
// (count + min(TL, W)) / 2.
__global__ void k_refine_kinds(const unsigned long long *unit_count, const uint32_t *u_W, uint32_t n, int force,
                               int mirror, uint32_t TL, uint8_t *u_kind, unsigned long long *gtotal) {
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    unsigned long long cells = unit_count[u];
    if (mirror) cells = (cells + min(TL, u_W[u])) / 2;
    if (u_kind[u] == KIND_LHASH && cells > LH_MAX && force != KIND_LHASH) u_kind[u] = KIND_GHASH;
    if (u_kind[u] == KIND_GHASH) atomicAdd(gtotal, unit_count[u]);
    // a forced LDS-hash unit over the limit seals its table once it holds
    // LH_MAX keys: the rest of its cells go to the global hash, sized for them
    else if (u_kind[u] == KIND_LHASH && cells > LH_MAX) atomicAdd(gtotal, cells - LH_MAX);
}

// fcut[f * n + w] = the first position of fiber f whose column is >= w * W
// (the fiber's end when none), w < n. One thread per fiber: its n cuts are
// consecutive words, so a wavefront's stores cover one contiguous run. The
// fiber's first and last columns settle every cut outside them (the last
// cut, past every column, among them) without a search; a cut inside advances
// from the previous one over the fiber's sorted columns, by a binary search
// down to 8 positions, then counts the columns below the cut in those 8 with
// independent loads.
__global__ void k_fiber_cuts(const uint32_t *fiber_ptr, const uint32_t *Bcol, uint32_t nfib, uint32_t n, uint32_t W,
                             uint32_t *fcut) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfib) return;
    uint32_t lo = fiber_ptr[f];
    const uint32_t hi = fiber_ptr[f + 1];
    uint32_t *fc = fcut + (size_t)f * n;
    fc[0] = lo;
    const uint32_t first = lo < hi ? Bcol[lo] : 0u, last = lo < hi ? Bcol[hi - 1] : 0u;
    for (uint32_t w = 1; w < n; w++) {
        const uint64_t c = (uint64_t)w * W;
        uint32_t a = lo, b = hi;
        if (a == b || (uint64_t)last < c) {
            a = hi;
        } else if ((uint64_t)first >= c) {
            a = lo;
        } else {
            while (b - a > 8) {
                const uint32_t mid = a + (b - a) / 2;
                if (Bcol[mid] < c) a = mid + 1;
                else b = mid;
            }
            uint32_t k = 0;
#pragma unroll
            for (uint32_t j = 0; j < 8; j++) k += (a + j < b && Bcol[a + j] < c) ? 1u : 0u;
            a += k;
        }
        fc[w] = a;
        lo = a;
    }
}

// ------------------------------------------------------------ launchers

void launch_fiber_cuts(const Index &ix, uint32_t n, uint32_t W, uint32_t *fcut, hipStream_t s) {
    if (!ix.nfib || !n || !ix.nB) return;
    k_fiber_cuts<<<(unsigned)div_up(ix.nfib, 256), 256, 0, s>>>(ix.fiber_ptr, ix.Bcol, ix.nfib, n, W, fcut);
}

void launch_plan_tiles(const Index &ix, const float *pred, const PlanParams &P, uint32_t *t_nwin, uint32_t *t_W,
                       uint8_t *t_kind, uint32_t *t_w0, hipStream_t s) {
    if (!ix.natiles) return;
    k_plan_tiles<<<(unsigned)div_up(ix.natiles, 256), 256, 0, s>>>((const uint64_t *)ix.st_products,
                                                                     (const uint64_t *)ix.st_maxgp, ix.st_direct, pred,
                                                                     ix.natiles, P, t_nwin, t_W, t_kind, t_w0);
    k_direct_runs<<<(unsigned)div_up(ix.natiles, 256), 256, 0, s>>>(t_kind, (const uint64_t *)ix.st_pprefix,
                                                                      ix.atile_id, ix.natiles, ix.logTL, P.group_rows,
                                                                      P.run, t_nwin);
}

void launch_emit_units(const uint32_t *t_nwin, const uint32_t *t_W, const uint8_t *t_kind, const uint32_t *t_ubase,
                       const uint32_t *t_w0, uint32_t natiles, uint64_t spanB, const float *pred, Units &U,
                       hipStream_t s) {
    if (!natiles) return;
    k_emit_units<<<(unsigned)div_up(natiles, 256), 256, 0, s>>>(t_nwin, t_W, t_kind, t_ubase, t_w0, natiles, spanB,
                                                                  pred, U.tile, U.c0, U.W, U.kind, U.multi, U.pred);
}

void launch_refine_kinds(const unsigned long long *unit_count, const uint32_t *u_W, uint32_t n, int force, int mirror,
                         uint32_t TL, uint8_t *u_kind, unsigned long long *gtotal, hipStream_t s) {
    if (!n) return;
    k_refine_kinds<<<(unsigned)div_up(n, 256), 256, 0, s>>>(unit_count, u_W, n, force, mirror, TL, u_kind, gtotal);
}

void launch_count(const AccArgs &A, unsigned grid, hipStream_t s) {
    if (!A.nunits) return;
    if (A.set_words != SYM_SLOTS && A.set_words != DENSE_SLOTS / 4)
        throw LibError{ERR_INVALID, "count pass: unsupported key-set size"};
    k_count<<<grid, ACC_THREADS, (size_t)A.set_words * 4, s>>>(A);
}

void launch_numeric(const AccArgs &A, unsigned grid, bool bitmap, bool sink, int mode, hipStream_t s) {
    if (!A.nunits) return;
#define FCC_NUM(BM, SK, MD) k_numeric<BM, SK, MD><<<grid, ACC_THREADS, 0, s>>>(A)
    if (bitmap) {
        if (sink) { if (mode) FCC_NUM(true, true, 1); else FCC_NUM(true, true, 0); }
        else { if (mode) FCC_NUM(true, false, 1); else FCC_NUM(true, false, 0); }
    } else {
        if (sink) { if (mode) FCC_NUM(false, true, 1); else FCC_NUM(false, true, 0); }
        else { if (mode) FCC_NUM(false, false, 1); else FCC_NUM(false, false, 0); }
    }
#undef FCC_NUM
}

void launch_gdrain(const AccArgs &A, bool sink, int mode, hipStream_t s) {
    uint64_t cap = A.gmask + 1;
    unsigned grid = (unsigned)std::min<uint64_t>(div_up(cap, 256), 4096);
    if (sink) k_gdrain<true, 0><<<grid, 256, 0, s>>>(A);
    else if (mode) k_gdrain<false, 1><<<grid, 256, 0, s>>>(A);
    else k_gdrain<false, 0><<<grid, 256, 0, s>>>(A);
}

}  // namespace fcc
