// Index build: extents, linearisation, right-operand CSR keyed by contraction
// key, left-operand tiles over the external index, per-tile statistics.
//
// Replaces TileIndexedTensor (index_tensors.hpp:491-547) and the reference's
// linearisers CoOrdinate::gather_linearize / remove_linearize
// (coordinate.hpp:723-733, 767-781). Instead of one hash map per tile, both
// operands are sorted on composite keys:
//   B: (key, col)                     -> fibers, CSR by contraction key
//   A: (tile, class, fiber, row%TL)   -> per tile: heavy groups, then light entries
#include "fcc_index.h"

namespace fcc {

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, size_t n, unsigned long long v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double *p, size_t n, double v) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

static unsigned fill_grid(size_t n) {
    size_t b = div_up(n ? n : 1, 256);
    return (unsigned)(b > 8192 ? 8192 : b);
}
__global__ void k_map_u32(const uint32_t *map, const uint32_t *idx, size_t n, uint32_t *out) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = map[idx[i]];
}

__global__ void k_perm_gather(const double *src, const uint32_t *perm, size_t n, double *dst) {
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
        dst[p] = src[perm[p]];
}

void perm_gather(const double *src, const uint32_t *perm, size_t n, double *dst, hipStream_t s) {
    if (n) k_perm_gather<<<(unsigned)std::min<size_t>((n + 255) / 256, 8192), 256, 0, s>>>(src, perm, n, dst);
}

void fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
    if (n) k_fill_u32<<<fill_grid(n), 256, 0, s>>>(p, n, v);
}

// Several fills in one launch: blockIdx.y picks the buffer (a short chain of
// small fills otherwise costs one launch each, exposed after a host wait).
__global__ void k_fill_many(FillSet f) {
    const FillOp op = f.op[blockIdx.y];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < op.n; i += (size_t)gridDim.x * blockDim.x) {
        if (op.bytes == 8) static_cast<unsigned long long *>(op.p)[i] = op.v;
        else static_cast<uint32_t *>(op.p)[i] = (uint32_t)op.v;
    }
}

void fill_many(const FillSet &f, hipStream_t s) {
    if (f.n <= 0) return;
    size_t most = 0;
    for (int i = 0; i < f.n; i++) most = std::max<size_t>(most, f.op[i].n);
    k_fill_many<<<dim3(fill_grid(most), (unsigned)f.n), 256, 0, s>>>(f);
}
void fill_u64(unsigned long long *p, size_t n, unsigned long long v, hipStream_t s) {
    if (n) k_fill_u64<<<fill_grid(n), 256, 0, s>>>(p, n, v);
}
void fill_f64(double *p, size_t n, double v, hipStream_t s) {
    if (n) k_fill_f64<<<fill_grid(n), 256, 0, s>>>(p, n, v);
}

// ---------------------------------------------------------------- extents
// out[m] = max coordinate of mode m (the reference's shape, contract.hpp:124-144);
// out[order] |= 1 if some value is zero, subnormal, or smaller than 2^-511 in
// magnitude (then a product can round to zero and the accumulators must track
// touched slots explicitly); |= 2 if some value is infinite or NaN (then no key
// may go through the dense-key GEMM, whose zero padding would turn inf * 0
// into NaN at cells the reference never touches with that value).
__global__ __launch_bounds__(256) void k_extents(const uint32_t *coords, const double *vals, int order, size_t nnz,
                                                uint32_t *out) {
    __shared__ uint32_t s_mx[4], s_fl[4];
    const int m = blockIdx.y;
    uint32_t mx = 0, flag = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t gid = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    // 16-byte loads over the aligned body of the mode's column (or of the
    // values), four in flight per thread; the few elements before the first
    // 16-byte boundary and after the last whole vector one at a time
    if (m < order) {
        const uint32_t *p = coords + (size_t)m * nnz;
        const size_t head = std::min<size_t>(nnz, ((16 - ((uintptr_t)p & 15)) & 15) / 4);
        const uint4 *q = reinterpret_cast<const uint4 *>(p + head);
        const size_t nq = (nnz - head) / 4;
        auto take4 = [&](uint4 w) {
            mx = max(mx, max(max(w.x, w.y), max(w.z, w.w)));
        };
        size_t j = gid;
        for (; j + 3 * stride < nq; j += 4 * stride) {
            uint4 w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) w[u] = q[j + u * stride];
#pragma unroll
            for (int u = 0; u < 4; u++) take4(w[u]);
        }
        for (; j < nq; j += stride) take4(q[j]);
        if (gid < head) mx = max(mx, p[gid]);
        for (size_t i = head + 4 * nq + gid; i < nnz; i += stride) mx = max(mx, p[i]);
    } else {
        auto take = [&](double v) {
            if (!(fabs(v) >= 0x1.0p-511) && !isnan(v)) flag |= 1;
            if (!isfinite(v)) flag |= 2;
        };
        const size_t head = std::min<size_t>(nnz, ((16 - ((uintptr_t)vals & 15)) & 15) / 8);
        const double2 *q = reinterpret_cast<const double2 *>(vals + head);
        const size_t nq = (nnz - head) / 2;
        size_t j = gid;
        for (; j + 3 * stride < nq; j += 4 * stride) {
            double2 w[4];
#pragma unroll
            for (int u = 0; u < 4; u++) w[u] = q[j + u * stride];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                take(w[u].x);
                take(w[u].y);
            }
        }
        for (; j < nq; j += stride) {
            take(q[j].x);
            take(q[j].y);
        }
        if (gid < head) take(vals[gid]);
        for (size_t i = head + 2 * nq + gid; i < nnz; i += stride) take(vals[i]);
    }
    for (int d = 32; d > 0; d >>= 1) {
        uint32_t o = __shfl_xor(mx, d, 64);
        mx = o > mx ? o : mx;
        flag |= __shfl_xor(flag, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_mx[threadIdx.x >> 6] = mx;
        s_fl[threadIdx.x >> 6] = flag;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; w++) {
            mx = s_mx[w] > mx ? s_mx[w] : mx;
            flag |= s_fl[w];
        }
        // most blocks hold the maximum the first ones already stored: read
        // first, so same-address atomics (serialised in L2) stay few
        if (m < order) {
            if (mx > __atomic_load_n(&out[m], __ATOMIC_RELAXED)) atomicMax(&out[m], mx);
        } else if (flag) {
            atomicOr(&out[order], flag);
        }
    }
}

void launch_extents(const Operand &t, uint32_t *dev_out, hipStream_t s) {
    if (t.nnz == 0) return;
    // at most 512 workgroups per mode: each ends in one atomic on the mode's
    // maximum, and same-address atomics serialise in L2
    unsigned gx = (unsigned)std::min<size_t>(div_up((size_t)t.nnz, 256 * 16), 512);
    dim3 grid(gx, t.order + 1);
    k_extents<<<grid, 256, 0, s>>>(t.coords, t.vals, t.order, (size_t)t.nnz, dev_out);
}

// -------------------------------------------------------------- linearise
// Both load every mode's coordinate first (unrolled to the largest order,
// the mode lists are uniform kernel arguments), so a thread waits for one
// memory round trip, not one per mode.
__device__ inline uint64_t lin_key(const uint32_t *coords, size_t nnz, size_t n, const int *modes, const uint64_t *R,
                                   int r) {
    uint32_t cv[kMaxOrder];
#pragma unroll
    for (int i = 0; i < kMaxOrder; i++) cv[i] = i < r ? coords[(size_t)modes[i] * nnz + n] : 0u;
    uint64_t k = 0;
#pragma unroll
    for (int i = 0; i < kMaxOrder; i++)
        if (i < r) k = k * R[i] + cv[i];
    return k;
}
__device__ inline uint32_t lin_ext(const uint32_t *coords, size_t nnz, size_t n, const int *xm, const uint32_t *xE,
                                   int nx) {
    uint32_t cv[kMaxOrder];
#pragma unroll
    for (int i = 0; i < kMaxOrder; i++) cv[i] = i < nx ? coords[(size_t)xm[i] * nnz + n] : 0u;
    uint32_t e = 0;
#pragma unroll
    for (int i = 0; i < kMaxOrder; i++)
        if (i < nx) e = e * xE[i] + cv[i];
    return e;
}

// B: comp = key << cbits | col  (coordinate.hpp:723-733 and 767-781)
// (the entry numbers, when the sort carries them, are made by its first pass)
__global__ void k_lin_B(const uint32_t *coords, size_t nnz, Geom g, uint64_t *comp) {
    size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (n >= nnz) return;
    uint64_t key = lin_key(coords, nnz, n, g.bm, g.R, g.r);
    uint32_t col = lin_ext(coords, nnz, n, g.xb, g.xbE, g.nxb);
    comp[n] = (key << g.cbits) | col;
}

// Sorted right entries: column, value and fiber-start flag. The value came
// with the sort, whose last pass wrote it straight into Bval (idx null), or,
// when the sort carried entry numbers, is gathered through them.
__global__ void k_gather_B(const uint64_t *comp, const uint32_t *idx, const double *vals, size_t n, int cbits,
                           uint32_t *Bcol, double *Bval, uint32_t *flag) {
    size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    uint64_t c = comp[p];
    Bcol[p] = (uint32_t)(c & ((1ull << cbits) - 1));
    if (idx) Bval[p] = vals[idx[p]];
    flag[p] = (p == 0 || (c >> cbits) != (comp[p - 1] >> cbits)) ? 1u : 0u;
}

__global__ void k_fibers(const uint64_t *comp, const uint32_t *flag, const uint32_t *fid, size_t n, int cbits,
                         uint32_t *fiber_ptr, uint64_t *fiber_key, uint32_t *Bfib) {
    size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t fl = flag[p], f = fid[p];
    if (fl) {
        fiber_ptr[f] = (uint32_t)p;
        if (fiber_key) fiber_key[f] = comp[p] >> cbits;  // (only the left join looks keys up)
    }
    Bfib[p] = f + fl - 1;
    if (p == n - 1) fiber_ptr[f + fl] = (uint32_t)n;
}

__global__ void k_keymap(const uint64_t *fiber_key, uint32_t nfib, uint32_t *keymap) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nfib) keymap[fiber_key[f]] = f;
}

__device__ inline uint32_t find_fiber(const uint64_t *fk, uint32_t nfib, uint64_t key) {
    uint32_t lo = 0, hi = nfib;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (fk[mid] < key) lo = mid + 1; else hi = mid;
    }
    return (lo < nfib && fk[lo] == key) ? lo : 0xFFFFFFFFu;
}

// A join: row, matching fiber (or none), global product count.
__global__ __launch_bounds__(256) void k_join_A(const uint32_t *coords, size_t nnz, Geom g, const uint32_t *keymap,
                                               const uint64_t *fiber_key, uint32_t nfib, const uint32_t *fiber_ptr,
                                               uint32_t *rowA, uint32_t *fidA, unsigned long long *products,
                                               unsigned int *nvalid) {
    __shared__ unsigned long long s_p[4];
    __shared__ unsigned int s_v[4];
    uint64_t prod = 0;
    uint32_t valid = 0;
    for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < nnz; n += (size_t)gridDim.x * blockDim.x) {
        uint64_t key = lin_key(coords, nnz, n, g.am, g.R, g.r);
        uint32_t f = keymap ? keymap[key] : find_fiber(fiber_key, nfib, key);
        rowA[n] = lin_ext(coords, nnz, n, g.xa, g.xaE, g.nxa);
        fidA[n] = f;
        if (f != 0xFFFFFFFFu) {
            prod += fiber_ptr[f + 1] - fiber_ptr[f];
            valid++;
        }
    }
    for (int d = 32; d > 0; d >>= 1) {
        prod += __shfl_xor(prod, d, 64);
        valid += __shfl_xor(valid, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_p[threadIdx.x >> 6] = prod;
        s_v[threadIdx.x >> 6] = valid;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        prod = s_p[0] + s_p[1] + s_p[2] + s_p[3];
        valid = s_v[0] + s_v[1] + s_v[2] + s_v[3];
        if (prod) atomicAdd(products, (unsigned long long)prod);
        if (valid) atomicAdd(nvalid, valid);
    }
}

// What the left build needs of a source entry (an A entry, or for a
// self-contraction a right position): its value, its fiber's range, whether
// the fiber's key is a dense key, and the fiber's column sketch. The
// composite kernels pack it into one 32-byte record while they read it in
// source order, so the gather after the sort makes a single random access per
// entry (uint4 halves: {value lo, value hi, fiber start, fiber end},
// {flags, sketch 0, sketch 1, 0}; flag bit 0: dense key).
__device__ __forceinline__ void put_lrec(uint4 *rec, size_t i, double v, uint32_t fp0, uint32_t fp1, uint32_t flags,
                                         uint2 sk) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    rec[2 * i] = make_uint4((uint32_t)b, (uint32_t)(b >> 32), fp0, fp1);
    rec[2 * i + 1] = make_uint4(flags, sk.x, sk.y, 0u);
}

// A composite: tile | class | fiber | row-in-tile; unmatched entries sort last.
__global__ void k_comp_A(const uint32_t *rowA, const uint32_t *fidA, const uint32_t *fiber_ptr, const double *vals,
                         const uint32_t *dk_hkey, const uint2 *fsk, size_t nnz, int logTL, int fbits,
                         uint32_t heavy_min, uint64_t *comp, uint4 *rec) {
    size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (n >= nnz) return;
    uint32_t f = fidA[n];
    uint64_t c = ~0ull;
    if (f != 0xFFFFFFFFu) {
        uint32_t row = rowA[n];
        uint64_t tile = row >> logTL;
        uint64_t li = row & ((1u << logTL) - 1);
        const uint32_t fp0 = fiber_ptr[f], fp1 = fiber_ptr[f + 1];
        uint64_t cls = (fp1 - fp0) >= heavy_min ? 0 : 1;
        c = (tile << (1 + fbits + logTL)) | (cls << (fbits + logTL)) | ((uint64_t)f << logTL) | li;
        const uint32_t dense = dk_hkey && dk_hkey[f] != 0xFFFFFFFFu ? 1u : 0u;
        put_lrec(rec, n, vals[n], fp0, fp1, dense, fsk ? fsk[f] : make_uint2(0u, 0u));
    }
    comp[n] = c;
}

// Left entries after the sort (ks: composite keys, vs: source entries), 4096
// per block in 16 coalesced rounds: the per-entry arrays, and the number of
// group starts (a new tile, class or fiber) and tile starts in the block
// (bsum: tiles << 32 | groups), which one small scan turns into each block's
// first group and tile. At a group start ex holds what its group descriptor
// and statistics need: {fiber length, flags, sketch}.
// Mirror mode: the source entry is the entry's own position in the right
// operand's (key, column) order, i.e. inside its fiber at the column equal to
// its row, so its fiber range starts there: the cells left of the diagonal
// are cut without a search.
// Dense keys: a light entry whose key the dense GEMM takes gets an empty
// fiber range (heavy groups of such keys get an empty descriptor later).
constexpr int LB_THREADS = 256;
constexpr int LB_ROUNDS = 16;
constexpr uint32_t LB_TILE = LB_THREADS * LB_ROUNDS;
constexpr int LB_BATCH = 4;  // rounds whose loads are issued together

__global__ __launch_bounds__(LB_THREADS) void k_left_entries(const uint64_t *ks, const uint32_t *vs, const uint4 *rec,
                                                             uint32_t m, int logTL, int fbits, bool mirror,
                                                             uint32_t *Ali, uint32_t *Arow, double *Aval,
                                                             uint32_t *Afib, uint2 *Afp, uint4 *ex,
                                                             unsigned long long *bsum) {
    __shared__ unsigned long long s_red[LB_THREADS / 64];
    const int tsh = 1 + fbits + logTL, csh = fbits + logTL;
    const uint64_t fmask = (1ull << fbits) - 1, lmask = (1ull << logTL) - 1;
    const uint32_t base = blockIdx.x * LB_TILE;
    uint32_t ng = 0, nt = 0;
    for (int r0 = 0; r0 < LB_ROUNDS; r0 += LB_BATCH) {
        uint64_t c[LB_BATCH], pv[LB_BATCH];
        uint32_t p[LB_BATCH];
        uint4 r1[LB_BATCH], r2[LB_BATCH];
#pragma unroll
        for (int j = 0; j < LB_BATCH; j++) {
            const uint32_t q = base + (r0 + j) * LB_THREADS + threadIdx.x;
            const uint32_t qc = q < m ? q : m - 1;  // clamped: loads stay in bounds, results unused
            c[j] = ks[qc];
            pv[j] = qc ? ks[qc - 1] : ~0ull;
            p[j] = vs[qc];
        }
#pragma unroll
        for (int j = 0; j < LB_BATCH; j++) {
            r1[j] = rec[2 * (size_t)p[j]];
            r2[j] = rec[2 * (size_t)p[j] + 1];
        }
#pragma unroll
        for (int j = 0; j < LB_BATCH; j++) {
            const uint32_t q = base + (r0 + j) * LB_THREADS + threadIdx.x;
            if (q >= m) continue;
            const uint64_t cc = c[j];
            const uint32_t li = (uint32_t)(cc & lmask);
            const uint64_t tile = cc >> tsh;
            const bool light = (cc >> csh) & 1;
            const bool gf = q == 0 || (cc >> logTL) != (pv[j] >> logTL);
            const bool tf = q == 0 || tile != (pv[j] >> tsh);
            Ali[q] = li;
            Arow[q] = (uint32_t)((tile << logTL) | li);
            Aval[q] = __longlong_as_double((long long)(((unsigned long long)r1[j].y << 32) | r1[j].x));
            Afib[q] = (uint32_t)((cc >> logTL) & fmask);
            const uint32_t fp0 = r1[j].z, fend = r1[j].w, flags = r2[j].x;
            const uint32_t x0 = mirror ? p[j] : fp0;
            Afp[q] = make_uint2(x0, ((flags & 1u) && light) ? x0 : fend);
            if (gf) ex[q] = make_uint4(fend - fp0, flags, r2[j].y, r2[j].z);
            ng += gf ? 1u : 0u;
            nt += tf ? 1u : 0u;
        }
    }
    unsigned long long v = ((unsigned long long)nt << 32) | ng;
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d, 64);
    if ((threadIdx.x & 63) == 0) s_red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < LB_THREADS / 64; w++) t += s_red[w];
        bsum[blockIdx.x] = t;
    }
}

// Groups and active tiles, over the same 4096-entry blocks: boff holds the
// exclusive prefix of bsum (tiles << 32 | groups before the block); inside
// the block each round numbers its starts with ballots. A group start writes
// the group (entry range start, fiber, tile, descriptor without its entry
// count, fiber length and flags, sketch); a tile start its id and first
// group; each tile's first light group is where the class turns light (light
// groups follow the heavy ones inside a tile), or the tile's group end when
// its last group is heavy.
// SINGLE (every group holds one entry: one-row tiles over distinct
// coordinates): the group's entry count (1) goes straight into its
// descriptor, and the per-tile statistics of k_tile_stats are summed here,
// runs of equal tiles reduced inside each wavefront, one set of atomics per
// run, so that pass is not needed, nor the per-group arrays only it reads
// (group_atile, gnf, gsk: left unwritten).
struct TileStatsOut {
    uint8_t *tile_dk;
    uint32_t *nnz, *keys, *sketch;
    unsigned long long *sumfib, *products, *maxgp;
};

template <bool SINGLE>
__global__ __launch_bounds__(LB_THREADS) void k_left_groups(const uint64_t *ks, const uint2 *Afp, const uint4 *ex,
                                                            const unsigned long long *boff, uint32_t m, int logTL,
                                                            int fbits, bool features, uint32_t *group_ptr,
                                                            uint32_t *group_fiber, uint32_t *group_atile,
                                                            uint4 *gdesc, uint2 *gnf, uint2 *gsk, uint32_t *atile_id,
                                                            uint32_t *atile_g0, uint32_t *atile_lg,
                                                            TileStatsOut so) {
    __shared__ uint32_t s_g[LB_THREADS / 64], s_t[LB_THREADS / 64];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    const int tsh = 1 + fbits + logTL, csh = fbits + logTL;
    const uint64_t fmask = (1ull << fbits) - 1;
    const unsigned long long bo = boff[blockIdx.x];
    uint32_t gb = (uint32_t)bo, tb = (uint32_t)(bo >> 32);  // starts before this round
    const uint32_t base = blockIdx.x * LB_TILE;
    for (int r = 0; r < LB_ROUNDS; r++) {
        const uint32_t q = base + r * LB_THREADS + threadIdx.x;
        const bool ok = q < m;
        uint64_t c = 0, prev = 0;
        bool gf = false, tf = false, tend = false;
        if (ok) {
            c = ks[q];
            prev = q ? ks[q - 1] : 0ull;
            gf = q == 0 || (c >> logTL) != (prev >> logTL);
            tf = q == 0 || (c >> tsh) != (prev >> tsh);
            tend = q + 1 == m || (ks[q + 1] >> tsh) != (c >> tsh);
        }
        const uint64_t bg = __ballot(gf), bt = __ballot(tf);
        if (lane == 0) {
            s_g[wave] = (uint32_t)__popcll(bg);
            s_t[wave] = (uint32_t)__popcll(bt);
        }
        __syncthreads();
        uint32_t wg = 0, wt = 0, rg = 0, rt = 0;
#pragma unroll
        for (int w = 0; w < LB_THREADS / 64; w++) {
            if (w < (int)wave) {
                wg += s_g[w];
                wt += s_t[w];
            }
            rg += s_g[w];
            rt += s_t[w];
        }
        __syncthreads();  // s_g / s_t are rewritten by the next round
        // group and tile of entry q: starts up to and including q, minus one
        const uint32_t g = gb + wg + (uint32_t)__popcll(bg & lt) + (gf ? 1u : 0u) - 1u;
        const uint32_t a = tb + wt + (uint32_t)__popcll(bt & lt) + (tf ? 1u : 0u) - 1u;
        uint4 x = make_uint4(0u, 0u, 0xFFFFFFFFu, 0xFFFFFFFFu);
        if (ok) {
            const bool light = (c >> csh) & 1;
            if (gf) {
                const uint2 fp = Afp[q];
                x = ex[q];
                group_ptr[g] = q;
                group_fiber[g] = (uint32_t)((c >> logTL) & fmask);
                gdesc[g] = make_uint4(fp.x, (x.y & 1u) ? 0u : fp.y - fp.x, q, SINGLE ? 1u : 0u);
                if (!SINGLE) {  // read by the statistics pass only
                    group_atile[g] = a;
                    gnf[g] = make_uint2(x.x, x.y);
                    if (features) gsk[g] = make_uint2(x.z, x.w);
                }
                if (light && (tf || !((prev >> csh) & 1))) atile_lg[a] = g;
                if (SINGLE && (x.y & 1u) && so.tile_dk) so.tile_dk[a] = 1;
            }
            if (tf) {
                atile_id[a] = (uint32_t)(c >> tsh);
                atile_g0[a] = g;
            }
            if (tend && !light) atile_lg[a] = g + 1;
            if (q + 1 == m) {
                group_ptr[g + 1] = m;
                atile_g0[a + 1] = g + 1;
            }
        }
        if (SINGLE) {
            // the statistics k_tile_stats takes from each group: entries m = 1,
            // products m * nf, distinct keys 1, fiber entries nf, sketch minima,
            // the largest group's products; lanes of one tile are contiguous
            const uint32_t ta = ok ? a : 0xFFFFFFFFu;
            uint32_t sm = ok ? 1u : 0u, h1 = x.z, h2 = x.w;
            uint64_t snf = ok ? x.x : 0u, mx = snf;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t oa = __shfl_down(ta, d, 64), om = __shfl_down(sm, d, 64);
                const uint64_t onf = __shfl_down(snf, d, 64), omx = __shfl_down(mx, d, 64);
                const uint32_t o1 = __shfl_down(h1, d, 64), o2 = __shfl_down(h2, d, 64);
                if (lane + d < 64 && oa == ta) {
                    sm += om;
                    snf += onf;
                    mx = omx > mx ? omx : mx;
                    h1 = o1 < h1 ? o1 : h1;
                    h2 = o2 < h2 ? o2 : h2;
                }
            }
            const uint32_t pa = __shfl_up(ta, 1, 64);
            if (ok && (lane == 0 || pa != ta)) {
                atomicAdd(&so.nnz[ta], sm);
                atomicAdd(&so.products[ta], (unsigned long long)snf);
                atomicMax(&so.maxgp[ta], (unsigned long long)mx);
                if (features) {
                    atomicAdd(&so.keys[ta], sm);
                    atomicAdd(&so.sumfib[ta], (unsigned long long)snf);
                    atomicMin(&so.sketch[2 * ta], h1);
                    atomicMin(&so.sketch[2 * ta + 1], h2);
                }
            }
        }
        gb += rg;
        tb += rt;
    }
}

// Heavy segments per active tile: a tile with fewer than seg_below heavy
// groups has its heavy groups cut every hseg fiber entries (its accumulate
// units spread long fibers over many wavefronts); other tiles get none.
__global__ void k_hseg_tile_count(const uint32_t *atile_g0, const uint32_t *atile_lg, const uint4 *gdesc,
                                  uint32_t natiles, uint32_t hseg, uint32_t seg_below, uint32_t *nseg) {
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a > natiles) return;
    uint32_t n = 0;
    if (a < natiles) {
        const uint32_t g0 = atile_g0[a], gl = atile_lg[a];
        if (gl - g0 < seg_below)
            for (uint32_t g = g0; g < gl; g++) n += (gdesc[g].y + hseg - 1) / hseg;
    }
    nseg[a] = n;  // nseg[natiles] = 0: the scan's total lands in atile_h0[natiles]
}

__global__ void k_hseg_tile_emit(const uint32_t *atile_g0, const uint32_t *atile_lg, const uint4 *gdesc,
                                 const uint32_t *atile_h0, uint32_t natiles, uint32_t hseg, uint32_t seg_below,
                                 uint4 *hdesc) {
    const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natiles) return;
    const uint32_t g0 = atile_g0[a], gl = atile_lg[a];
    if (gl - g0 >= seg_below) return;
    uint32_t o = atile_h0[a];
    for (uint32_t g = g0; g < gl; g++) {
        const uint4 d = gdesc[g];
        for (uint32_t s0 = 0; s0 < d.y; s0 += hseg) hdesc[o++] = make_uint4(d.x + s0, min(hseg, d.y - s0), d.z, d.w);
    }
}

// A tile is DIRECT-able when no two of its entries share a row (then no two
// products meet at one output). Checked for tiles of at most 64 rows holding
// at most one entry per row.
// (never a tile holding a dense key: its outputs meet the dense block)
__global__ void k_tile_direct(const uint32_t *atile_g0, const uint32_t *group_ptr, const uint32_t *Ali,
                              const uint32_t *st_nnz, const uint8_t *tile_dk, uint32_t natiles, uint32_t TL,
                              uint8_t *direct) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natiles) return;
    uint8_t ok = 0;
    if (TL <= DIRECT_MAX_TL && st_nnz[a] <= TL && !(tile_dk && tile_dk[a])) {
        const uint32_t e0 = group_ptr[atile_g0[a]], e1 = group_ptr[atile_g0[a + 1]];
        uint64_t seen = 0;
        ok = 1;
        for (uint32_t e = e0; e < e1; e++) {
            const uint64_t bit = 1ull << Ali[e];
            if (seen & bit) {
                ok = 0;
                break;
            }
            seen |= bit;
        }
    }
    direct[a] = ok;
}

// Mirror mode: outputs of each DIRECT tile. Its entries sit on distinct rows,
// so every product is a distinct cell; entry (row r, fiber f) keeps the fiber
// columns c >= r (its fiber range already starts there), each written twice
// except the diagonal one. Other tiles keep their products.
__global__ void k_direct_outputs(const uint32_t *atile_g0, const uint32_t *group_ptr, const uint32_t *Arow,
                                 const uint2 *Afp, const uint32_t *Bcol, const uint8_t *st_direct,
                                 const uint64_t *st_products, uint32_t natiles, uint64_t *st_outd) {
    uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= natiles) return;
    if (!st_direct[a]) {
        st_outd[a] = st_products[a];
        return;
    }
    const uint32_t e0 = group_ptr[atile_g0[a]], e1 = group_ptr[atile_g0[a + 1]];
    uint64_t n = 0;
    for (uint32_t e = e0; e < e1; e++) {
        const uint2 fp = Afp[e];  // starts at the entry's own column (its row)
        if (fp.x < fp.y) n += 2ull * (fp.y - fp.x) - (Bcol[fp.x] == Arow[e] ? 1u : 0u);
    }
    st_outd[a] = n;
}

__device__ inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// Per fiber: two min-hashes over up to 4 evenly spaced columns (fibers are
// never empty). A group's sketch is its fiber's, so it is computed once per
// fiber rather than once per (tile, fiber) group.
__global__ void k_fiber_sketch(const uint32_t *fiber_ptr, const uint32_t *Bcol, uint32_t nfib, uint2 *fsk) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfib) return;
    const uint32_t f0 = fiber_ptr[f], f1 = fiber_ptr[f + 1];
    const uint32_t step = (f1 - f0 + 3) / 4;
    uint32_t h1 = 0xFFFFFFFFu, h2 = 0xFFFFFFFFu;
    for (uint32_t p = f0; p < f1; p += step) {
        const uint32_t c = Bcol[p];
        const uint32_t x1 = hash32(c * 2u + 1u), x2 = hash32(c ^ 0x5bd1e995u);
        h1 = x1 < h1 ? x1 : h1;
        h2 = x2 < h2 ? x2 : h2;
    }
    fsk[f] = make_uint2(h1, h2);
}

// Per group: its entry count (the last field of its descriptor) and,
// aggregated per active tile (block-wide when the block's groups share a
// tile, else per run of equal tiles in a wavefront), the tile statistics:
//  nnz      A entries of the tile          (num_nnz_in_tile, index_tensors.hpp:562)
//  keys     distinct contraction keys      (num_active_columns, index_tensors.hpp:604)
//  sumfib   summed right-fiber lengths
//  products sum of m * |fiber| = an upper bound on the tile's output count
//  sketch   two min-hashes over sampled fiber columns (distinct-column estimate)
// gnf: per group {fiber length, flags} (flag bit 0: a dense key), gsk its
// fiber's sketch (planner features only). A dense key's group leaves the tile
// traversal (its descriptor and its light entries' ranges are already empty)
// and its tile is flagged in tile_dk (its outputs meet the dense block); the
// statistics keep it. Every read is in group order.
__global__ __launch_bounds__(256) void k_tile_stats(const uint32_t *group_ptr, const uint32_t *group_atile,
                                                   const uint2 *gnf, const uint2 *gsk, uint8_t *tile_dk,
                                                   uint32_t ngroups, uint4 *gdesc, uint32_t *st_nnz,
                                                   uint32_t *st_keys, unsigned long long *st_sumfib,
                                                   unsigned long long *st_products, uint32_t *st_sketch,
                                                   unsigned long long *st_maxgp) {
    __shared__ uint32_t s_a0;
    __shared__ unsigned long long s_red[4][4];
    __shared__ uint32_t s_h[4][2];
    __shared__ unsigned long long s_mx[4];
    const bool fsk = gsk != nullptr;
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = g < ngroups;
    uint32_t a = ok ? group_atile[g] : 0xFFFFFFFFu;
    uint64_t m = 0, nf = 0;
    uint32_t h1 = 0xFFFFFFFFu, h2 = 0xFFFFFFFFu;
    if (ok) {
        const uint32_t e0 = group_ptr[g];
        m = group_ptr[g + 1] - e0;
        const uint2 x = gnf[g];
        nf = x.x;
        reinterpret_cast<uint32_t *>(gdesc)[4 * (size_t)g + 3] = (uint32_t)m;
        if ((x.y & 1u) && tile_dk) tile_dk[a] = 1;
        if (fsk) {
            const uint2 hh = gsk[g];
            h1 = hh.x;
            h2 = hh.y;
        }
    }
    if (threadIdx.x == 0) s_a0 = a;
    __syncthreads();
    const uint32_t a0 = s_a0;
    // whole block in one tile (large tiles): one set of atomics per block
    if (__syncthreads_and(!ok || a == a0) && a0 != 0xFFFFFFFFu) {
        uint64_t sm = m, snf = nf, sp = m * nf, sk = ok ? 1 : 0, mx = m * nf;
        for (int d = 32; d > 0; d >>= 1) {
            uint64_t om = __shfl_xor(mx, d, 64);
            mx = om > mx ? om : mx;
            sm += __shfl_xor(sm, d, 64);
            snf += __shfl_xor(snf, d, 64);
            sp += __shfl_xor(sp, d, 64);
            sk += __shfl_xor(sk, d, 64);
            uint32_t o1 = __shfl_xor(h1, d, 64), o2 = __shfl_xor(h2, d, 64);
            h1 = o1 < h1 ? o1 : h1;
            h2 = o2 < h2 ? o2 : h2;
        }
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0) {
            s_red[w][0] = sm;
            s_red[w][1] = snf;
            s_red[w][2] = sp;
            s_red[w][3] = sk;
            s_h[w][0] = h1;
            s_h[w][1] = h2;
            s_mx[w] = mx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int v = 1; v < 4; v++) {
                sm += s_red[v][0];
                snf += s_red[v][1];
                sp += s_red[v][2];
                sk += s_red[v][3];
                h1 = s_h[v][0] < h1 ? s_h[v][0] : h1;
                h2 = s_h[v][1] < h2 ? s_h[v][1] : h2;
                mx = s_mx[v] > mx ? s_mx[v] : mx;
            }
            atomicMax(&st_maxgp[a0], (unsigned long long)mx);
            atomicAdd(&st_nnz[a0], (uint32_t)sm);
            atomicAdd(&st_products[a0], (unsigned long long)sp);
            if (fsk) {  // planner features
                atomicAdd(&st_keys[a0], (uint32_t)sk);
                atomicAdd(&st_sumfib[a0], (unsigned long long)snf);
                atomicMin(&st_sketch[2 * a0], h1);
                atomicMin(&st_sketch[2 * a0 + 1], h2);
            }
        }
        return;
    }
    // Small tiles: groups are in tile order, so the lanes of one tile are
    // contiguous. A segmented reduction towards each run's first lane (the
    // partner d lanes on belongs to the same run exactly when its tile is
    // equal) leaves one set of atomics per tile per wavefront.
    const uint32_t lane = threadIdx.x & 63;
    uint32_t sm = (uint32_t)m, sk = ok ? 1u : 0u;
    uint64_t snf = nf, sp = m * nf, mx = m * nf;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t oa = __shfl_down(a, d, 64);
        const uint32_t om = __shfl_down(sm, d, 64), ok2 = __shfl_down(sk, d, 64);
        const uint64_t onf = __shfl_down(snf, d, 64), op = __shfl_down(sp, d, 64), omx = __shfl_down(mx, d, 64);
        const uint32_t o1 = __shfl_down(h1, d, 64), o2 = __shfl_down(h2, d, 64);
        if (lane + d < 64 && oa == a) {
            sm += om;
            sk += ok2;
            snf += onf;
            sp += op;
            mx = omx > mx ? omx : mx;
            h1 = o1 < h1 ? o1 : h1;
            h2 = o2 < h2 ? o2 : h2;
        }
    }
    const uint32_t pa = __shfl_up(a, 1, 64);
    if (ok && (lane == 0 || pa != a)) {
        atomicAdd(&st_nnz[a], sm);
        atomicAdd(&st_products[a], (unsigned long long)sp);
        atomicMax(&st_maxgp[a], (unsigned long long)mx);
        if (fsk) {  // planner features
            atomicAdd(&st_keys[a], sk);
            atomicAdd(&st_sumfib[a], (unsigned long long)snf);
            atomicMin(&st_sketch[2 * a], h1);
            atomicMin(&st_sketch[2 * a + 1], h2);
        }
    }
}

// ---------------------------------------------------------------- host side

static unsigned grid_for(size_t n, unsigned bs = 256) { return (unsigned)div_up(n ? n : 1, bs); }

// Self-contraction (same tensor, same mode lists): the left join is read off
// the right operand's sort. Sorted position p holds original entry vs[p], its
// column (= the left row) and its fiber.
__global__ __launch_bounds__(256) void k_self_join(const uint32_t *vs, const uint32_t *Bfib, const uint32_t *Bcol,
                                                  const uint32_t *fiber_ptr, size_t n, uint32_t *rowA, uint32_t *fidA,
                                                  uint32_t *vsB, unsigned long long *products) {
    __shared__ unsigned long long s_p[4];
    uint64_t prod = 0;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const uint32_t f = Bfib[p];
        if (vs) {  // entry numbers carried by the sort
            const uint32_t o = vs[p];
            if (rowA) {  // per original entry (random writes): only when asked for
                rowA[o] = Bcol[p];
                fidA[o] = f;
            }
            vsB[p] = o;
        }
        prod += fiber_ptr[f + 1] - fiber_ptr[f];
    }
    for (int d = 32; d > 0; d >>= 1) prod += __shfl_xor(prod, d, 64);
    if ((threadIdx.x & 63) == 0) s_p[threadIdx.x >> 6] = prod;
    __syncthreads();
    if (threadIdx.x == 0) {
        prod = s_p[0] + s_p[1] + s_p[2] + s_p[3];
        if (prod) atomicAdd(products, (unsigned long long)prod);
    }
}

// A self-contraction's product count when no per-entry join is wanted: the
// sum over fibers of (fiber length)^2, one thread per fiber (nfib is read on
// the device, before the host has it).
__global__ __launch_bounds__(256) void k_fiber_products(const uint32_t *fiber_ptr, const uint32_t *nfib,
                                                        unsigned long long *products) {
    __shared__ unsigned long long s_p[4];
    const uint32_t nf = *nfib;
    uint64_t prod = 0;
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
        const uint64_t l = fiber_ptr[f + 1] - fiber_ptr[f];
        prod += l * l;
    }
    for (int d = 32; d > 0; d >>= 1) prod += __shfl_xor(prod, d, 64);
    if ((threadIdx.x & 63) == 0) s_p[threadIdx.x >> 6] = prod;
    __syncthreads();
    if (threadIdx.x == 0) {
        prod = s_p[0] + s_p[1] + s_p[2] + s_p[3];
        if (prod) atomicAdd(products, (unsigned long long)prod);
    }
}

void build_right(DevPool &pool, const Geom &g, const Operand &B, Index &ix, std::vector<std::shared_ptr<void>> &hold,
                 hipStream_t s, SelfJoin *self) {
    size_t n = (size_t)B.nnz;
    ix.nB = (uint32_t)n;
    auto keep = [&](size_t bytes) -> void * {
        void *p = pool.get(bytes + 16);
        DevPool *pl = &pool;
        hold.emplace_back(p, [pl](void *q) { pl->put(q); });
        return p;
    };
    // The sort carries the values themselves (no gather through a
    // permutation afterwards) unless the entry numbers are wanted: plans keep
    // the permutation, sampled self-contractions join per original entry.
    const bool by_value = !ix.keep_perm && !(self && self->rowA);
    Tmp<uint64_t> k1(pool, n), k2(pool, n);
    Tmp<uint32_t> v1(pool, by_value ? 0 : n), v2(pool, by_value ? 0 : n);
    ix.Bcol = (uint32_t *)keep(n * 4);
    ix.Bval = (double *)keep(n * 8);
    // by value: the passes alternate between Bval and one spare array, in the
    // order that makes the last pass write Bval
    Tmp<uint64_t> w2(pool, by_value ? n : 0);
    uint64_t *bv = reinterpret_cast<uint64_t *>(ix.Bval);
    const int nbits = g.kbits + g.cbits;
    const bool odd = n > 1 && ((nbits + 7) / 8) % 2 == 1;
    k_lin_B<<<grid_for(n), 256, 0, s>>>(B.coords, n, g, k1);
    uint64_t *ks, *ws = nullptr;
    uint32_t *vs = nullptr;
    if (by_value) {
        radix_sort_keys_values(pool, k1, reinterpret_cast<const uint64_t *>(B.vals), odd ? w2.get() : bv, k2,
                               odd ? bv : w2.get(), n, nbits, s, &ks, &ws);
        if (ws != bv) throw LibError{ERR_INVALID, "build_right: sorted values not in place"};
    } else {
        radix_sort_pairs(pool, k1, v1, k2, v2, n, nbits, s, &ks, &vs, 0, true);
    }
    Tmp<uint32_t> flag(pool, n), fid(pool, n);
    k_gather_B<<<grid_for(n), 256, 0, s>>>(ks, vs, B.vals, n, g.cbits, ix.Bcol, ix.Bval, flag);
    if (ix.keep_perm) {
        ix.permB = (uint32_t *)keep(n * 4);
        if (n) FCC_HIP(hipMemcpyAsync(ix.permB, vs, n * 4, hipMemcpyDeviceToDevice, s));
    }
    Tmp<uint32_t> total(pool, 1);
    scan_u32(pool, flag, fid, n, total, s);
    ix.fiber_ptr = (uint32_t *)keep((n + 1) * 4);
    // a self-contraction reads its left side off this sort: no key lookups
    ix.fiber_key = self ? nullptr : (uint64_t *)keep(n * 8);
    ix.Bfib = (uint32_t *)keep(n * 4);
    k_fibers<<<grid_for(n), 256, 0, s>>>(ks, flag, fid, n, g.cbits, ix.fiber_ptr, ix.fiber_key, ix.Bfib);
    if (self) {
        Tmp<unsigned long long> prod(pool, 1);
        fill_u64(prod, 1, 0, s);
        self->bfib = ix.Bfib;
        if (vs)
            k_self_join<<<std::min(grid_for(n), 2048u), 256, 0, s>>>(vs, ix.Bfib, ix.Bcol, ix.fiber_ptr, n,
                                                                      self->rowA, self->fidA, self->vsB, prod);
        else
            k_fiber_products<<<std::min(grid_for(n), 1024u), 256, 0, s>>>(ix.fiber_ptr, total, prod);
        pool.read(&self->products, prod.get(), 8, s);
    }
    pool.read(&ix.nfib, total.get(), 4, s);
    pool.sync(s);
}

// key -> fiber hash map for key spaces too large for a direct table
__device__ __forceinline__ uint64_t hmix(uint64_t z) {
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    return z ^ (z >> 33);
}

__global__ void k_hmap_build(const uint64_t *fiber_key, uint32_t nfib, unsigned long long *hk, uint32_t *hv,
                             uint64_t mask) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfib) return;
    unsigned long long key = fiber_key[f];
    uint64_t h = hmix(key) & mask;
    for (;;) {  // keys are unique and the table is at most half full: terminates
        unsigned long long old = atomicCAS(&hk[h], ~0ull, key);
        if (old == ~0ull) {
            hv[h] = f;
            return;
        }
        h = (h + 1) & mask;
    }
}

__device__ __forceinline__ uint32_t hmap_find(const unsigned long long *hk, const uint32_t *hv, uint64_t mask,
                                              uint64_t key) {
    uint64_t h = hmix(key) & mask;
    for (;;) {
        unsigned long long k = hk[h];
        if (k == key) return hv[h];
        if (k == ~0ull) return 0xFFFFFFFFu;
        h = (h + 1) & mask;
    }
}

__global__ __launch_bounds__(256) void k_join_A_hash(const uint32_t *coords, size_t nnz, Geom g,
                                                    const unsigned long long *hk, const uint32_t *hv, uint64_t mask,
                                                    const uint32_t *fiber_ptr, uint32_t *rowA, uint32_t *fidA,
                                                    unsigned long long *products, unsigned int *nvalid) {
    __shared__ unsigned long long s_p[4];
    __shared__ unsigned int s_v[4];
    uint64_t prod = 0;
    uint32_t valid = 0;
    for (size_t n = blockIdx.x * (size_t)blockDim.x + threadIdx.x; n < nnz; n += (size_t)gridDim.x * blockDim.x) {
        uint64_t key = lin_key(coords, nnz, n, g.am, g.R, g.r);
        uint32_t f = hmap_find(hk, hv, mask, key);
        rowA[n] = lin_ext(coords, nnz, n, g.xa, g.xaE, g.nxa);
        fidA[n] = f;
        if (f != 0xFFFFFFFFu) {
            prod += fiber_ptr[f + 1] - fiber_ptr[f];
            valid++;
        }
    }
    for (int d = 32; d > 0; d >>= 1) {
        prod += __shfl_xor(prod, d, 64);
        valid += __shfl_xor(valid, d, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_p[threadIdx.x >> 6] = prod;
        s_v[threadIdx.x >> 6] = valid;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        prod = s_p[0] + s_p[1] + s_p[2] + s_p[3];
        valid = s_v[0] + s_v[1] + s_v[2] + s_v[3];
        if (prod) atomicAdd(products, (unsigned long long)prod);
        if (valid) atomicAdd(nvalid, valid);
    }
}

void join_left(DevPool &pool, const Geom &g, const Operand &A, const Index &ix, uint32_t *rowA, uint32_t *fidA,
               uint64_t *products, uint32_t *nvalid, hipStream_t s) {
    size_t n = (size_t)A.nnz;
    Tmp<uint32_t> keymap;
    if (g.K <= (1ull << 26) && ix.nfib) {
        keymap = Tmp<uint32_t>(pool, g.K);
        fill_u32(keymap, g.K, 0xFFFFFFFFu, s);
        if (ix.nfib) k_keymap<<<grid_for(ix.nfib), 256, 0, s>>>(ix.fiber_key, ix.nfib, keymap);
    }
    Tmp<unsigned long long> cnt(pool, 2);
    fill_u64(cnt, 2, 0, s);
    if (g.K <= (1ull << 26) || ix.nfib == 0) {
        k_join_A<<<std::min(grid_for(n), 2048u), 256, 0, s>>>(A.coords, n, g, keymap.get(), ix.fiber_key, ix.nfib,
                                                               ix.fiber_ptr, rowA, fidA, cnt.get(),
                                                               (unsigned int *)(cnt.get() + 1));
    } else {
        uint64_t cap = pow2_ceil((uint64_t)ix.nfib * 2);
        Tmp<unsigned long long> hk(pool, cap);
        Tmp<uint32_t> hv(pool, cap);
        fill_u64(hk, cap, ~0ull, s);
        k_hmap_build<<<grid_for(ix.nfib), 256, 0, s>>>(ix.fiber_key, ix.nfib, hk, hv, cap - 1);
        k_join_A_hash<<<std::min(grid_for(n), 2048u), 256, 0, s>>>(A.coords, n, g, hk, hv, cap - 1, ix.fiber_ptr,
                                                                    rowA, fidA, cnt.get(),
                                                                    (unsigned int *)(cnt.get() + 1));
    }
    unsigned long long h[2];
    pool.read(h, cnt.get(), 16, s);
    pool.sync(s);
    *products = h[0];
    *nvalid = (uint32_t)(h[1] & 0xFFFFFFFFu);
}

// Left tiles from the sorted composite keys (tile | class | fiber | row-in-tile)
// and the original entry index of each sorted position.
void build_hseg(DevPool &pool, Index &ix, std::vector<std::shared_ptr<void>> &hold, hipStream_t s) {
    auto keep = [&](size_t bytes) -> void * {
        void *p = pool.get(bytes + 16);
        DevPool *pl = &pool;
        hold.emplace_back(p, [pl](void *q) { pl->put(q); });
        return p;
    };
    const size_t T = ix.natiles;
    Tmp<uint32_t> nseg(pool, T + 1), nh(pool, 1);
    ix.atile_h0 = (uint32_t *)keep((T + 1) * 4);
    k_hseg_tile_count<<<grid_for(T + 1), 256, 0, s>>>(ix.atile_g0, ix.atile_lg, ix.gdesc, (uint32_t)T, ix.hseg,
                                                      ix.seg_below, nseg);
    scan_u32(pool, nseg, ix.atile_h0, T + 1, nh, s);
    pool.read(&ix.nhseg, nh.get(), 4, s);
    pool.sync(s);
    ix.hdesc = (uint4 *)keep((size_t)ix.nhseg * 16);
    if (T && ix.nhseg)
        k_hseg_tile_emit<<<grid_for(T), 256, 0, s>>>(ix.atile_g0, ix.atile_lg, ix.gdesc, ix.atile_h0, (uint32_t)T,
                                                     ix.hseg, ix.seg_below, ix.hdesc);
}

// After the left sort: ks the sorted composite keys, vs their source entries
// (A entries, or for a self-contraction right positions), rec the source
// records (put_lrec); perm_map (when not null) maps vs to A's entry order for
// plans.
static void left_from_sorted(DevPool &pool, const uint4 *rec, const uint64_t *ks, const uint32_t *vs,
                             const uint32_t *perm_map, uint32_t nvalid, int logTL, int fbits, Index &ix,
                             std::vector<std::shared_ptr<void>> &hold, hipStream_t s) {
    auto keep = [&](size_t bytes) -> void * {
        void *p = pool.get(bytes + 16);
        DevPool *pl = &pool;
        hold.emplace_back(p, [pl](void *q) { pl->put(q); });
        return p;
    };
    size_t m = nvalid;
    ix.Ali = (uint32_t *)keep(m * 4);
    ix.Arow = (uint32_t *)keep(m * 4);
    ix.Aval = (double *)keep(m * 8);
    ix.Afib = (uint32_t *)keep(m * 4);
    ix.Afp = (uint2 *)keep(m * 8);
    const size_t nb = div_up(m, LB_TILE);
    Tmp<uint4> ex(pool, m);
    Tmp<unsigned long long> bsum(pool, nb), boff(pool, nb), tot(pool, 1);
    if (m)
        k_left_entries<<<(unsigned)nb, LB_THREADS, 0, s>>>(ks, vs, rec, (uint32_t)m, logTL, fbits, ix.mirror, ix.Ali,
                                                            ix.Arow, ix.Aval, ix.Afib, ix.Afp, ex, bsum);
    if (ix.keep_perm) {
        ix.permA = (uint32_t *)keep(m * 4);
        if (m && perm_map) k_map_u32<<<grid_for(m), 256, 0, s>>>(perm_map, vs, m, ix.permA);
        else if (m) FCC_HIP(hipMemcpyAsync(ix.permA, vs, m * 4, hipMemcpyDeviceToDevice, s));
    }
    scan_u64(pool, (const uint64_t *)bsum.get(), (uint64_t *)boff.get(), nb, (uint64_t *)tot.get(), s);
    uint64_t h = 0;
    pool.read(&h, tot.get(), 8, s);
    pool.sync(s);
    ix.ngroups = (uint32_t)h;
    ix.natiles = (uint32_t)(h >> 32);
    ix.group_ptr = (uint32_t *)keep((size_t)(ix.ngroups + 1) * 4);
    ix.group_fiber = (uint32_t *)keep((size_t)ix.ngroups * 4);
    ix.group_atile = (uint32_t *)keep((size_t)ix.ngroups * 4);
    ix.gdesc = (uint4 *)keep((size_t)ix.ngroups * 16);
    ix.atile_id = (uint32_t *)keep((size_t)ix.natiles * 4);
    ix.atile_g0 = (uint32_t *)keep((size_t)(ix.natiles + 1) * 4);
    ix.atile_lg = (uint32_t *)keep((size_t)ix.natiles * 4);
    // entry counts + statistics: one pass over the groups, or, when every
    // group holds a single entry, summed by the group pass itself
    // (FCC_SINGLE_GROUPS=0 keeps the separate statistics pass: tests compare the two)
    const char *sg = getenv("FCC_SINGLE_GROUPS");
    const bool single = m && ix.ngroups == m && !(sg && atoi(sg) == 0);
    Tmp<uint2> gnf(pool, single ? 0 : ix.ngroups), gsk(pool, (ix.features && !single) ? ix.ngroups : 0);
    size_t T = ix.natiles;
    if (ix.dk_hkey) {
        ix.tile_dk = (uint8_t *)keep(T + 4);
        fill_u32(reinterpret_cast<uint32_t *>(ix.tile_dk), (T + 4) / 4, 0, s);
    }
    ix.st_nnz = (uint32_t *)keep(T * 4);
    ix.st_keys = (uint32_t *)keep(T * 4);
    ix.st_sumfib = (uint64_t *)keep(T * 8);
    ix.st_products = (uint64_t *)keep(T * 8);
    ix.st_sketch = (uint32_t *)keep(T * 8);
    ix.st_maxgp = (uint64_t *)keep(T * 8);
    if (T) {
        FillSet f;
        f.add(ix.st_nnz, T, 0, 4);
        f.add(ix.st_keys, T, 0, 4);
        f.add(ix.st_sumfib, T, 0, 8);
        f.add(ix.st_products, T, 0, 8);
        f.add(ix.st_sketch, 2 * T, 0xFFFFFFFFull, 4);
        f.add(ix.st_maxgp, T, 0, 8);
        fill_many(f, s);
    }
    const TileStatsOut so{ix.tile_dk,
                          ix.st_nnz,
                          ix.st_keys,
                          ix.st_sketch,
                          (unsigned long long *)ix.st_sumfib,
                          (unsigned long long *)ix.st_products,
                          (unsigned long long *)ix.st_maxgp};
    if (m) {
        if (single)
            k_left_groups<true><<<(unsigned)nb, LB_THREADS, 0, s>>>(
                ks, ix.Afp, ex, boff, (uint32_t)m, logTL, fbits, ix.features, ix.group_ptr, ix.group_fiber,
                ix.group_atile, ix.gdesc, gnf, gsk, ix.atile_id, ix.atile_g0, ix.atile_lg, so);
        else
            k_left_groups<false><<<(unsigned)nb, LB_THREADS, 0, s>>>(
                ks, ix.Afp, ex, boff, (uint32_t)m, logTL, fbits, ix.features, ix.group_ptr, ix.group_fiber,
                ix.group_atile, ix.gdesc, gnf, gsk, ix.atile_id, ix.atile_g0, ix.atile_lg, so);
    } else {
        // no valid left entry: no groups, no tiles; the one-entry offset
        // arrays still hold their end (read by the segment table below)
        fill_u32(ix.atile_g0, 1, 0, s);
        fill_u32(ix.group_ptr, 1, 0, s);
    }
    // the planner features (key counts, fiber sums, column sketches) only
    // when a learned planner or statistics collection reads them
    if (T && !single)
        k_tile_stats<<<grid_for(ix.ngroups), 256, 0, s>>>(ix.group_ptr, ix.group_atile, gnf,
                                                          ix.features ? gsk.get() : nullptr, ix.tile_dk, ix.ngroups,
                                                          ix.gdesc, ix.st_nnz, ix.st_keys,
                                                          (unsigned long long *)ix.st_sumfib,
                                                          (unsigned long long *)ix.st_products, ix.st_sketch,
                                                          (unsigned long long *)ix.st_maxgp);
    build_hseg(pool, ix, hold, s);
    ix.st_direct = (uint8_t *)keep(T + 1);
    if (T)
        k_tile_direct<<<grid_for(T), 256, 0, s>>>(ix.atile_g0, ix.group_ptr, ix.Ali, ix.st_nnz, ix.tile_dk,
                                                   (uint32_t)T, ix.TL, ix.st_direct);
    // DIRECT tiles' outputs: their products, or in mirror mode the products
    // at or above the diagonal counted twice off it
    ix.st_outd = ix.st_products;
    if (ix.mirror) {
        ix.st_outd = (uint64_t *)keep(T * 8 + 8);
        if (T)
            k_direct_outputs<<<grid_for(T), 256, 0, s>>>(ix.atile_g0, ix.group_ptr, ix.Arow, ix.Afp, ix.Bcol,
                                                          ix.st_direct, ix.st_products, (uint32_t)T, ix.st_outd);
    }
    // exclusive prefix of st_outd (the output count of a run of DIRECT tiles)
    ix.st_pprefix = (uint64_t *)keep((T + 1) * 8);
    if (T) scan_u64(pool, ix.st_outd, ix.st_pprefix, T, ix.st_pprefix + T, s);
    else fill_u64((unsigned long long *)ix.st_pprefix, 1, 0, s);
}

// per-fiber column sketches for the planner features (null when not wanted)
static Tmp<uint2> fiber_sketches(DevPool &pool, const Index &ix, hipStream_t s) {
    Tmp<uint2> fsk(pool, ix.features ? ix.nfib : 0);
    if (ix.features && ix.nfib)
        k_fiber_sketch<<<grid_for(ix.nfib), 256, 0, s>>>(ix.fiber_ptr, ix.Bcol, ix.nfib, fsk);
    return fsk;
}

void build_left(DevPool &pool, const Geom &g, const Operand &A, const uint32_t *rowA, const uint32_t *fidA,
                uint32_t nvalid, int logTL, Index &ix, std::vector<std::shared_ptr<void>> &hold, hipStream_t s,
                const std::function<void()> *after_sort) {
    size_t n = (size_t)A.nnz;
    ix.logTL = logTL;
    ix.TL = 1u << logTL;
    ix.nAv = nvalid;
    int fbits = std::max(1, ceil_log2((uint64_t)ix.nfib + 1));
    int tbits = std::max(0, g.rbits - logTL);
    int nbits = tbits + 1 + fbits + logTL + 1;
    if (nbits > 64) throw LibError{ERR_INVALID, "left composite key exceeds 64 bits"};
    Tmp<uint64_t> k1(pool, n), k2(pool, n);
    Tmp<uint32_t> v1(pool, n), v2(pool, n);
    Tmp<uint4> rec(pool, 2 * n);
    {
        Tmp<uint2> fsk = fiber_sketches(pool, ix, s);
        k_comp_A<<<grid_for(n), 256, 0, s>>>(rowA, fidA, ix.fiber_ptr, A.vals, ix.dk_hkey,
                                             ix.features ? fsk.get() : nullptr, n, logTL, fbits, ix.heavy_min, k1,
                                             rec);
    }
    uint64_t *ks;
    uint32_t *vs;
    radix_sort_pairs(pool, k1, v1, k2, v2, n, std::min(nbits, 64), s, &ks, &vs, 0, true);
    if (after_sort) (*after_sort)();
    left_from_sorted(pool, rec, ks, vs, nullptr, nvalid, logTL, fbits, ix, hold, s);
}

// Self-contraction: the right operand is already sorted by (fiber, column)
// and column == left row, so a stable sort on (tile, class) alone yields the
// left order (tile, class, fiber, row-in-tile). The whole composite key is
// written here in the right operand's order (coalesced reads), and the sort
// touches only the (tile, class) bits.
__global__ void k_self_composite(const uint32_t *Bcol, const uint32_t *Bfib, const uint32_t *fiber_ptr, size_t n,
                                 int logTL, int fbits, uint32_t heavy_min, uint64_t *ks) {
    size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t f = Bfib[p], c = Bcol[p];
    const uint64_t cls = (fiber_ptr[f + 1] - fiber_ptr[f]) >= heavy_min ? 0 : 1;
    const uint64_t tile = c >> logTL, li = c & ((1u << logTL) - 1);
    ks[p] = (tile << (1 + fbits + logTL)) | (cls << (fbits + logTL)) | ((uint64_t)f << logTL) | li;
}

// The entries' source records, in the right operand's order: the left-tile
// pass reads them by position (the sort's payload). Written after the sort
// when a long block is forked there, so that it starts without waiting for
// them (order-5 self-contraction, 26M entries: step 16.2 -> 15.9 ms); written
// before the sort otherwise (beside the 0.54 ms block of an order-4 one they
// lengthened it to 0.66 ms).
__global__ void k_self_records(const uint32_t *Bfib, const uint32_t *fiber_ptr, const double *Bval,
                               const uint32_t *dk_hkey, const uint2 *fsk, size_t n, uint4 *rec) {
    size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t f = Bfib[p];
    const uint32_t dense = dk_hkey && dk_hkey[f] != 0xFFFFFFFFu ? 1u : 0u;
    put_lrec(rec, p, Bval[p], fiber_ptr[f], fiber_ptr[f + 1], dense, fsk ? fsk[f] : make_uint2(0u, 0u));
}

void build_left_self(DevPool &pool, const Geom &g, const Operand &A, const SelfJoin &sj, int logTL, Index &ix,
                     std::vector<std::shared_ptr<void>> &hold, hipStream_t s, const std::function<void()> *after_sort,
                     bool records_late) {
    size_t n = (size_t)A.nnz;
    ix.logTL = logTL;
    ix.TL = 1u << logTL;
    ix.nAv = (uint32_t)n;
    int fbits = std::max(1, ceil_log2((uint64_t)ix.nfib + 1));
    int tbits = std::max(0, g.rbits - logTL);
    if (tbits + 1 + fbits + logTL + 1 > 64) throw LibError{ERR_INVALID, "left composite key exceeds 64 bits"};
    Tmp<uint64_t> k1(pool, n), k2(pool, n);
    Tmp<uint32_t> v1(pool, n), v2(pool, n);
    Tmp<uint4> rec(pool, 2 * n);
    auto records = [&]() {
        Tmp<uint2> fsk = fiber_sketches(pool, ix, s);
        k_self_records<<<grid_for(n), 256, 0, s>>>(sj.bfib, ix.fiber_ptr, ix.Bval, ix.dk_hkey,
                                                   ix.features ? fsk.get() : nullptr, n, rec);
    };
    if (!records_late) records();
    k_self_composite<<<grid_for(n), 256, 0, s>>>(ix.Bcol, sj.bfib, ix.fiber_ptr, n, logTL, fbits, ix.heavy_min, k1);
    uint64_t *ks;
    uint32_t *vs;
    radix_sort_pairs(pool, k1, v1, k2, v2, n, tbits + 1, s, &ks, &vs, fbits + logTL, true);
    if (after_sort) (*after_sort)();
    if (records_late) records();
    left_from_sorted(pool, rec, ks, vs, sj.vsB, (uint32_t)n, logTL, fbits, ix, hold, s);
}

}  // namespace fcc
