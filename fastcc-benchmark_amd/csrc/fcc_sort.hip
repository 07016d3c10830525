// Device-wide exclusive scan and stable LSD radix sort of (u64 key, u32 value)
// pairs, written for CDNA (64-lane wavefronts) -- plus the caching allocator.
//
// These replace the reference's hash-map tile index build
// (index_tensors.hpp:506-547, ankerl::unordered_dense maps): grouping by
// sorting the linearised composite keys instead of hashing them.
#include "fcc_internal.h"

#include <cstring>

namespace fcc {

// ------------------------------------------------------------ caching pool

DevPool::~DevPool() {
    release_all();
    if (pin_) (void)hipHostFree(pin_);
}

void DevPool::init_pinned() {
    if (!pin_ && hipHostMalloc((void **)&pin_, kPinBytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        pin_ = nullptr;
    }
}

void DevPool::read(void *dst, const void *src, size_t bytes, hipStream_t s) {
    if (!arena_) init_pinned();  // arena mode allocates nothing: without staging, direct copies
    const size_t off = (pin_used_ + 15) & ~size_t(15);
    if (!pin_ || off + bytes > kPinBytes) {  // no room: a direct copy
        FCC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        return;
    }
    FCC_HIP(hipMemcpyAsync(pin_ + off, src, bytes, hipMemcpyDeviceToHost, s));
    pend_.push_back(Pending{dst, off, bytes});
    pin_used_ = off + bytes;
}

void DevPool::sync(hipStream_t s) {
    const hipError_t e = hipStreamSynchronize(s);
    if (e == hipSuccess)
        for (const Pending &p : pend_) memcpy(p.dst, pin_ + p.off, p.bytes);
    discard();
    if (e != hipSuccess) throw HipError{e, "hipStreamSynchronize (read-back)"};
}

void *DevPool::get(size_t bytes) {
    if (bytes == 0) bytes = 16;
    bytes = (bytes + 255) & ~size_t(255);
    if (arena_) {
        if (bytes > arena_bytes_ - arena_used_)
            throw LibError{ERR_INVALID, "the contraction needs more than the " + std::to_string(arena_bytes_ >> 20) +
                                            " MiB no-allocation scratch (" + std::to_string(arena_used_) +
                                            " bytes taken, " + std::to_string(bytes) + " more asked)"};
        void *p = arena_ + arena_used_;
        arena_used_ += bytes;
        return p;
    }
    auto it = free_.lower_bound(bytes);
    if (it != free_.end() && it->first <= bytes * 2 + (1u << 20)) {
        void *p = it->second;
        size_t sz = it->first;
        free_.erase(it);
        used_[p] = sz;
        in_use_ += sz;
        return p;
    }
    if (budget_ && cached_ + bytes > budget_) {
        // over the budget: give cached blocks back first
        for (auto &kv : free_) {
            (void)hipFree(kv.second);
            cached_ -= kv.first;
        }
        free_.clear();
        if (cached_ + bytes > budget_)
            throw LibError{ERR_OOM, "device memory budget of " + std::to_string(budget_) + " bytes exceeded (" +
                                        std::to_string(cached_) + " held, " + std::to_string(bytes) + " asked)"};
    }
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        // give cached blocks back and retry once
        (void)hipGetLastError();
        for (auto &kv : free_) {
            (void)hipFree(kv.second);
            cached_ -= kv.first;
        }
        free_.clear();
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw LibError{ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed"};
        }
    }
    cached_ += bytes;
    used_[p] = bytes;
    in_use_ += bytes;
    return p;
}

void DevPool::put(void *p) {
    auto it = used_.find(p);
    if (it == used_.end()) return;
    free_.emplace(it->second, p);
    in_use_ -= it->second;
    used_.erase(it);
}

void DevPool::release_all() {
    for (auto &kv : free_) (void)hipFree(kv.second);
    for (auto &kv : used_) (void)hipFree(kv.first);
    free_.clear();
    used_.clear();
    cached_ = 0;
    in_use_ = 0;
}

// ------------------------------------------------------------------- scan

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <typename T> __device__ inline T wave_incl_scan(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        T o = __shfl_up(v, d, 64);
        if (lane >= d) v += o;
    }
    return v;
}

// Exclusive block scan of one value per thread; returns the prefix and writes
// the block total to *total.
template <typename T> __device__ inline T block_excl_scan(T v, T *total) {
    __shared__ T wsum[SCAN_THREADS / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    T inc = wave_incl_scan(v);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SCAN_THREADS / 64; i++) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

template <typename T> __global__ __launch_bounds__(SCAN_THREADS) void k_block_reduce(const T *in, size_t n, T *sums) {
    size_t base = (size_t)blockIdx.x * SCAN_TILE;
    T acc = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        size_t k = base + threadIdx.x + (size_t)i * SCAN_THREADS;
        if (k < n) acc += in[k];
    }
    T tot;
    (void)block_excl_scan(acc, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// A tile is staged through LDS: coalesced global loads and stores, and each
// thread scans SCAN_ITEMS consecutive elements read back from LDS. One pad
// slot per SCAN_ITEMS elements spreads a thread's run over the banks
// (thread t's run starts at t * (SCAN_ITEMS + 1)). in may equal out: a block
// reads its whole tile before it writes.
__device__ __forceinline__ int scan_pad(int e) { return e + e / SCAN_ITEMS; }

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_block_scan(const T *in, T *out, size_t n, const T *offsets,
                                                             T *total) {
    __shared__ T s[SCAN_TILE + SCAN_THREADS];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE;
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int e = i * SCAN_THREADS + t;
        s[scan_pad(e)] = base + e < n ? in[base + e] : T(0);
    }
    __syncthreads();
    T v[SCAN_ITEMS];
    T sum = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
        const T x = s[t * (SCAN_ITEMS + 1) + j];
        v[j] = sum;
        sum += x;
    }
    T tot;
    const T pre = block_excl_scan(sum, &tot);  // its barriers also end every read of s above
    const T off = (offsets ? offsets[blockIdx.x] : T(0)) + pre;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) s[t * (SCAN_ITEMS + 1) + j] = v[j] + off;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; i++) {
        const int e = i * SCAN_THREADS + t;
        if (base + e < n) out[base + e] = s[scan_pad(e)];
    }
    if (total && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
        *total = (offsets ? offsets[blockIdx.x] : T(0)) + tot;
}

template <typename T> __global__ void k_set_scalar(T *p, T v) { *p = v; }

template <typename T>
static void scan_impl(DevPool &pool, const T *in, T *out, size_t n, T *total_dev, hipStream_t s) {
    if (n == 0) {
        if (total_dev) k_set_scalar<T><<<1, 1, 0, s>>>(total_dev, T(0));
        return;
    }
    size_t nb = div_up(n, SCAN_TILE);
    if (nb == 1) {
        k_block_scan<T><<<1, SCAN_THREADS, 0, s>>>(in, out, n, nullptr, total_dev);
        return;
    }
    Tmp<T> sums(pool, nb);
    k_block_reduce<T><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, n, sums);
    scan_impl<T>(pool, sums, sums, nb, nullptr, s);
    k_block_scan<T><<<(unsigned)nb, SCAN_THREADS, 0, s>>>(in, out, n, sums, total_dev);
}

void scan_u32(DevPool &pool, const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev, hipStream_t s) {
    scan_impl<uint32_t>(pool, in, out, n, total_dev, s);
}
void scan_u64(DevPool &pool, const uint64_t *in, uint64_t *out, size_t n, uint64_t *total_dev, hipStream_t s) {
    scan_impl<uint64_t>(pool, in, out, n, total_dev, s);
}

// -------------------------------------------------------------- radix sort
// One pass = histogram per 4096-element tile, exclusive scan of the
// digit-major histogram table, stable scatter. Inside a tile each wavefront
// owns 1024 consecutive elements and ranks them in 16 rounds of 64 (lane
// order = element order): equal digits within a round are found with 8
// ballots, and a wavefront-private LDS histogram carries the ranks across its
// rounds without block barriers. One block-wide scan over (digit, wavefront)
// then places every element; the tile is written out digit run by digit run.

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_WAVE_ELEMS = RS_TILE / RS_WAVES;  // 1024

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t *keys, size_t n, int shift, uint32_t *hist,
                                                        unsigned nblocks) {
    // four histograms per wavefront, lane l counting into copy l % 4:
    // skewed digits (Zipf keys, many lanes on one digit) serialise 4x less on
    // one address; rows are 257 words apart so equal digits of different
    // copies sit in different banks
    constexpr int HC = RS_WAVES * 4, HS = 257;
    __shared__ uint32_t h[HC * HS];
    for (int c = 0; c < HC; c++) h[c * HS + threadIdx.x] = 0;
    __syncthreads();
    uint32_t *hc = h + ((threadIdx.x >> 6) * 4 + (threadIdx.x & 3)) * HS;
    size_t t0 = (size_t)blockIdx.x * RS_TILE;
    uint32_t dg[RS_ITEMS];  // all loads in flight before the first atomic
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        size_t idx = t0 + (size_t)j * RS_THREADS + threadIdx.x;
        dg[j] = idx < n ? (uint32_t)(keys[idx] >> shift) & 255u : 256u;
    }
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++)
        if (dg[j] < 256u) atomicAdd(&hc[dg[j]], 1u);
    __syncthreads();
    uint32_t c = 0;
#pragma unroll
    for (int k = 0; k < HC; k++) c += h[k * HS + threadIdx.x];
    hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = c;
}

template <typename VT>
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t *keys, const VT *vals, uint64_t *okeys,
                                                           VT *ovals, size_t n, int shift, const uint32_t *hist_scan,
                                                           unsigned nblocks) {
    __shared__ uint64_t s_key[RS_TILE];
    __shared__ VT s_val[RS_TILE];
    __shared__ uint32_t s_wh[RS_WAVES][256];  // per wavefront: digit counts, then its digit bases
    __shared__ uint32_t s_gbase[256];         // global start of this tile's run of digit d
    __shared__ uint32_t s_lbase[256];         // start of digit d's run inside the tile
    __shared__ uint32_t s_tot[RS_WAVES];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t t0 = (size_t)blockIdx.x * RS_TILE;
    const uint32_t tn = (uint32_t)(n - t0 < (size_t)RS_TILE ? n - t0 : (size_t)RS_TILE);
    // wavefront w owns tile elements [w * 1024, (w + 1) * 1024); item j of a
    // lane is element w * 1024 + j * 64 + lane (coalesced loads)
    uint64_t rk[RS_ITEMS];
    VT rv[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        const uint32_t li = (uint32_t)(wave * RS_WAVE_ELEMS + j * 64 + lane);
        rk[j] = li < tn ? keys[t0 + li] : 0ull;
        rv[j] = li < tn ? (vals ? vals[t0 + li] : VT(t0 + li)) : VT(0);  // no vals: the entry numbers
    }
#pragma unroll
    for (int w = 0; w < RS_WAVES; w++) s_wh[w][tid] = 0;
    __syncthreads();
    // ranks inside the wavefront's sub-tile: count of earlier equal digits
    uint32_t *wh = s_wh[wave];
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    uint32_t rank[RS_ITEMS];
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        const bool valid = (uint32_t)(wave * RS_WAVE_ELEMS + j * 64 + lane) < tn;
        const uint32_t d = (uint32_t)(rk[j] >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t bb = __ballot(bit);
            peers &= bit ? bb : ~bb;
        }
        const uint32_t before = (uint32_t)__popcll(peers & lt_mask);
        // every lane reads the running count before the group's first lane
        // adds the group (one wavefront: LDS operations stay in order)
        const uint32_t run = valid ? wh[d] : 0u;
        rank[j] = run + before;
        if (valid && before == 0) wh[d] = run + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // digit `tid`: tile total, exclusive scan over digits, per-wavefront bases
    {
        uint32_t c[RS_WAVES], tot = 0;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            c[w] = s_wh[w][tid];
            tot += c[w];
        }
        uint32_t incl = tot;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(incl, d, 64);
            if (lane >= d) incl += o;
        }
        if (lane == 63) s_tot[wave] = incl;
        __syncthreads();
        uint32_t off = 0;
        for (int w = 0; w < wave; w++) off += s_tot[w];
        const uint32_t lb = off + incl - tot;
        s_lbase[tid] = lb;
        s_gbase[tid] = hist_scan[(size_t)tid * nblocks + blockIdx.x];
        uint32_t acc = lb;
#pragma unroll
        for (int w = 0; w < RS_WAVES; w++) {
            s_wh[w][tid] = acc;
            acc += c[w];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; j++) {
        if ((uint32_t)(wave * RS_WAVE_ELEMS + j * 64 + lane) < tn) {
            const uint32_t d = (uint32_t)(rk[j] >> shift) & 255u;
            const uint32_t pos = wh[d] + rank[j];
            s_key[pos] = rk[j];
            s_val[pos] = rv[j];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < tn; i += RS_THREADS) {
        const uint64_t k = s_key[i];
        const uint32_t d = (uint32_t)(k >> shift) & 255u;
        const uint32_t pos = s_gbase[d] + (i - s_lbase[d]);
        okeys[pos] = k;
        ovals[pos] = s_val[i];
    }
}

template <typename VT>
__global__ void k_iota(VT *v, size_t n) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) v[i] = VT(i);
}

// vals_in null: the payload is each key's input position (0, 1, 2, ...),
// made by the first pass instead of read
template <typename VT>
static void radix_sort_impl(DevPool &pool, uint64_t *keys, const VT *vals_in, VT *vals, uint64_t *keys_alt,
                            VT *vals_alt, size_t n, int nbits, hipStream_t s, uint64_t **keys_out, VT **vals_out,
                            int lo_bit) {
    if (n >= (1ull << 32)) throw LibError{ERR_INVALID, "radix sort limited to < 2^32 elements"};
    uint64_t *ka = keys, *kb = keys_alt;
    VT *va = vals, *vb = vals_alt;
    const VT *src = vals_in;  // the first pass's payload (then the alternates)
    if (n > 1 && nbits > 0) {
        unsigned nblocks = (unsigned)div_up(n, RS_TILE);
        size_t hn = (size_t)256 * nblocks;
        Tmp<uint32_t> hist(pool, hn);
        for (int shift = lo_bit; shift < lo_bit + nbits; shift += 8) {
            k_rs_hist<<<nblocks, RS_THREADS, 0, s>>>(ka, n, shift, hist, nblocks);
            scan_u32(pool, hist, hist, hn, nullptr, s);
            k_rs_scatter<VT><<<nblocks, RS_THREADS, 0, s>>>(ka, src, kb, vb, n, shift, hist, nblocks);
            std::swap(ka, kb);
            std::swap(va, vb);
            src = va;
        }
    } else if (n && src != va) {
        if (src) FCC_HIP(hipMemcpyAsync(va, src, n * sizeof(VT), hipMemcpyDeviceToDevice, s));
        else k_iota<VT><<<(unsigned)div_up(n, 256), 256, 0, s>>>(va, n);
    }
    *keys_out = ka;
    *vals_out = va;
}

void radix_sort_pairs(DevPool &pool, uint64_t *keys, uint32_t *vals, uint64_t *keys_alt, uint32_t *vals_alt,
                      size_t n, int nbits, hipStream_t s, uint64_t **keys_out, uint32_t **vals_out, int lo_bit,
                      bool positions) {
    radix_sort_impl<uint32_t>(pool, keys, positions ? nullptr : vals, vals, keys_alt, vals_alt, n, nbits, s,
                              keys_out, vals_out, lo_bit);
}

void radix_sort_keys_values(DevPool &pool, uint64_t *keys, const uint64_t *vals_in, uint64_t *vals,
                            uint64_t *keys_alt, uint64_t *vals_alt, size_t n, int nbits, hipStream_t s,
                            uint64_t **keys_out, uint64_t **vals_out) {
    radix_sort_impl<uint64_t>(pool, keys, vals_in, vals, keys_alt, vals_alt, n, nbits, s, keys_out, vals_out, 0);
}

}  // namespace fcc
