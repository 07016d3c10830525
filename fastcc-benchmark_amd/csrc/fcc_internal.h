// Internal declarations shared by the HIP translation units of libfastcc_hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace fcc {

constexpr int kMaxOrder = 8;

// Contraction geometry, computed on the host from the inferred extents and
// passed to kernels by value.
struct Geom {
    int dA, dB, r, dC;
    int am[kMaxOrder], bm[kMaxOrder];      // paired contraction modes
    int nxa, nxb;
    int xa[kMaxOrder], xb[kMaxOrder];      // external modes, ascending
    uint32_t xaE[kMaxOrder], xbE[kMaxOrder];
    uint64_t R[kMaxOrder];                 // contraction-key radix per pair
    uint64_t spanA, spanB, K;              // external spans (< 2^32), key span
    int rbits, cbits, kbits;               // bit widths of row, col, key
};

// Error codes mirrored from fcc.h
enum : int { OK = 0, ERR_INVALID = -1, ERR_OOM = -2, ERR_OVERFLOW = -3, ERR_HIP = -4 };

// Caching device allocator: blocks are kept across calls so a steady-state
// contraction performs no hipMalloc.
class DevPool {
  public:
    ~DevPool();
    void *get(size_t bytes);
    void put(void *p);
    void release_all();
    size_t bytes_cached() const { return cached_; }
    size_t bytes_in_use() const { return in_use_; }
    // device bytes the pool may hold (cached + in use); 0 = no limit. An
    // allocation beyond it first returns cached blocks, then fails with
    // ERR_OOM, as a failed hipMalloc does.
    void set_budget(size_t bytes) { budget_ = bytes; }
    // Arena mode: every get() is carved from [base, base + bytes) in order
    // (no device allocation; put() of such a block is a no-op, the arena is
    // reused whole by the next begin_arena). Exceeding it throws ERR_INVALID.
    void begin_arena(void *base, size_t bytes) {
        arena_ = (uint8_t *)base;
        arena_bytes_ = bytes;
        arena_used_ = 0;
    }
    void end_arena() { arena_ = nullptr; }
    bool in_arena() const { return arena_ != nullptr; }
    size_t arena_peak() const { return arena_used_; }
    // the pinned staging buffer of read(), allocated up front (context creation)
    void init_pinned();

    // Small device-to-host reads staged through pinned host memory (a
    // pageable destination makes every copy a slower synchronous one):
    // read() queues a copy on stream s, sync(s) waits for s and delivers every
    // queued read to its destination; discard() drops queued reads.
    void read(void *dst, const void *src, size_t bytes, hipStream_t s);
    void sync(hipStream_t s);
    void discard() {
        pend_.clear();
        pin_used_ = 0;
    }

  private:
    std::multimap<size_t, void *> free_;
    std::map<void *, size_t> used_;
    size_t cached_ = 0;
    size_t in_use_ = 0;
    size_t budget_ = 0;
    uint8_t *arena_ = nullptr;
    size_t arena_bytes_ = 0, arena_used_ = 0;
    static constexpr size_t kPinBytes = 1 << 16;
    struct Pending {
        void *dst;
        size_t off, bytes;
    };
    uint8_t *pin_ = nullptr;
    size_t pin_used_ = 0;
    std::vector<Pending> pend_;
};

// RAII temporary from the pool.
template <typename T> struct Tmp {
    DevPool *pool = nullptr;
    T *p = nullptr;
    Tmp() = default;
    Tmp(DevPool &pl, size_t n) : pool(&pl) { p = (T *)pl.get(n * sizeof(T) + 16); }
    Tmp(const Tmp &) = delete;
    Tmp &operator=(const Tmp &) = delete;
    Tmp(Tmp &&o) noexcept : pool(o.pool), p(o.p) { o.p = nullptr; }
    Tmp &operator=(Tmp &&o) noexcept {
        reset();
        pool = o.pool; p = o.p; o.p = nullptr;
        return *this;
    }
    ~Tmp() { reset(); }
    void reset() { if (p && pool) pool->put(p); p = nullptr; }
    T *get() const { return p; }
    operator T *() const { return p; }
};

struct HipError {
    hipError_t code;
    std::string where;
};
struct LibError {
    int code;
    std::string msg;
};

#define FCC_HIP(x)                                                       \
    do {                                                                 \
        hipError_t e_ = (x);                                             \
        if (e_ != hipSuccess) throw ::fcc::HipError{e_, #x};             \
    } while (0)

inline int ceil_log2(uint64_t x) {
    int b = 0;
    while (b < 64 && (1ULL << b) < x) b++;
    return b;
}
inline uint64_t pow2_ceil(uint64_t x) { return x <= 1 ? 1 : 1ULL << ceil_log2(x); }
inline uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------- scan / sort
// Exclusive scan of n elements (in may equal out). *total_dev (optional)
// receives the sum on the device.
void scan_u32(DevPool &pool, const uint32_t *in, uint32_t *out, size_t n, uint32_t *total_dev,
              hipStream_t s);
void scan_u64(DevPool &pool, const uint64_t *in, uint64_t *out, size_t n, uint64_t *total_dev,
              hipStream_t s);

// Stable LSD radix sort of (key, value) pairs on key bits [lo_bit, lo_bit +
// nbits) (8-bit digits; key bits above lo_bit + nbits must be zero, bits below
// lo_bit keep their input order). Returns pointers to the sorted arrays
// (either the inputs or the alternates). positions: the values are the keys'
// input positions 0..n-1, generated by the sort (vals is only a buffer).
void radix_sort_pairs(DevPool &pool, uint64_t *keys, uint32_t *vals, uint64_t *keys_alt,
                      uint32_t *vals_alt, size_t n, int nbits, hipStream_t s,
                      uint64_t **keys_out, uint32_t **vals_out, int lo_bit = 0, bool positions = false);
// The same with a 64-bit payload whose first pass reads vals_in (left
// unchanged) and whose later passes ping-pong between vals and vals_alt.
void radix_sort_keys_values(DevPool &pool, uint64_t *keys, const uint64_t *vals_in, uint64_t *vals,
                            uint64_t *keys_alt, uint64_t *vals_alt, size_t n, int nbits, hipStream_t s,
                            uint64_t **keys_out, uint64_t **vals_out);

// ------------------------------------------------------------------- index
struct Operand {
    int order;
    int64_t nnz;
    const uint32_t *coords;  // device SoA
    const double *vals;      // device
};

// Accumulator kinds (match FCC_ACC_*).
// DIRECT: a tile whose entries all sit on distinct rows; every product is
// then a distinct output and is written straight out, no accumulator.
enum : uint8_t { KIND_DENSE = 1, KIND_LHASH = 2, KIND_GHASH = 3, KIND_DIRECT = 4 };
// CBITS: count pass only -- a hash unit counted through one LDS bit per cell
// of its rows x (padded) window when that fits the launch's key-set LDS.
constexpr uint8_t KIND_CBITS = 5;
constexpr uint32_t DIRECT_MAX_TL = 64;  // rows per tile up to which DIRECT tiles are detected
constexpr uint64_t DIRECT_RUN = 65536;  // products per unit made of consecutive DIRECT tiles

constexpr uint32_t HEAVY_MIN = 64;      // default: fibers at least this long are processed wave-per-group
constexpr uint32_t DENSE_SLOTS = 9984;  // f64 LDS accumulator slots (78 KB: two workgroups per CU)
constexpr uint32_t LH_SLOTS = 4096;     // LDS hash slots (f64 value + u32 key)
constexpr uint32_t LH_MAX = 3072;       // LDS hash distinct-key limit before sealing
constexpr uint32_t SYM_SLOTS = 16384;   // LDS hash-set slots of the count pass (u32 keys)
constexpr uint32_t SYM_MAX = 12288;
constexpr uint32_t HASH_TARGET = 2048;  // planned distinct outputs per hash unit
constexpr int ACC_THREADS = 512;
constexpr uint32_t HSEG = 512;  // fiber entries per heavy segment

// Device-side index of one contraction: B's CSR keyed by contraction key
// ("fibers") and A's entries grouped by (left tile, fiber).
struct Index {
    // B: nB entries sorted by (key, col); fiber f = [fiber_ptr[f], fiber_ptr[f+1])
    uint32_t nB = 0, nfib = 0;
    uint32_t *Bcol = nullptr;
    double *Bval = nullptr;
    uint32_t *fiber_ptr = nullptr;
    uint64_t *fiber_key = nullptr;
    uint32_t *Bfib = nullptr;  // nB: the fiber of each position
    // A: nAv entries with a matching fiber, sorted by (tile, class, fiber, row-in-tile)
    uint32_t nAv = 0, ngroups = 0, natiles = 0;
    uint32_t *Ali = nullptr;   // row within tile
    uint32_t *Arow = nullptr;  // global row
    double *Aval = nullptr;
    uint32_t *Afib = nullptr;  // fiber id per entry
    uint2 *Afp = nullptr;      // fiber [start, end) per entry
    uint32_t *group_ptr = nullptr;    // ngroups+1
    uint4 *gdesc = nullptr;           // per group: fiber start, fiber length, first entry, entries
    uint32_t *group_fiber = nullptr;
    uint32_t *group_atile = nullptr;
    uint32_t *atile_id = nullptr;     // tile index of active tile a
    uint32_t *atile_g0 = nullptr;     // natiles+1: first group of active tile a
    uint32_t *atile_lg = nullptr;     // first light group of active tile a
    // heavy groups cut into segments of at most HSEG fiber entries, so long
    // fibers spread over many wavefronts; tile a's segments are
    // [atile_h0[a], atile_h0[a + 1])
    uint32_t nhseg = 0;
    uint32_t hseg = HSEG;             // fiber entries per segment
    uint32_t seg_below = 16;          // only tiles with fewer heavy groups than this get segments
                                      // (the accumulate kernels walk the others group by group)
    // dense keys (dk_hkey[fiber] != ~0u: contracted by the dense GEMM): the
    // left build takes their groups out of the tile traversal and flags
    // their tiles in tile_dk (natiles + 1 bytes); null: no dense keys
    const uint32_t *dk_hkey = nullptr;
    uint8_t *tile_dk = nullptr;
    uint4 *hdesc = nullptr;           // per segment: fiber start, length, first entry, entries
    uint32_t *atile_h0 = nullptr;     // natiles+1
    // per active tile statistics (planner features)
    uint64_t *st_products = nullptr;  // sum over groups of m * |fiber|
    uint64_t *st_sumfib = nullptr;    // sum over groups of |fiber|
    uint32_t *st_nnz = nullptr;       // A entries in tile
    uint32_t *st_keys = nullptr;      // groups (distinct contraction keys) in tile
    uint32_t *st_rows = nullptr;      // distinct rows in tile
    uint32_t *st_sketch = nullptr;    // distinct-column sketch (KMV estimate)
    uint64_t *st_maxgp = nullptr;     // largest single-group product m * |fiber| (lower bound on outputs)
    uint8_t *st_direct = nullptr;     // 1: the tile's entries sit on distinct rows (DIRECT-able)
    uint64_t *st_outd = nullptr;      // outputs of the tile when DIRECT: st_products, or in mirror
                                      // mode its products at or above the diagonal, counted twice
                                      // off it (aliases st_products when not mirrored)
    uint64_t *st_pprefix = nullptr;   // natiles+1: exclusive prefix sum of st_outd
    int logTL = 0;
    uint32_t TL = 1;
    uint64_t products = 0;
    uint32_t heavy_min = HEAVY_MIN;  // fiber length from which groups are processed wave-per-group
    bool features = true;            // per-tile planner features (keys, fiber sums, sketches) wanted
    // mirror mode: a self-contraction (one tensor, the same mode lists on both
    // sides) has a symmetric output C(x, y) = C(y, x); each unordered pair is
    // computed once, at its cell with column >= row, and written twice
    bool mirror = false;
    // kept for plans (new values on the same coordinates): the input entry
    // behind each sorted position, so Bval/Aval can be gathered again
    bool keep_perm = false;
    uint32_t *permB = nullptr;  // nB
    uint32_t *permA = nullptr;  // nAv
};

struct Units {
    uint32_t n = 0;
    uint32_t *tile = nullptr;  // active tile index
    uint32_t *c0 = nullptr;    // first column of the window
    uint32_t *W = nullptr;     // window width
    uint8_t *kind = nullptr;
    uint8_t *multi = nullptr;  // 1 when the tile has more than one window
    float *pred = nullptr;     // predicted outputs (learned planner)
};

// Everything an accumulate kernel needs, passed by value.
struct AccArgs {
    const uint32_t *Bcol;
    const double *Bval;
    const uint32_t *fiber_ptr;
    const uint32_t *Ali;
    const double *Aval;
    const uint2 *Afp;
    const uint32_t *Afib;         // fiber per left entry
    // dense column windows of width fcut_W: fcut[f * fcut_n + w] = first
    // position of fiber f at a column >= w * fcut_W (w = 0..fcut_n-1), or null
    const uint32_t *fcut;
    uint32_t fcut_n;
    uint32_t fcut_W;
    const uint32_t *group_ptr;
    const uint4 *gdesc;
    const uint4 *hdesc;
    const uint32_t *atile_h0;
    const uint32_t *group_fiber;
    const uint32_t *atile_id;
    const uint32_t *atile_g0;
    const uint32_t *atile_lg;
    const uint64_t *st_products;  // per active tile
    const uint64_t *st_pprefix;   // natiles+1 exclusive prefix of st_products (DIRECT run output counts)
    const uint32_t *Arow;         // global row per left entry
    uint32_t natiles;
    uint32_t nunits;
    const uint32_t *unit_tile, *unit_c0, *unit_W;
    const uint8_t *unit_kind, *unit_multi;
    const uint32_t *unit_order;  // processing order (heaviest first) or null
    int logTL;
    uint64_t spanB;
    int nxa, nxb;
    uint32_t xaE[kMaxOrder], xbE[kMaxOrder];
    // shared global hash (overflow / GHASH units)
    unsigned long long *gkeys;
    double *gvals;
    uint32_t *gunit;
    uint64_t gmask;
    unsigned long long *gcount;
    uint32_t *gflag;
    // count pass
    unsigned long long *unit_count;
    // placement
    const unsigned long long *unit_off;  // exact planner
    unsigned long long *unit_cursor;     // exact planner, GHASH entries
    unsigned long long *cursor;          // learned planner
    uint64_t capacity;
    uint32_t *fail_flag;
    uint32_t *ocoords;
    double *ovals;
    uint64_t ostride;
    // sink: per group g, sink[4g]=count sink[4g+1]=hash, sinkf[4g]=val_sum
    // sinkf[4g+1]=val_wsum (sinkf aliases sink + 2)
    unsigned long long *sink;
    double *sinkf;
    // groups of rows (independent contractions fused into one launch): group
    // g holds left rows [g*group_rows, (g+1)*group_rows); 0 = one group.
    // Sink summaries are kept per group, hashed on the row within the group.
    uint64_t group_rows;
    unsigned int *group_left;            // units still to finish per group, or null
    unsigned long long *group_done;      // wall clock when a group's last unit finished
    unsigned long long *clock_start;     // wall clock when the first unit started
    unsigned int *work;        // persistent work counter
    int row_own;               // dense units: row-owned plain updates (experimental)
    uint32_t claim_div;        // heavy claims: segments / claim_div per claim (0: 64 per claim)
    uint32_t seg_below;        // tiles with fewer heavy groups than this work on segments
    uint32_t set_words;        // count pass: LDS words of its key set / byte map (launch size)
    unsigned long long *block_counted;  // count pass: units counted from the dense-key block
    // dense-key block over left rows x right columns: presence bits dk_P
    // (row pitch dk_pw words) and values dk_C (row pitch dk_rb); units of the
    // active tiles flagged in dk_tile start their accumulator from it.
    // dk_tile == null: no block.
    const uint32_t *dk_P;
    const double *dk_C;
    const uint8_t *dk_tile;
    uint64_t dk_pw, dk_rb;
    uint64_t dk_slab;    // dk_C holds dk_splits partial blocks dk_slab doubles apart; a value is their sum
    uint32_t dk_splits;
    uint64_t spanA;
    // mirror mode (self-contractions, spanA == spanB): only cells (r, c) with
    // c >= r are accumulated; each is written as (r, c) and, off the
    // diagonal, as (c, r)
    int mirror;
    // the index was built for mirroring (fiber ranges start at the entries'
    // own columns) but this pass wants whole fibers: ranges start at the
    // fiber's first entry (the sampled exact count that judges the planner)
    int unclipped;
};

// Mirror mode: outputs written for cell (r, c): none below the diagonal, one
// on it, two (the cell and its transpose) above it.
__device__ __forceinline__ uint32_t mirror_weight(uint64_t r, uint64_t c) { return c < r ? 0u : (c == r ? 1u : 2u); }

}  // namespace fcc
