// Particle sort by cell on gfx950: stable LSD radix sort of (cell, index) pairs, then one
// permutation gather per particle dat (the p2c dat itself is written from the sorted keys).
//
// Reference: opp_particle_sort / the Sort_* fill policies (opp_lib/src/core/opp_particle_organize_core.cpp:98-121,
// opp_lib/src/omp/opp_particle_mover.cpp:223-265, hip opp_lib/src/hip/opp_particle_sorter.cpp:188-267) order the
// particles by p2c with removed particles (MAX_CELL_INDEX) last. Their tie order differs between backends
// (std::sort / thrust radix); this build's contract is stable by (p2c, index).
//
// Each pass sorts by one 8-bit digit. A tile is 256 * sort_items consecutive elements (tuning), each
// wave of its block a contiguous part of them; ranks come from wave-level digit matching (8 ballots
// give the lanes holding the same digit) against the wave's own digit counts, so equal digits keep
// their index order and no LDS atomic is contended by nearly sorted keys. Per-tile digit counts (a
// histogram pass that needs no ranks) are scanned digit-major across the tiles to give every
// (digit, tile) its output offset. Every kernel reads the live count on the device: the histogram and
// scatter grids (sized by the capacity, capped at a few times the resident blocks) stride over the
// live tiles only, the counts are laid out with the live tile count as row stride, and the scan
// covers only those rows.

#include "oppx_internal.h"

namespace oppx {
namespace {

constexpr int kB = 256;
constexpr int kItems = 16;           // elements per thread of the scan kernels
constexpr int kElems = kB * kItems;
constexpr int kMinSortItems = 4;     // smallest radix block (elements per thread) the scratch is sized for
constexpr int kBins = 256;
constexpr int kMaxDats = 16;
constexpr int kRemSlots = 256;  // removed-particle counters (block id modulo kRemSlots)
constexpr unsigned kGridMult = 4;  // grids that stride over live work: this many times the resident blocks

struct GatherTable {
    int n;
    int64_t stride;
    const void* in[kMaxDats];
    void* out[kMaxDats];
    int dim[kMaxDats];
    int eb[kMaxDats];
};

__device__ __forceinline__ int64_t live_n(const oppx_move_counters* ctr, int64_t n_max) {
    const int64_t n = ctr->size;
    return n < 0 ? 0 : (n > n_max ? n_max : n);
}

// lanes of the wave (among `active`) holding the same 8-bit digit as this lane
__device__ __forceinline__ unsigned long long digit_peers(uint32_t d, unsigned long long active) {
    unsigned long long peers = active;
#pragma unroll
    for (int bit = 0; bit < 8; bit++) {
        const unsigned long long m = __ballot((d >> bit) & 1u);
        peers &= ((d >> bit) & 1u) ? m : ~m;
    }
    return peers;
}

// Key of a particle: its cell, MAX_CELL_INDEX mapped to n_cells (the last bucket). The first digit
// pass reads the keys from p2c and takes the particle index as the value, so no key/value arrays
// are written beforehand; it also counts the removed particles.
__device__ __forceinline__ uint32_t cell_key(int32_t c, uint32_t n_cells) {
    return c == kMaxCell ? n_cells : (uint32_t)c;
}

// Per-block digit counts of one pass. All of a thread's keys are loaded before any is counted
// (independent loads in flight). Counting needs no ranks: a wave whose live lanes share one digit adds
// once (nearly sorted keys), any other wave adds per lane.
template <int ITEMS, bool FIRST>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(8))) sort_hist(const uint32_t* __restrict__ keys, const int32_t* __restrict__ p2c,
                                                uint32_t n_cells, const oppx_move_counters* ctr, int64_t n_max,
                                                int shift, uint32_t* __restrict__ hist,
                                                unsigned long long* __restrict__ n_removed) {
    __shared__ uint32_t s_h[kBins];
    const int64_t n = live_n(ctr, n_max);
    const int lane = threadIdx.x & (kWave - 1);
    // the grid strides over the tiles below the live count; tile t's count of digit d goes to
    // hist[d * tiles + t], so the scan covers the live tiles only
    const int64_t tiles = (n + kB * ITEMS - 1) / (kB * ITEMS);
    unsigned long long removed = 0;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        s_h[threadIdx.x] = 0;  // each thread clears the word it wrote out last tile
        __syncthreads();
        const int64_t base = t * (kB * ITEMS);
        uint32_t key[ITEMS];
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const int64_t i = base + r * kB + threadIdx.x;
            key[r] = i < n ? (FIRST ? cell_key(p2c[i], n_cells) : keys[i]) : 0u;
        }
#pragma unroll
        for (int r = 0; r < ITEMS; r++) {
            const bool ok = base + r * kB + threadIdx.x < n;
            const uint32_t k = key[r];
            if (FIRST) removed += (unsigned long long)__popcll(__ballot(ok && k == n_cells));
            const uint32_t d = (k >> shift) & 0xFFu;
            const unsigned long long act = __ballot(ok);
            if (act == 0ull) continue;  // uniform over the wave
            const int lead = __ffsll((long long)act) - 1;
            const uint32_t d_lead = (uint32_t)__shfl((int)d, lead);
            if (__ballot(ok && d == d_lead) == act) {
                if (lane == lead) atomicAdd(&s_h[d], (uint32_t)__popcll(act));
            } else if (ok) {
                atomicAdd(&s_h[d], 1u);
            }
        }
        __syncthreads();
        hist[(int64_t)threadIdx.x * tiles + t] = s_h[threadIdx.x];  // digit-major rows of the live tiles
    }
    // per-wave removed counts land on one of kRemSlots counters (block id modulo kRemSlots): one
    // shared counter would serialise every wave's atomic in one L2 channel
    if (FIRST && lane == 0 && removed) atomicAdd(n_removed + (blockIdx.x % kRemSlots), removed);
}

// ---- one digit pass over a tile of 256 * ITEMS consecutive elements ----
// Wave w of the block owns the 64 * ITEMS consecutive elements [w * 64 * ITEMS, (w + 1) * 64 * ITEMS)
// of the tile and takes them 64 at a time (round r: element w * 64 * ITEMS + r * 64 + lane), so an
// element's stable rank among the tile's elements of its digit is: that digit's elements in the
// waves before (summed once, after all rounds), plus the wave's own in earlier rounds (a per-wave
// count in LDS, updated round by round without block barriers), plus the lanes below holding the
// same digit (8 ballots). The tile is then laid out in LDS in digit order and each digit's run is
// written to its global offset with consecutive threads on consecutive addresses, so the global
// writes are whole runs instead of one scattered word per lane.
template <int ITEMS>
struct TileLds {
    uint32_t run[kBins];               // the tile's elements of each digit, then the digit's start in the tile
    uint32_t wave[kB / kWave][kBins];  // per-wave digit counts, then their exclusive prefix over the waves
    uint32_t dst[kBins];               // global offset of each digit's run
    uint32_t tot[kB / kWave];          // wave totals of the digit-start scan
    uint32_t key[kB * ITEMS], val[kB * ITEMS];
};

template <int ITEMS>
__device__ __forceinline__ int tile_elem(int tid, int r) {
    return (int)(tid / kWave) * (kWave * ITEMS) + r * kWave + (int)(tid & (kWave - 1));
}

template <int ITEMS, bool FIRST>
__device__ __forceinline__ void tile_load(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                          const int32_t* __restrict__ p2c, uint32_t n_cells, int64_t base, int m, int tid,
                                          uint32_t (&key)[ITEMS], uint32_t (&val)[ITEMS]) {
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const int li = tile_elem<ITEMS>(tid, r);
        const bool ok = li < m;
        if (FIRST) {  // keys straight from p2c, the particle index as the value
            key[r] = ok ? cell_key(p2c[base + li], n_cells) : 0u;
            val[r] = (uint32_t)(base + li);
        } else {
            key[r] = ok ? kin[base + li] : 0u;
            val[r] = ok ? vin[base + li] : 0u;
        }
    }
}

// Ranks of the tile's elements inside their wave's part; leaves the tile's count of every digit in
// L.run and the exclusive prefix over the waves in L.wave. L.wave must be zeroed (and that made
// visible) by the caller.
template <int ITEMS>
__device__ __forceinline__ void tile_rank(TileLds<ITEMS>& L, int tid, const uint32_t (&key)[ITEMS], int m, int shift,
                                          uint32_t (&rank)[ITEMS]) {
    const int w = tid / kWave, lane = tid & (kWave - 1);
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const bool ok = tile_elem<ITEMS>(tid, r) < m;
        const uint32_t d = (key[r] >> shift) & 0xFFu;
        const unsigned long long peers = digit_peers(d, __ballot(ok));
        // every lane reads the wave's count before the lowest lane of each digit group adds the group
        // (one wave's LDS operations complete in program order)
        const uint32_t before = L.wave[w][d];
        rank[r] = before + (uint32_t)__popcll(peers & below);
        if (ok && lane == __ffsll((long long)peers) - 1) L.wave[w][d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    uint32_t run = 0;
    for (int q = 0; q < kB / kWave; q++) {
        const uint32_t c = L.wave[q][tid];
        L.wave[q][tid] = run;
        run += c;
    }
    L.run[tid] = run;
    __syncthreads();
}

// Digit starts inside the tile: exclusive scan of L.run (one digit per thread), in place.
template <int ITEMS>
__device__ __forceinline__ void tile_digit_starts(TileLds<ITEMS>& L, int tid) {
    const int w = tid / kWave, lane = tid & (kWave - 1);
    const uint32_t c = L.run[tid];
    uint32_t inc = c;
#pragma unroll
    for (int dd = 1; dd < kWave; dd <<= 1) {
        const uint32_t t = __shfl_up(inc, dd);
        if (lane >= dd) inc += t;
    }
    if (lane == kWave - 1) L.tot[w] = inc;
    __syncthreads();
    uint32_t off = 0;
    for (int q = 0; q < w; q++) off += L.tot[q];
    L.run[tid] = off + inc - c;
    __syncthreads();
}

// Digit-ordered layout in LDS, then the runs to L.dst[d] + (position in the run). Writes at or past
// `limit` are dropped (never taken by a correct sort; a bound on every store). The last pass writes no
// keys, except as sorted p2c into p2c_out (nullable; key n_cells back to MAX_CELL_INDEX).
template <int ITEMS, bool LAST>
__device__ __forceinline__ void tile_write(TileLds<ITEMS>& L, int tid, const uint32_t (&key)[ITEMS], const uint32_t (&val)[ITEMS],
                                           const uint32_t (&rank)[ITEMS], int m, int shift, uint64_t limit,
                                           uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                           int32_t* __restrict__ p2c_out, uint32_t n_cells) {
    const int w = tid / kWave;
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        if (tile_elem<ITEMS>(tid, r) < m) {
            const uint32_t d = (key[r] >> shift) & 0xFFu;
            const uint32_t pos = L.run[d] + L.wave[w][d] + rank[r];
            L.key[pos] = key[r];
            L.val[pos] = val[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ITEMS; r++) {
        const int p = r * kB + tid;
        if (p < m) {
            const uint32_t k = L.key[p];
            const uint32_t d = (k >> shift) & 0xFFu;
            const uint64_t dst = (uint64_t)L.dst[d] + (uint32_t)(p - (int)L.run[d]);
            if (dst < limit) {
                if (!LAST) kout[dst] = k;
                else if (p2c_out) p2c_out[dst] = k == n_cells ? kMaxCell : (int32_t)k;
                vout[dst] = L.val[p];
            }
        }
    }
}

// Scatter of one digit pass with the (digit, block) offsets from the scanned per-block histograms.
// Four waves per SIMD: the occupancy the 16-key tile's LDS allows (four 39 KB blocks per CU), reached
// without spilling.
template <int ITEMS, bool LAST, bool FIRST>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(4)))
sort_scatter(const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, const int32_t* __restrict__ p2c,
             uint32_t n_cells, const oppx_move_counters* ctr, int64_t n_max, int shift,
             const uint32_t* __restrict__ offs, uint32_t* __restrict__ kout,
             uint32_t* __restrict__ vout, int32_t* __restrict__ p2c_out) {
    __shared__ TileLds<ITEMS> L;
    const int64_t n = live_n(ctr, n_max);
    // the grid (at most the blocks that fit the chip at once) strides over the tiles that hold live
    // elements: a tile past the live count is never dispatched, where each such block would still have
    // had to be given its LDS (40 KB) before it could leave
    const int64_t tiles = (n + kB * ITEMS - 1) / (kB * ITEMS);
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t base = t * (kB * ITEMS);
        const int m = (int)((n - base) < kB * ITEMS ? (n - base) : kB * ITEMS);  // elements of the tile
        // the thread index, opaque to the compiler inside each tile: the per-element indexes and LDS
        // addresses derived from it are recomputed per tile instead of held in registers across the loop
        int tid = (int)threadIdx.x;
        asm volatile("" : "+v"(tid));
        for (int q = 0; q < kB / kWave; q++) L.wave[q][tid] = 0;
        L.dst[tid] = offs[(int64_t)tid * tiles + t];
        uint32_t key[ITEMS], val[ITEMS], rank[ITEMS];
        tile_load<ITEMS, FIRST>(kin, vin, p2c, n_cells, base, m, tid, key, val);
        __syncthreads();
        tile_rank<ITEMS>(L, tid, key, m, shift, rank);
        tile_digit_starts<ITEMS>(L, tid);
        tile_write<ITEMS, LAST>(L, tid, key, val, rank, m, shift, (uint64_t)n, kout, vout, p2c_out, n_cells);
        __syncthreads();  // every thread is done with the tile's LDS before the next tile fills it
    }
}

// ---- exclusive scan of a uint32 array (chunks of kElems) ----
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t* total) {
    __shared__ uint32_t s_w[kB / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t t = __shfl_up(inc, d);
        if (lane >= d) inc += t;
    }
    if (lane == kWave - 1) s_w[w] = inc;
    __syncthreads();
    uint32_t off = 0, all = 0;
    for (int q = 0; q < kB / kWave; q++) {
        if (q < w) off += s_w[q];
        all += s_w[q];
    }
    __syncthreads();
    *total = all;
    return off + inc - v;
}

// Entries of the digit counts of the live tiles (kBins rows of live tiles of per_tile elements).
__device__ __forceinline__ int64_t live_hist_n(const oppx_move_counters* ctr, int64_t n_max, int64_t per_tile) {
    return (int64_t)kBins * ((live_n(ctr, n_max) + per_tile - 1) / per_tile);
}

// The scan of the live digit counts: blocks of chunks past them leave at once.
__global__ void __launch_bounds__(kB) scan_reduce(const uint32_t* __restrict__ a, const oppx_move_counters* ctr,
                                                  int64_t n_max, int64_t per_tile, uint32_t* __restrict__ sums) {
    const int64_t n = live_hist_n(ctr, n_max, per_tile);
    if ((int64_t)blockIdx.x * kElems >= n) return;
    uint32_t s = 0;
    for (int r = 0; r < kItems; r++) {  // the chunk's total: any order, so consecutive threads on consecutive words
        const int64_t i = (int64_t)blockIdx.x * kElems + (int64_t)r * kB + threadIdx.x;
        if (i < n) s += a[i];
    }
    uint32_t tot;
    block_scan_excl(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kB) scan_sums(uint32_t* __restrict__ sums, const oppx_move_counters* ctr,
                                                int64_t n_max, int64_t per_tile) {
    const int64_t nb = (live_hist_n(ctr, n_max, per_tile) + kElems - 1) / kElems;
    uint32_t carry = 0;
    for (int64_t base = 0; base < nb; base += kB) {
        const int64_t j = base + threadIdx.x;
        const uint32_t v = (j < nb) ? sums[j] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan_excl(v, &tot);
        if (j < nb) sums[j] = carry + ex;
        carry += tot;
    }
}

__global__ void __launch_bounds__(kB) scan_apply(uint32_t* __restrict__ a, const oppx_move_counters* ctr,
                                                 int64_t n_max, int64_t per_tile, const uint32_t* __restrict__ sums) {
    // The chunk passes through LDS so that global loads and stores are coalesced while each thread scans
    // kItems consecutive words; word e sits at e + e / kItems (thread t's words from 17 t on: distinct
    // banks across the wave).
    __shared__ uint32_t s_a[kElems + kElems / kItems];
    const int64_t n = live_hist_n(ctr, n_max, per_tile);
    const int64_t base = (int64_t)blockIdx.x * kElems;
    if (base >= n) return;
    for (int r = 0; r < kItems; r++) {
        const int e = r * kB + threadIdx.x;
        s_a[e + e / kItems] = base + e < n ? a[base + e] : 0u;
    }
    __syncthreads();
    const int t0 = threadIdx.x * (kItems + 1);
    uint32_t loc[kItems];
    uint32_t s = 0;
    for (int r = 0; r < kItems; r++) {
        loc[r] = s_a[t0 + r];
        s += loc[r];
    }
    uint32_t tot;
    uint32_t run = sums[blockIdx.x] + block_scan_excl(s, &tot);
    for (int r = 0; r < kItems; r++) {
        s_a[t0 + r] = run;
        run += loc[r];
    }
    __syncthreads();
    for (int r = 0; r < kItems; r++) {
        const int e = r * kB + threadIdx.x;
        if (base + e < n) a[base + e] = s_a[e + e / kItems];
    }
}

template <typename T>
__device__ __forceinline__ void gather_rec(const void* in, void* out, int dim, int64_t stride, int64_t from, int64_t to) {
    const T* s = reinterpret_cast<const T*>(in);
    T* d = reinterpret_cast<T*>(out);
    for (int q = 0; q < dim; q++) d[q * stride + to] = s[q * stride + from];
}

__global__ void __launch_bounds__(kB) sort_gather(const uint32_t* __restrict__ perm, const oppx_move_counters* ctr,
                                                  int64_t n_max, GatherTable T, int32_t* __restrict__ perm_out) {
    const int64_t n = live_n(ctr, n_max);
    const int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int64_t from = perm[j];
    if (from >= T.stride) return;  // never read outside the dats
    for (int q = 0; q < T.n; q++) {
        if (T.eb[q] == 8) gather_rec<double>(T.in[q], T.out[q], T.dim[q], T.stride, from, j);
        else gather_rec<int32_t>(T.in[q], T.out[q], T.dim[q], T.stride, from, j);
    }
    if (perm_out) perm_out[j] = (int32_t)from;
}
__global__ void zero_u64(unsigned long long* p) { p[threadIdx.x] = 0ull; }  // kRemSlots threads

__global__ void lower_bound_kernel(const int32_t* __restrict__ p2c, const oppx_move_counters* ctr, int64_t n_max,
                                   const int32_t* __restrict__ keys, int64_t k, int64_t* __restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= k) return;
    const int64_t n = live_n(ctr, n_max);
    const int32_t key = keys[j];
    int64_t lo = 0, hi = n;  // first index with p2c >= key in [lo, hi]
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (p2c[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    out[j] = lo;
}

__global__ void sort_finalize(oppx_move_counters* ctr, int64_t n_max, const unsigned long long* n_removed, int drop) {
    const int64_t n = live_n(ctr, n_max);
    int64_t r = 0;
    for (int q = 0; q < kRemSlots; q++) r += (int64_t)n_removed[q];
    ctr->removed = (int32_t)r;
    ctr->holes_low = 0;
    if (drop) ctr->size = (int32_t)(n - r);
}

struct SortScratch {
    uint32_t *ka, *va, *kb, *vb, *hist, *sums;
    unsigned long long* n_removed;
    int64_t nblocks, hist_n, scan_blocks;
};

// int32 words of the histogram region: per-block histograms at the smallest radix block
int64_t hist_words(int64_t n_max) {
    const int64_t nb = n_max > 0 ? (n_max + kB * kMinSortItems - 1) / (kB * kMinSortItems) : 1;
    return nb * kBins;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

SortScratch carve(void* base, int64_t n_max, int items) {
    SortScratch s;
    const int64_t per_block = (int64_t)kB * items;
    s.nblocks = (n_max + per_block - 1) / per_block;
    if (s.nblocks < 1) s.nblocks = 1;
    s.hist_n = s.nblocks * kBins;
    s.scan_blocks = (s.hist_n + kElems - 1) / kElems;
    char* p = reinterpret_cast<char*>(base);
    const size_t arr = align256((size_t)(n_max > 0 ? n_max : 1) * sizeof(uint32_t));
    s.ka = reinterpret_cast<uint32_t*>(p); p += arr;
    s.va = reinterpret_cast<uint32_t*>(p); p += arr;
    s.kb = reinterpret_cast<uint32_t*>(p); p += arr;
    s.vb = reinterpret_cast<uint32_t*>(p); p += arr;
    s.hist = reinterpret_cast<uint32_t*>(p); p += align256((size_t)hist_words(n_max) * sizeof(uint32_t));
    s.sums = reinterpret_cast<uint32_t*>(p); p += align256((size_t)s.scan_blocks * sizeof(uint32_t));
    s.n_removed = reinterpret_cast<unsigned long long*>(p);
    return s;
}

size_t scratch_need(int64_t n_max) {
    const int64_t per_block = (int64_t)kB * kMinSortItems;  // the largest histogram any sort_items needs
    const int64_t nb = n_max > 0 ? (n_max + per_block - 1) / per_block : 1;
    const int64_t scan_blocks = (nb * kBins + kElems - 1) / kElems;
    return 4 * align256((size_t)(n_max > 0 ? n_max : 1) * 4) + align256((size_t)hist_words(n_max) * 4) +
           align256((size_t)scan_blocks * 4) + kRemSlots * sizeof(unsigned long long);
}

}  // namespace
}  // namespace oppx

using namespace oppx;

extern "C" {

size_t oppx_sort_scratch_bytes(int64_t n_max) { return n_max < 0 ? 0 : scratch_need(n_max); }

int oppx_sort_by_cell(int64_t n_max, int64_t stride, oppx_move_counters* counters, int64_t n_cells,
                      const int32_t* p2c, const oppx_dat_desc* dats, const oppx_dat_desc* out_dats, int n_dats,
                      int32_t* perm_out, int drop_removed, void* scratch, size_t scratch_bytes, void* stream) {
    OPPX_ARG_CHECK(counters && p2c && scratch, "oppx_sort_by_cell: null pointer");
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max, "oppx_sort_by_cell: need 0 <= n_max <= stride");
    OPPX_ARG_CHECK(n_cells > 0 && n_cells < kMaxCell, "oppx_sort_by_cell: n_cells out of range");
    OPPX_ARG_CHECK(n_max < (int64_t)0xFFFFFFFFll, "oppx_sort_by_cell: n_max beyond 32-bit indices");
    OPPX_ARG_CHECK(n_dats >= 0 && n_dats <= kMaxDats && (n_dats == 0 || (dats && out_dats)),
                   "oppx_sort_by_cell: 0..16 dats with output buffers");
    OPPX_ARG_CHECK(scratch_bytes >= scratch_need(n_max), "oppx_sort_by_cell: scratch too small");
    GatherTable T{};
    T.n = 0;
    T.stride = stride;
    int32_t* p2c_out = nullptr;  // the p2c dat itself is written from the keys by the last pass
    for (int q = 0; q < n_dats; q++) {
        OPPX_ARG_CHECK(dats[q].base && out_dats[q].base && dats[q].dim > 0 && dats[q].dim == out_dats[q].dim &&
                           dats[q].elem_bytes == out_dats[q].elem_bytes &&
                           (dats[q].elem_bytes == 4 || dats[q].elem_bytes == 8) && dats[q].base != out_dats[q].base,
                       "oppx_sort_by_cell: bad dat descriptor (dims/sizes must match, output must be a separate buffer)");
        if (dats[q].base == p2c && dats[q].dim == 1 && dats[q].elem_bytes == 4 && !p2c_out) {
            p2c_out = reinterpret_cast<int32_t*>(out_dats[q].base);
            continue;
        }
        T.in[T.n] = dats[q].base;
        T.out[T.n] = out_dats[q].base;
        T.dim[T.n] = dats[q].dim;
        T.eb[T.n] = dats[q].elem_bytes;
        T.n++;
    }
    hipStream_t s = as_stream(stream);
    const int items = g_tuning.sort_items;
    OPPX_ARG_CHECK(items == 4 || items == 8 || items == 16 || items == 32,
                   "oppx_sort_by_cell: tuning.sort_items must be 4, 8, 16 or 32");
    SortScratch sc = carve(scratch, n_max, items);
    int bits = 1;
    while (bits < 32 && ((int64_t)1 << bits) <= n_cells) bits++;  // keys are in [0, n_cells]
    const int passes = (bits + 7) / 8;
    zero_u64<<<1, kRemSlots, 0, s>>>(sc.n_removed);  // a kernel, not a memset node, so the sort can be captured
    OPPX_LAUNCH_CHECK("sort_reset");
    if (n_max > 0) {
        uint32_t *kin = sc.ka, *vin = sc.va, *kout = sc.kb, *vout = sc.vb;
        const uint32_t kc = (uint32_t)n_cells;
        // no dat left to gather: the last digit pass writes the permutation straight into perm_out
        const bool direct = T.n == 0 && perm_out != nullptr;
        for (int pass = 0; pass < passes; pass++) {
            const int shift = 8 * pass;
            const bool first = pass == 0;
            const bool last = pass == passes - 1;  // the keys of the last pass are not needed
            uint32_t* vdst = (last && direct) ? reinterpret_cast<uint32_t*>(perm_out) : vout;
            {
#define OPPX_HIST(I, F)                                                                              \
    sort_hist<I, F><<<stride_grid(reinterpret_cast<const void*>(sort_hist<I, F>), kB, sc.nblocks, kGridMult), kB, 0, \
                      s>>>(kin, p2c, kc, counters, n_max, shift, sc.hist, sc.n_removed)
                if (items == 4) { if (first) OPPX_HIST(4, true); else OPPX_HIST(4, false); }
                else if (items == 8) { if (first) OPPX_HIST(8, true); else OPPX_HIST(8, false); }
                else if (items == 16) { if (first) OPPX_HIST(16, true); else OPPX_HIST(16, false); }
                else { if (first) OPPX_HIST(32, true); else OPPX_HIST(32, false); }
#undef OPPX_HIST
                OPPX_LAUNCH_CHECK("sort_hist");
                const int64_t per_tile = (int64_t)kB * items;
                scan_reduce<<<(unsigned)sc.scan_blocks, kB, 0, s>>>(sc.hist, counters, n_max, per_tile, sc.sums);
                scan_sums<<<1, kB, 0, s>>>(sc.sums, counters, n_max, per_tile);
                scan_apply<<<(unsigned)sc.scan_blocks, kB, 0, s>>>(sc.hist, counters, n_max, per_tile, sc.sums);
                OPPX_LAUNCH_CHECK("sort_scan");
#define OPPX_SCATTER(I, LST, F)                                                                      \
    do {                                                                                              \
        auto kern = sort_scatter<I, LST, F>;                                                          \
        const unsigned res = resident_blocks(reinterpret_cast<const void*>(kern), kB);               \
        const unsigned g = (int64_t)res < sc.nblocks ? res : (unsigned)sc.nblocks;                    \
        kern<<<g, kB, 0, s>>>(kin, vin, p2c, kc, counters, n_max, shift, sc.hist, kout, vdst,            \
                              p2c_out);                                                              \
    } while (0)
#define OPPX_SCATTER_LF(I)                                                     \
    do {                                                                       \
        if (last) { if (first) OPPX_SCATTER(I, true, true); else OPPX_SCATTER(I, true, false); } \
        else { if (first) OPPX_SCATTER(I, false, true); else OPPX_SCATTER(I, false, false); }   \
    } while (0)
                if (items == 4) OPPX_SCATTER_LF(4);
                else if (items == 8) OPPX_SCATTER_LF(8);
                else if (items == 16) OPPX_SCATTER_LF(16);
                else OPPX_SCATTER_LF(32);
#undef OPPX_SCATTER_LF
#undef OPPX_SCATTER
                OPPX_LAUNCH_CHECK("sort_scatter");
            }
            uint32_t* t = kin; kin = kout; kout = t;
            t = vin; vin = vout; vout = t;
        }
        if (!direct && (T.n > 0 || perm_out)) {
            sort_gather<<<grid_for(n_max, kB), kB, 0, s>>>(vin, counters, n_max, T, perm_out);
            OPPX_LAUNCH_CHECK("sort_gather");
        }
    }
    sort_finalize<<<1, 1, 0, s>>>(counters, n_max, sc.n_removed, drop_removed);
    OPPX_LAUNCH_CHECK("sort_finalize");
    return OPPX_OK;
}

int oppx_gather_dats(int64_t n_max, int64_t stride, const oppx_move_counters* counters, const int32_t* perm,
                     const oppx_dat_desc* dats, const oppx_dat_desc* out_dats, int n_dats, void* stream) {
    OPPX_ARG_CHECK(counters && perm && n_dats > 0 && n_dats <= kMaxDats && dats && out_dats,
                   "oppx_gather_dats: bad arguments");
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max, "oppx_gather_dats: need 0 <= n_max <= stride");
    GatherTable T{};
    T.n = n_dats;
    T.stride = stride;
    for (int q = 0; q < n_dats; q++) {
        OPPX_ARG_CHECK(dats[q].base && out_dats[q].base && dats[q].dim > 0 && dats[q].dim == out_dats[q].dim &&
                           dats[q].elem_bytes == out_dats[q].elem_bytes &&
                           (dats[q].elem_bytes == 4 || dats[q].elem_bytes == 8) && dats[q].base != out_dats[q].base,
                       "oppx_gather_dats: bad dat descriptor");
        T.in[q] = dats[q].base;
        T.out[q] = out_dats[q].base;
        T.dim[q] = dats[q].dim;
        T.eb[q] = dats[q].elem_bytes;
    }
    if (n_max == 0) return OPPX_OK;
    sort_gather<<<grid_for(n_max, kB), kB, 0, as_stream(stream)>>>(reinterpret_cast<const uint32_t*>(perm),
                                                                   counters, n_max, T, nullptr);
    OPPX_LAUNCH_CHECK("gather_dats");
    return OPPX_OK;
}

int oppx_lower_bound(int64_t n_max, const oppx_move_counters* counters, const int32_t* p2c, const int32_t* keys,
                     int64_t k, int64_t* out, void* stream) {
    OPPX_ARG_CHECK(counters && p2c && keys && out && n_max >= 0 && k >= 0, "oppx_lower_bound: bad arguments");
    if (k == 0) return OPPX_OK;
    lower_bound_kernel<<<grid_for(k, kB), kB, 0, as_stream(stream)>>>(p2c, counters, n_max, keys, k, out);
    OPPX_LAUNCH_CHECK("lower_bound");
    return OPPX_OK;
}

}  // extern "C"
