// Particle compaction on gfx950: hole-fill after a move.
//
// Reference semantics (opp_lib/src/core/opp_particle_organize_core.cpp:124-218; identical live range
// in omp opp_lib/src/omp/opp_particle_mover.cpp:153-218 and hip opp_lib/src/hip/opp_particle_sorter.cpp:271-359):
// with N particles, R removed and S = N - R, the k-th hole below S (ascending) receives the k-th
// live record at or above S (descending). No sort of hole indices is needed here: the move pass
// leaves a removal bitmap (one 64-bit ballot word per wave), and
//   #holes below index v        = wp[v/64] + popcount(word & low bits)      (wp = word prefix)
//   rank of live v >= S from top = (N-1-v) - (R - #holes below v)
// so every live tail record finds its destination hole by a binary search over wp and a
// select inside one word. Passes: block popcounts -> scan -> word prefix -> tail copy -> finalize.

#include <mutex>

#include "oppx_internal.h"

// Scratch of oppx_holefill_noalloc: a file-scope device array, one copy per GPU, held until the process
// ends. In host code this name stands for the array but is no device address: kernels receive the
// address hipGetSymbolAddress returns for the current device (oppx::noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[OPPX_HOLEFILL_NOALLOC_SCRATCH_BYTES];

namespace oppx {
namespace {

constexpr int kB = 256;
constexpr int kMaxDats = 16;

struct DatTable {
    int n;
    int64_t stride;
    void* base[kMaxDats];
    int dim[kMaxDats];
    int eb[kMaxDats];
};

__device__ __forceinline__ int64_t words_of(int64_t n) { return (n + 63) >> 6; }

// live count before compaction, clamped to the host's bound (grids and scratch are sized by it)
__device__ __forceinline__ int64_t live_count(const oppx_move_counters* ctr, int64_t n_max) {
    const int64_t n = ctr->size;
    return n < 0 ? 0 : (n > n_max ? n_max : n);
}

// exclusive block scan of one uint32 per thread (kB threads); returns the prefix, total in *tot
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* tot) {
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
    *tot = all;
    return off + inc - v;
}

__global__ void __launch_bounds__(kB) hf_block_sums(const uint64_t* __restrict__ bits,
                                                    const oppx_move_counters* __restrict__ ctr,
                                                    uint32_t* __restrict__ bsum, int64_t n_max) {
    const int64_t W = words_of(live_count(ctr, n_max));
    const int64_t w = (int64_t)blockIdx.x * kB + threadIdx.x;
    const uint32_t c = (w < W) ? (uint32_t)__popcll(bits[w]) : 0u;
    uint32_t tot;
    block_exclusive_scan(c, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block of kB threads: exclusive scan of nb block sums in place; total = removed count
__global__ void __launch_bounds__(kB) hf_scan_blocks(uint32_t* __restrict__ bsum, int64_t nb,
                                                     oppx_move_counters* __restrict__ ctr) {
    uint32_t carry = 0;
    for (int64_t base = 0; base < nb; base += kB) {
        const int64_t j = base + threadIdx.x;
        const uint32_t v = (j < nb) ? bsum[j] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, &tot);
        if (j < nb) bsum[j] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) ctr->removed = (int32_t)carry;
}

__global__ void __launch_bounds__(kB) hf_word_prefix(const uint64_t* __restrict__ bits,
                                                     const oppx_move_counters* __restrict__ ctr,
                                                     const uint32_t* __restrict__ bsum, uint32_t* __restrict__ wp,
                                                     int64_t n_max) {
    const int64_t W = words_of(live_count(ctr, n_max));
    const int64_t w = (int64_t)blockIdx.x * kB + threadIdx.x;
    const uint32_t c = (w < W) ? (uint32_t)__popcll(bits[w]) : 0u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_scan(c, &tot);
    if (w < W) wp[w] = bsum[blockIdx.x] + ex;
}

template <typename T>
__device__ __forceinline__ void move_record(void* base, int dim, int64_t stride, int64_t from, int64_t to) {
    T* p = reinterpret_cast<T*>(base);
    for (int d = 0; d < dim; d++) p[d * stride + to] = p[d * stride + from];
}

__global__ void __launch_bounds__(kB) hf_fill_tail(const uint64_t* __restrict__ bits,
                                                   const oppx_move_counters* __restrict__ ctr,
                                                   const uint32_t* __restrict__ wp, DatTable T, int64_t n_max) {
    const int64_t N = live_count(ctr, n_max), R = ctr->removed;
    if (R <= 0 || R > N) return;
    const int64_t S = N - R, W = words_of(N);
    for (int64_t v = S + (int64_t)blockIdx.x * kB + threadIdx.x; v < N; v += (int64_t)gridDim.x * kB) {
        const uint64_t word = bits[v >> 6];
        const int j = (int)(v & 63);
        if ((word >> j) & 1ull) continue;  // a hole itself
        const int64_t below = (int64_t)wp[v >> 6] + __popcll(word & ((1ull << j) - 1ull));
        const int64_t k = (N - 1 - v) - (R - below);
        // k-th hole: largest word w with wp[w] <= k, then the (k - wp[w])-th set bit of it
        int64_t lo = 0, hi = W - 1;
        while (lo < hi) {
            const int64_t mid = (lo + hi + 1) >> 1;
            if ((int64_t)wp[mid] <= k) lo = mid;
            else hi = mid - 1;
        }
        uint64_t hw = bits[lo];
        for (int64_t r = k - (int64_t)wp[lo]; r > 0 && hw; r--) hw &= hw - 1ull;
        if (hw == 0ull) continue;  // inconsistent bitmap: never write outside [0, S)
        const int64_t h = lo * 64 + (__ffsll((long long)hw) - 1);
        if (h < 0 || h >= S) continue;
        for (int q = 0; q < T.n; q++) {
            if (T.eb[q] == 8) move_record<double>(T.base[q], T.dim[q], T.stride, v, h);
            else move_record<int32_t>(T.base[q], T.dim[q], T.stride, v, h);
        }
    }
}

__global__ void hf_finalize(const uint64_t* __restrict__ bits, oppx_move_counters* __restrict__ ctr,
                            const uint32_t* __restrict__ wp, int64_t n_max) {
    const int64_t N = live_count(ctr, n_max), R = ctr->removed;
    if (R <= 0 || R > N) {
        ctr->holes_low = 0;
        return;
    }
    const int64_t S = N - R;
    const int j = (int)(S & 63);
    const uint64_t lowmask = j ? ((1ull << j) - 1ull) : 0ull;
    ctr->holes_low = (int32_t)(wp[S >> 6] + __popcll(bits[S >> 6] & lowmask));
    ctr->size = (int32_t)S;
}

struct Scratch {
    uint32_t* bsum;
    uint32_t* wp;
    int64_t nb;
};

// bsum (block counts, padded to 64 words), wp (one prefix per bitmap word) and 64 spare words
constexpr size_t scratch_need(int64_t n_max) {
    const int64_t W = (n_max + 63) / 64;
    const int64_t nb = (W + kB - 1) / kB < 1 ? 1 : (W + kB - 1) / kB;
    return (size_t)(((nb + 63) / 64) * 64 + W + 64) * sizeof(uint32_t);
}
static_assert(scratch_need(OPPX_HOLEFILL_NOALLOC_MAX_PARTICLES) <= OPPX_HOLEFILL_NOALLOC_SCRATCH_BYTES,
              "the static scratch must hold the hole fill of the largest set oppx_holefill_noalloc accepts");

Scratch carve(void* scratch, int64_t n_max) {
    const int64_t W = (n_max + 63) / 64;
    const int64_t nb = (W + kB - 1) / kB;
    Scratch s;
    s.nb = nb > 0 ? nb : 1;
    s.bsum = reinterpret_cast<uint32_t*>(scratch);
    s.wp = s.bsum + ((s.nb + 63) / 64) * 64;
    return s;
}

int holefill_impl(int64_t n_max, int64_t stride, oppx_move_counters* counters, const uint64_t* hole_bits,
                  const oppx_dat_desc* dats, int n_dats, void* work, hipStream_t s) {
    DatTable T{};
    T.n = n_dats;
    T.stride = stride;
    for (int q = 0; q < n_dats; q++) {
        OPPX_ARG_CHECK(dats[q].base && dats[q].dim > 0, "oppx_holefill: bad dat descriptor");
        OPPX_ARG_CHECK(dats[q].elem_bytes == 4 || dats[q].elem_bytes == 8, "oppx_holefill: elem_bytes 4 or 8");
        T.base[q] = dats[q].base;
        T.dim[q] = dats[q].dim;
        T.eb[q] = dats[q].elem_bytes;
    }
    // every scratch word a pass reads was written by an earlier pass of the same call: nothing to clear
    Scratch sc = carve(work, n_max);
    hf_block_sums<<<(unsigned)sc.nb, kB, 0, s>>>(hole_bits, counters, sc.bsum, n_max);
    OPPX_LAUNCH_CHECK("hf_block_sums");
    hf_scan_blocks<<<1, kB, 0, s>>>(sc.bsum, sc.nb, counters);
    OPPX_LAUNCH_CHECK("hf_scan_blocks");
    hf_word_prefix<<<(unsigned)sc.nb, kB, 0, s>>>(hole_bits, counters, sc.bsum, sc.wp, n_max);
    OPPX_LAUNCH_CHECK("hf_word_prefix");
    hf_fill_tail<<<1024, kB, 0, s>>>(hole_bits, counters, sc.wp, T, n_max);
    OPPX_LAUNCH_CHECK("hf_fill_tail");
    hf_finalize<<<1, 1, 0, s>>>(hole_bits, counters, sc.wp, n_max);
    OPPX_LAUNCH_CHECK("hf_finalize");
    return OPPX_OK;
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices];  // address of `scratch` on each device, once looked up

}  // namespace

// Device address of the static scratch array on `device`, which must be current: looked up with
// hipGetSymbolAddress on the first call for the device, then kept.
int noalloc_scratch(int device, void** out) {
    OPPX_ARG_CHECK(device >= 0 && device < kMaxDevices, "oppx: device index beyond the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device] == nullptr) {
        void* p = nullptr;
        OPPX_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        OPPX_ARG_CHECK(p != nullptr, "oppx: hipGetSymbolAddress returned no address for the static scratch");
        g_scratch_dev[device] = p;
    }
    if (out) *out = g_scratch_dev[device];
    return OPPX_OK;
}

}  // namespace oppx

using namespace oppx;

extern "C" {

// a move covers whole tiles: at most ceil(n_max/64) words plus one partial block of the largest tile
int64_t oppx_hole_bits_words(int64_t n_max) { return (n_max > 0 ? (n_max + 63) / 64 : 0) + 32; }

size_t oppx_holefill_scratch_bytes(int64_t n_max) { return n_max < 0 ? 0 : scratch_need(n_max); }

int oppx_holefill(int64_t n_max, int64_t stride, oppx_move_counters* counters, const uint64_t* hole_bits,
                  const oppx_dat_desc* dats, int n_dats, void* work, size_t work_bytes, void* stream) {
    OPPX_ARG_CHECK(counters && hole_bits && work, "oppx_holefill: null pointer");
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max, "oppx_holefill: need 0 <= n_max <= stride");
    OPPX_ARG_CHECK(n_dats >= 1 && n_dats <= kMaxDats && dats, "oppx_holefill: 1..16 dats required");
    OPPX_ARG_CHECK(work_bytes >= oppx_holefill_scratch_bytes(n_max), "oppx_holefill: scratch too small");
    return holefill_impl(n_max, stride, counters, hole_bits, dats, n_dats, work, as_stream(stream));
}

int oppx_holefill_noalloc(int64_t n_max, int64_t stride, oppx_move_counters* counters, const uint64_t* hole_bits,
                          const oppx_dat_desc* dats, int n_dats, void* stream) {
    OPPX_ARG_CHECK(counters && hole_bits, "oppx_holefill_noalloc: null pointer");
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max, "oppx_holefill_noalloc: need 0 <= n_max <= stride");
    OPPX_ARG_CHECK(n_max <= OPPX_HOLEFILL_NOALLOC_MAX_PARTICLES,
                   "oppx_holefill_noalloc: n_max exceeds OPPX_HOLEFILL_NOALLOC_MAX_PARTICLES (the static scratch "
                   "holds the hole fill of at most 134217728 particles); use oppx_holefill with a workspace");
    OPPX_ARG_CHECK(n_dats >= 1 && n_dats <= kMaxDats && dats, "oppx_holefill_noalloc: 1..16 dats required");
    int dev = 0;
    OPPX_HIP_TRY(hipGetDevice(&dev));
    void* work = nullptr;
    const int rc = noalloc_scratch(dev, &work);  // looked up by oppx_init, or here on a device's first call
    if (rc != OPPX_OK) return rc;
    return holefill_impl(n_max, stride, counters, hole_bits, dats, n_dats, work, as_stream(stream));
}

}  // extern "C"
