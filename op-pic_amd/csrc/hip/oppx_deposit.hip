// OPP_INC charge deposit in FP32 with compensated (Kahan) sums over cell-sorted particles.
//
// Reference semantics: deposit_charge_on_nodes (apps/app_fempic/kernels.h:164-175) through the
// double-indirect c2n[p2c] map; the reference's segmented-reduction variant sorts (node, value)
// pairs and reduces by key (apps_codegen/app_fempic_cg/hip/deposit_charge_on_nodes_kernel_loop.hpp:98-158,
// opp_lib/include/device_kernels/hip_inline_kernels.h:396-535). Here particles are already ordered
// by cell (oppx_sort_by_cell), so:
//   pass 1: the heads of the runs of equal cells are compacted (block counts, scan, scatter), then
//           one lane per run sums lc[0..3] of its run in index order (FP32 Kahan), from a tile of lc
//           staged in LDS with coalesced loads -- every lane of a wave owns a run instead of one lane
//           in ~(run length) holding it. A run longer than
//           kLongRun particles is summed by a whole wave instead: lane l takes particles
//           head + l, head + l + 64, ... (Kahan), and the 64 lane sums are added in lane order, so a
//           long run no longer holds its wave for its whole length. Either way the order depends only
//           on the run itself (not on where it sits in the particle set);
//   pass 2: one lane per node sums its (cell, vertex) partials in the fixed node->cell order
//           (FP32 Kahan) and stores the result.
// The order of every addition is fixed by the data, so the result does not depend on scheduling
// or on other work sharing the GPU.

#include "oppx_internal.h"

namespace oppx {
namespace {

constexpr int kB = 256;
constexpr unsigned kGridMult = 4;  // grids that stride over live work: this many times the resident blocks

struct Kahan {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

__global__ void __launch_bounds__(kB) zero_partials(float* __restrict__ part, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) part[i] = 0.0f;
}

constexpr int kRunItems = 8;                  // particles per thread in the head compaction
constexpr int kRunElems = kB * kRunItems;      // particles per tile

__device__ __forceinline__ int64_t live_of(const oppx_move_counters* ctr, int64_t n_max) {
    const int64_t n = ctr->size;
    return n < 0 ? 0 : (n > n_max ? n_max : n);
}

// A block's p2c tile (its kRunElems entries and the one before) staged in LDS with coalesced loads,
// one padding word per 8 so that thread t's 8 consecutive entries (from word 9t) sit in distinct banks.
struct RunTile {
    int32_t v[kRunElems + 1 + (kRunElems + 1) / 8 + 1];
};

__device__ __forceinline__ int tile_at(int e) { return e + (e >> 3); }

__device__ __forceinline__ void load_run_tile(RunTile& T, const int32_t* __restrict__ p2c, int64_t base0, int64_t n) {
    for (int t = threadIdx.x; t < kRunElems + 1; t += kB) {
        const int64_t j = base0 - 1 + t;
        T.v[tile_at(t)] = (j >= 0 && j < n) ? p2c[j] : -1;  // -1: never a valid cell
    }
    __syncthreads();
}

// element base0 + e of the tile (e in [0, kRunElems)) starts a run of a valid cell
__device__ __forceinline__ bool run_head(const RunTile& T, int e, int64_t n_cells) {
    const int32_t c = T.v[tile_at(e + 1)];
    return (unsigned)c < (unsigned)n_cells && T.v[tile_at(e)] != c;
}

// exclusive block scan of one uint32 per thread (NT threads); total in *tot
template <int NT = kB>
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* tot) {
    __shared__ uint32_t s_w[NT / kWave];
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
    for (int q = 0; q < NT / kWave; q++) {
        if (q < w) off += s_w[q];
        all += s_w[q];
    }
    __syncthreads();
    *tot = all;
    return off + inc - v;
}

// Also checks the precondition: a valid cell right after a larger cell or after an entry that is not a
// valid cell (a removed particle) means p2c is not sorted, and two runs of one cell would overwrite each
// other's partials; every such entry is counted in ctr->faults.
// The grid strides over the tiles below the live count.
__global__ void __launch_bounds__(kB) run_count(const int32_t* __restrict__ p2c, oppx_move_counters* ctr,
                                                int64_t n_max, int64_t n_cells, uint32_t* __restrict__ bsum) {
    __shared__ RunTile T;
    const int64_t n = live_of(ctr, n_max);
    const int64_t tiles = (n + kRunElems - 1) / kRunElems;
    const int e0 = (int)threadIdx.x * kRunItems;
    int unsorted = 0;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t base0 = t * kRunElems;
        load_run_tile(T, p2c, base0, n);
        uint32_t c = 0;
        for (int r = 0; r < kRunItems; r++) {
            const int64_t j = base0 + e0 + r;
            if (j >= n) break;
            if (run_head(T, e0 + r, n_cells)) {
                c++;
                const int32_t prev = T.v[tile_at(e0 + r)], cur = T.v[tile_at(e0 + r + 1)];
                if (j > 0 && (!((unsigned)prev < (unsigned)n_cells) || prev > cur)) unsorted++;
            }
        }
        uint32_t tot;
        block_scan(c, &tot);  // ends with a barrier: the tile may be overwritten after it
        if (threadIdx.x == 0) bsum[t] = tot;
    }
    unsorted = wave_sum_i32(unsorted);
    if ((threadIdx.x & (kWave - 1)) == 0 && unsorted) atomicAdd(&ctr->faults, unsorted);
}

// single block of kScanThreads: exclusive scan of the live tiles' counts in place; the total goes to
// n_runs[0], and the long-run count n_runs[1] starts at zero
constexpr int kScanThreads = 1024;
__global__ void __launch_bounds__(kScanThreads) run_scan(uint32_t* __restrict__ bsum, const oppx_move_counters* ctr,
                                                         int64_t n_max, uint32_t* __restrict__ n_runs) {
    if (threadIdx.x == 0) n_runs[1] = 0u;
    const int64_t nb = (live_of(ctr, n_max) + kRunElems - 1) / kRunElems;
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
        const int64_t j = b0 + threadIdx.x;
        const uint32_t v = j < nb ? bsum[j] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan<kScanThreads>(v, &tot);
        if (j < nb) bsum[j] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *n_runs = carry;
}

__global__ void __launch_bounds__(kB) run_compact(const int32_t* __restrict__ p2c, const oppx_move_counters* ctr,
                                                  int64_t n_max, int64_t n_cells, const uint32_t* __restrict__ bsum,
                                                  int32_t* __restrict__ heads) {
    __shared__ RunTile T;
    const int64_t n = live_of(ctr, n_max);
    const int64_t tiles = (n + kRunElems - 1) / kRunElems;
    const int e0 = (int)threadIdx.x * kRunItems;
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t base0 = t * kRunElems;
        load_run_tile(T, p2c, base0, n);
        uint32_t c = 0;
        for (int r = 0; r < kRunItems; r++)
            if (base0 + e0 + r < n && run_head(T, e0 + r, n_cells)) c++;
        uint32_t tot;
        uint32_t pos = bsum[t] + block_scan(c, &tot);
        for (int r = 0; r < kRunItems; r++)
            if (base0 + e0 + r < n && run_head(T, e0 + r, n_cells)) heads[pos++] = (int32_t)(base0 + e0 + r);
        __syncthreads();  // every thread is done with the tile before the next one is loaded
    }
}

constexpr int64_t kLongRun = 64;  // runs longer than this are summed by a whole wave

// End of run r (one past its last particle): the next head (valid cells sort before every invalid or
// removed entry, so runs are back to back), or for the last run the first index holding another cell.
__device__ __forceinline__ int64_t run_end(const int32_t* __restrict__ p2c, const int32_t* __restrict__ heads,
                                           int64_t r, int64_t n_runs, int64_t n, int32_t c) {
    if (r + 1 < n_runs) return heads[r + 1];
    int64_t e = heads[r];
    while (e < n && p2c[e] == c) e++;
    return e;
}

// lc index of the particle at sorted position q: perm[q] with PERM (lc left in the order before the
// sort), else q. An entry outside [0, stride) gives particle 0 (never outside the dat) and counts in bad.
template <bool PERM>
__device__ __forceinline__ int perm_at(const int32_t* __restrict__ perm, int64_t q, int64_t stride, int& bad) {
    if (!PERM) return (int)q;
    const int32_t p = perm[q];
    if ((uint64_t)(int64_t)p < (uint64_t)stride) return p;
    bad++;
    return 0;
}

// Kahan sums of lc[0..3] of particles j, j + step, ... below e, in that order, through perm_at. Four
// particles at a time: their indexes, then their 16 lc words, are loaded before any of them is added, so
// one lane keeps 16 loads in flight; the additions keep index order. Returns the bad permutation entries.
template <bool PERM>
__device__ __forceinline__ int run_sums(const double* __restrict__ lc, const int32_t* __restrict__ perm,
                                        int64_t stride, int64_t j, int64_t e, int64_t step, Kahan (&k)[4]) {
    int bad = 0;
    int64_t q = j;
    for (; q + 3 * step < e; q += 4 * step) {
        int64_t p[4];
#pragma unroll
        for (int u = 0; u < 4; u++) p[u] = PERM ? (int64_t)perm_at<PERM>(perm, q + u * step, stride, bad) : q + u * step;
        float v[4][4];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int d = 0; d < 4; d++) v[u][d] = (float)lc[d * stride + p[u]];
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int d = 0; d < 4; d++) k[d].add(v[u][d]);
    }
    for (; q < e; q += step) {
        const int64_t pq = PERM ? (int64_t)perm_at<PERM>(perm, q, stride, bad) : q;
#pragma unroll
        for (int d = 0; d < 4; d++) k[d].add((float)lc[d * stride + pq]);
    }
    return bad;
}

// One block per tile of kRunElems particles (the tiles of run_count): the lc of the tile's particles and
// of the kLongRun after it, in FP32, is staged in LDS with coalesced loads (with PERM, loads through the
// permutation), then one lane per run that starts in the tile (heads[bsum[t] .. bsum[t + 1])) sums the
// run from LDS in index order (Kahan); a run of at most kLongRun particles ends within kLongRun of the
// tile's end. Longer runs are listed for cell_runs_long. The sums are those of lane-per-run loops over
// lc in global memory, bit for bit. lc_f32 (nullable) receives the staged FP32 values of every live
// particle in sorted order (16 B each, coalesced), from which cell_runs_long then reads instead of
// gathering through the permutation a second time. The grid strides over the live tiles.
constexpr int kLcThreads = 512;
struct LcTile {
    float v[4][kRunElems + kLongRun];
};

template <bool PERM>
__global__ void __launch_bounds__(kLcThreads) cell_runs(const int32_t* __restrict__ p2c, const double* __restrict__ lc,
                                                        const int32_t* __restrict__ perm, int64_t stride,
                                                        oppx_move_counters* ctr, int64_t n_max,
                                                        const int32_t* __restrict__ heads,
                                                        const uint32_t* __restrict__ bsum, uint32_t* __restrict__ n_runs,
                                                        int32_t* __restrict__ long_runs, float* __restrict__ part,
                                                        float4* __restrict__ lc_f32) {
    __shared__ LcTile T;
    const int64_t n = live_of(ctr, n_max);
    const int64_t tiles = (n + kRunElems - 1) / kRunElems;
    const int64_t nr = (int64_t)n_runs[0];
    int bad = 0, bad_over = 0;  // bad permutation entries of the tile's own particles / of the overlap
    for (int64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const int64_t base = t * kRunElems;
        for (int e = threadIdx.x; e < kRunElems + kLongRun; e += kLcThreads) {
            const int64_t q = base + e;
            float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, v3 = 0.0f;
            if (q < n) {
                const int64_t p = PERM ? (int64_t)perm_at<PERM>(perm, q, stride, e < kRunElems ? bad : bad_over) : q;
                v0 = (float)lc[p];
                v1 = (float)lc[stride + p];
                v2 = (float)lc[2 * stride + p];
                v3 = (float)lc[3 * stride + p];
            }
            T.v[0][e] = v0;
            T.v[1][e] = v1;
            T.v[2][e] = v2;
            T.v[3][e] = v3;
            if (lc_f32 && e < kRunElems && q < n) lc_f32[q] = make_float4(v0, v1, v2, v3);  // sorted FP32 copy
        }
        __syncthreads();
        const int64_t r_end = t + 1 < tiles ? (int64_t)bsum[t + 1] : nr;
        for (int64_t r = (int64_t)bsum[t] + threadIdx.x; r < r_end; r += kLcThreads) {
            const int64_t j = heads[r];
            const int32_t c = p2c[j];
            const int64_t e = run_end(p2c, heads, r, nr, n, c);
            if (e - j > kLongRun) {  // summed by cell_runs_long
                long_runs[atomicAdd(n_runs + 1, 1u)] = (int32_t)r;
                continue;
            }
            Kahan k0, k1, k2, k3;
            for (int q = (int)(j - base); q < (int)(e - base); q++) {
                k0.add(T.v[0][q]);
                k1.add(T.v[1][q]);
                k2.add(T.v[2][q]);
                k3.add(T.v[3][q]);
            }
            float4 v;
            v.x = k0.s; v.y = k1.s; v.z = k2.s; v.w = k3.s;
            *reinterpret_cast<float4*>(part + 4 * (int64_t)c) = v;
        }
        __syncthreads();  // every lane is done with the tile before the next one is staged
    }
    if (PERM && bad) atomicAdd(&ctr->faults, bad);
}

// One wave per long run (waves stride over the list): lane l sums particles head + l + 64 t in index
// order (Kahan), then lane 0 adds the 64 lane sums in lane order (Kahan).
template <bool PERM, bool F32>
__global__ void __launch_bounds__(kB) cell_runs_long(const int32_t* __restrict__ p2c, const double* __restrict__ lc,
                                                     const int32_t* __restrict__ perm, const float4* __restrict__ lc_f32,
                                                     int64_t stride, const oppx_move_counters* ctr, int64_t n_max,
                                                     const int32_t* __restrict__ heads,
                                                     const uint32_t* __restrict__ n_runs,
                                                     const int32_t* __restrict__ long_runs, float* __restrict__ part) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave = ((int64_t)blockIdx.x * kB + threadIdx.x) / kWave;
    const int64_t n_waves = (int64_t)gridDim.x * (kB / kWave);
    const int64_t nr = (int64_t)n_runs[0], nl = (int64_t)n_runs[1];
    const int64_t n = live_of(ctr, n_max);
    for (int64_t k = wave; k < nl; k += n_waves) {
        const int64_t r = long_runs[k];
        const int64_t j = heads[r];
        const int32_t c = p2c[j];
        const int64_t e = run_end(p2c, heads, r, nr, n, c);
        Kahan ks[4];
        if constexpr (F32) {  // the sorted FP32 copy: the same values, read in the same order
            for (int64_t q = j + lane; q < e; q += kWave) {
                const float4 v = lc_f32[q];
                ks[0].add(v.x);
                ks[1].add(v.y);
                ks[2].add(v.z);
                ks[3].add(v.w);
            }
        } else {
            run_sums<PERM>(lc, perm, stride, j + lane, e, kWave, ks);  // bad entries are counted by cell_runs
        }
        Kahan t0, t1, t2, t3;
        for (int l = 0; l < kWave; l++) {  // the same on every lane: lane l's sums, in lane order
            t0.add(__shfl(ks[0].s, l));
            t1.add(__shfl(ks[1].s, l));
            t2.add(__shfl(ks[2].s, l));
            t3.add(__shfl(ks[3].s, l));
        }
        if (lane == 0) {
            float4 v;
            v.x = t0.s; v.y = t1.s; v.z = t2.s; v.w = t3.s;
            *reinterpret_cast<float4*>(part + 4 * (int64_t)c) = v;
        }
    }
}

struct RunScratch {
    uint32_t* bsum;
    uint32_t* n_runs;  // [0] runs, [1] long runs
    int32_t* heads;
    int32_t* long_runs;
    int64_t nb;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t run_scratch_need(int64_t n_max) {
    const int64_t nb = n_max > 0 ? (n_max + kRunElems - 1) / kRunElems : 1;
    return align256((size_t)nb * 4) + 256 + 2 * align256((size_t)(n_max > 0 ? n_max : 1) * 4);
}

RunScratch run_carve(void* base, int64_t n_max) {
    RunScratch r;
    r.nb = n_max > 0 ? (n_max + kRunElems - 1) / kRunElems : 1;
    char* p = reinterpret_cast<char*>(base);
    r.bsum = reinterpret_cast<uint32_t*>(p);
    p += align256((size_t)r.nb * 4);
    r.n_runs = reinterpret_cast<uint32_t*>(p);
    p += 256;
    r.heads = reinterpret_cast<int32_t*>(p);
    p += align256((size_t)(n_max > 0 ? n_max : 1) * 4);
    r.long_runs = reinterpret_cast<int32_t*>(p);
    return r;
}

__global__ void __launch_bounds__(kB) node_gather(const int32_t* __restrict__ node_start,
                                                  const int32_t* __restrict__ node_end,
                                                  const int32_t* __restrict__ node_ref, int64_t n_nodes,
                                                  const float* __restrict__ part, double* __restrict__ charge,
                                                  float* __restrict__ charge_f32) {
    const int64_t v = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (v >= n_nodes) return;
    Kahan k;
    const int32_t e1 = node_end ? node_end[v] : node_start[v + 1];
    int32_t e = node_start[v];
    for (; e + 4 <= e1; e += 4) {  // four loads in flight, then the four adds in order
        const int32_t r0 = node_ref[e], r1 = node_ref[e + 1], r2 = node_ref[e + 2], r3 = node_ref[e + 3];
        const float p0 = part[r0], p1 = part[r1], p2 = part[r2], p3 = part[r3];
        k.add(p0);
        k.add(p1);
        k.add(p2);
        k.add(p3);
    }
    for (; e < e1; e++) k.add(part[node_ref[e]]);
    if (charge) charge[v] = (double)k.s;
    if (charge_f32) charge_f32[v] = k.s;
}

// The deposit over cell-sorted p2c, with lc in the same order (perm null) or in the order before the sort
// (perm[q]: the source index of sorted slot q).
int deposit_f32(const char* name, int64_t n_max, int64_t stride, oppx_move_counters* counters, const int32_t* p2c,
                const double* lc, const int32_t* perm, float* lc_f32, int64_t n_cells, const int32_t* node_start,
                const int32_t* node_end, const int32_t* node_ref, int64_t n_nodes, float* cell_partials,
                double* node_charge, float* node_charge_f32, void* scratch, size_t scratch_bytes, void* stream) {
    const std::string fn(name);
    OPPX_ARG_CHECK(counters && p2c && lc && node_start && node_ref && cell_partials && scratch,
                   fn + ": null pointer");
    OPPX_ARG_CHECK(node_charge || node_charge_f32, fn + ": no output");
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max && n_cells > 0 && n_cells < kMaxCell && n_nodes > 0,
                   fn + ": bad sizes");
    OPPX_ARG_CHECK(n_max < (int64_t)0x7FFFFFFF, fn + ": n_max beyond 32-bit indices");
    OPPX_ARG_CHECK(scratch_bytes >= run_scratch_need(n_max), fn + ": scratch too small");
    hipStream_t s = as_stream(stream);
    const int64_t np = 4 * n_cells;
    int64_t gz = grid_for(np, kB);
    if (gz > 4096) gz = 4096;
    zero_partials<<<(unsigned)gz, kB, 0, s>>>(cell_partials, np);
    OPPX_LAUNCH_CHECK("zero_partials");
    if (n_max > 0) {
        RunScratch rs = run_carve(scratch, n_max);
        // grids that stride over the live tiles
        run_count<<<stride_grid(reinterpret_cast<const void*>(run_count), kB, rs.nb, kGridMult), kB, 0, s>>>(
            p2c, counters, n_max, n_cells, rs.bsum);
        run_scan<<<1, kScanThreads, 0, s>>>(rs.bsum, counters, n_max, rs.n_runs);
        run_compact<<<stride_grid(reinterpret_cast<const void*>(run_compact), kB, rs.nb, kGridMult), kB, 0, s>>>(
            p2c, counters, n_max, n_cells, rs.bsum, rs.heads);
        int64_t gl = grid_for(n_max, kLongRun * (kB / kWave));  // at most one long run per kLongRun particles
        if (gl > 2048) gl = 2048;
        float4* f32 = reinterpret_cast<float4*>(lc_f32);
        if (perm) {
            auto kern = cell_runs<true>;
            kern<<<stride_grid(reinterpret_cast<const void*>(kern), kLcThreads, rs.nb, kGridMult), kLcThreads, 0, s>>>(
                p2c, lc, perm, stride, counters, n_max, rs.heads, rs.bsum, rs.n_runs, rs.long_runs, cell_partials, f32);
            if (f32)
                cell_runs_long<true, true><<<(unsigned)gl, kB, 0, s>>>(p2c, lc, perm, f32, stride, counters, n_max,
                                                                        rs.heads, rs.n_runs, rs.long_runs,
                                                                        cell_partials);
            else
                cell_runs_long<true, false><<<(unsigned)gl, kB, 0, s>>>(p2c, lc, perm, nullptr, stride, counters,
                                                                         n_max, rs.heads, rs.n_runs, rs.long_runs,
                                                                         cell_partials);
        } else {
            auto kern = cell_runs<false>;
            kern<<<stride_grid(reinterpret_cast<const void*>(kern), kLcThreads, rs.nb, kGridMult), kLcThreads, 0, s>>>(
                p2c, lc, nullptr, stride, counters, n_max, rs.heads, rs.bsum, rs.n_runs, rs.long_runs, cell_partials,
                nullptr);
            cell_runs_long<false, false><<<(unsigned)gl, kB, 0, s>>>(p2c, lc, nullptr, nullptr, stride, counters,
                                                                      n_max, rs.heads, rs.n_runs, rs.long_runs,
                                                                      cell_partials);
        }
        OPPX_LAUNCH_CHECK("cell_runs");
    }
    node_gather<<<grid_for(n_nodes, kB), kB, 0, s>>>(node_start, node_end, node_ref, n_nodes, cell_partials,
                                                     node_charge, node_charge_f32);
    OPPX_LAUNCH_CHECK("node_gather");
    return OPPX_OK;
}

}  // namespace
}  // namespace oppx

using namespace oppx;

extern "C" {

size_t oppx_fempic_deposit_scratch_bytes(int64_t n_max) { return n_max < 0 ? 0 : run_scratch_need(n_max); }

int oppx_fempic_deposit_sorted_f32(int64_t n_max, int64_t stride, oppx_move_counters* counters, const int32_t* p2c,
                                   const double* lc, int64_t n_cells, const int32_t* node_start, const int32_t* node_end,
                                   const int32_t* node_ref, int64_t n_nodes, float* cell_partials, double* node_charge,
                                   float* node_charge_f32, void* scratch, size_t scratch_bytes, void* stream) {
    return deposit_f32("oppx_fempic_deposit_sorted_f32", n_max, stride, counters, p2c, lc, nullptr, nullptr, n_cells,
                       node_start, node_end, node_ref, n_nodes, cell_partials, node_charge, node_charge_f32, scratch,
                       scratch_bytes, stream);
}

int oppx_fempic_deposit_sorted_f32_perm(int64_t n_max, int64_t stride, oppx_move_counters* counters,
                                        const int32_t* p2c, const double* lc, const int32_t* perm, float* lc_f32,
                                        int64_t n_cells, const int32_t* node_start, const int32_t* node_end,
                                        const int32_t* node_ref, int64_t n_nodes, float* cell_partials,
                                        double* node_charge, float* node_charge_f32, void* scratch,
                                        size_t scratch_bytes, void* stream) {
    if (!perm) {
        set_error("oppx_fempic_deposit_sorted_f32_perm: null perm");
        return OPPX_ERR_ARG;
    }
    if (lc_f32 && (reinterpret_cast<uintptr_t>(lc_f32) & 15)) {
        set_error("oppx_fempic_deposit_sorted_f32_perm: lc_f32 not 16-byte aligned");
        return OPPX_ERR_ARG;
    }
    return deposit_f32("oppx_fempic_deposit_sorted_f32_perm", n_max, stride, counters, p2c, lc, perm, lc_f32, n_cells,
                       node_start, node_end, node_ref, n_nodes, cell_partials, node_charge, node_charge_f32, scratch,
                       scratch_bytes, stream);
}

}  // extern "C"
