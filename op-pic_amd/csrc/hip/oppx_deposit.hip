// OPP_INC charge deposit in FP32 with compensated (Kahan) sums over cell-sorted particles.
//
// Reference semantics: deposit_charge_on_nodes (apps/app_fempic/kernels.h:164-175) through the
// double-indirect c2n[p2c] map; the reference's segmented-reduction variant sorts (node, value)
// pairs and reduces by key (apps_codegen/app_fempic_cg/hip/deposit_charge_on_nodes_kernel_loop.hpp:98-158,
// opp_lib/include/device_kernels/hip_inline_kernels.h:396-535). Here particles are already ordered
// by cell (oppx_sort_by_cell), so:
//   pass 1: the heads of the runs of equal cells are compacted (block counts, scan, scatter), then
//           one lane per run sums lc[0..3] of its run in index order (FP32 Kahan) -- every lane of a
//           wave owns a run instead of one lane in ~(run length) holding it;
//   pass 2: one lane per node sums its (cell, vertex) partials in the fixed node->cell order
//           (FP32 Kahan) and stores the result.
// The order of every addition is fixed by the data, so the result does not depend on scheduling
// or on other work sharing the GPU.

#include "oppx_internal.h"

namespace oppx {
namespace {

constexpr int kB = 256;

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

constexpr int kRunItems = 16;                 // particles per thread in the head compaction
constexpr int kRunElems = kB * kRunItems;      // particles per block

__device__ __forceinline__ int64_t live_of(const oppx_move_counters* ctr, int64_t n_max) {
    const int64_t n = ctr->size;
    return n < 0 ? 0 : (n > n_max ? n_max : n);
}

__device__ __forceinline__ bool run_head(const int32_t* __restrict__ p2c, int64_t j, int64_t n_cells) {
    const int32_t c = p2c[j];
    return (unsigned)c < (unsigned)n_cells && (j == 0 || p2c[j - 1] != c);
}

// exclusive block scan of one uint32 per thread (kB threads); total in *tot
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t* tot) {
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

__global__ void __launch_bounds__(kB) run_count(const int32_t* __restrict__ p2c, const oppx_move_counters* ctr,
                                                int64_t n_max, int64_t n_cells, uint32_t* __restrict__ bsum) {
    const int64_t n = live_of(ctr, n_max);
    const int64_t base = (int64_t)blockIdx.x * kRunElems + (int64_t)threadIdx.x * kRunItems;
    uint32_t c = 0;
    for (int r = 0; r < kRunItems; r++)
        if (base + r < n && run_head(p2c, base + r, n_cells)) c++;
    uint32_t tot;
    block_scan(c, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of nb block counts in place; the total goes to *n_runs
__global__ void __launch_bounds__(kB) run_scan(uint32_t* __restrict__ bsum, int64_t nb, uint32_t* __restrict__ n_runs) {
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kB) {
        const int64_t j = b0 + threadIdx.x;
        const uint32_t v = j < nb ? bsum[j] : 0u;
        uint32_t tot;
        const uint32_t ex = block_scan(v, &tot);
        if (j < nb) bsum[j] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *n_runs = carry;
}

__global__ void __launch_bounds__(kB) run_compact(const int32_t* __restrict__ p2c, const oppx_move_counters* ctr,
                                                  int64_t n_max, int64_t n_cells, const uint32_t* __restrict__ bsum,
                                                  int32_t* __restrict__ heads) {
    const int64_t n = live_of(ctr, n_max);
    const int64_t base = (int64_t)blockIdx.x * kRunElems + (int64_t)threadIdx.x * kRunItems;
    uint32_t c = 0;
    for (int r = 0; r < kRunItems; r++)
        if (base + r < n && run_head(p2c, base + r, n_cells)) c++;
    uint32_t tot;
    uint32_t pos = bsum[blockIdx.x] + block_scan(c, &tot);
    for (int r = 0; r < kRunItems; r++)
        if (base + r < n && run_head(p2c, base + r, n_cells)) heads[pos++] = (int32_t)(base + r);
}

__global__ void __launch_bounds__(kB) cell_runs(const int32_t* __restrict__ p2c, const double* __restrict__ lc,
                                                int64_t stride, const oppx_move_counters* ctr, int64_t n_max,
                                                const int32_t* __restrict__ heads, const uint32_t* __restrict__ n_runs,
                                                float* __restrict__ part) {
    const int64_t r = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (r >= (int64_t)*n_runs) return;
    const int64_t n = live_of(ctr, n_max);
    const int64_t j = heads[r];
    const int32_t c = p2c[j];
    Kahan k0, k1, k2, k3;
    for (int64_t q = j; q < n && p2c[q] == c; q++) {
        k0.add((float)lc[q]);
        k1.add((float)lc[stride + q]);
        k2.add((float)lc[2 * stride + q]);
        k3.add((float)lc[3 * stride + q]);
    }
    float4 v;
    v.x = k0.s; v.y = k1.s; v.z = k2.s; v.w = k3.s;
    *reinterpret_cast<float4*>(part + 4 * (int64_t)c) = v;
}

struct RunScratch {
    uint32_t* bsum;
    uint32_t* n_runs;
    int32_t* heads;
    int64_t nb;
};

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t run_scratch_need(int64_t n_max) {
    const int64_t nb = n_max > 0 ? (n_max + kRunElems - 1) / kRunElems : 1;
    return align256((size_t)nb * 4) + 256 + align256((size_t)(n_max > 0 ? n_max : 1) * 4);
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
    for (int32_t e = node_start[v]; e < e1; e++) k.add(part[node_ref[e]]);
    if (charge) charge[v] = (double)k.s;
    if (charge_f32) charge_f32[v] = k.s;
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
    OPPX_ARG_CHECK(counters && p2c && lc && node_start && node_ref && cell_partials && scratch,
                   "oppx_fempic_deposit_sorted_f32: null pointer");
    OPPX_ARG_CHECK(node_charge || node_charge_f32, "oppx_fempic_deposit_sorted_f32: no output");
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max && n_cells > 0 && n_cells < kMaxCell && n_nodes > 0,
                   "oppx_fempic_deposit_sorted_f32: bad sizes");
    OPPX_ARG_CHECK(n_max < (int64_t)0x7FFFFFFF, "oppx_fempic_deposit_sorted_f32: n_max beyond 32-bit indices");
    OPPX_ARG_CHECK(scratch_bytes >= run_scratch_need(n_max), "oppx_fempic_deposit_sorted_f32: scratch too small");
    hipStream_t s = as_stream(stream);
    const int64_t np = 4 * n_cells;
    int64_t gz = grid_for(np, kB);
    if (gz > 4096) gz = 4096;
    zero_partials<<<(unsigned)gz, kB, 0, s>>>(cell_partials, np);
    OPPX_LAUNCH_CHECK("zero_partials");
    if (n_max > 0) {
        RunScratch rs = run_carve(scratch, n_max);
        run_count<<<(unsigned)rs.nb, kB, 0, s>>>(p2c, counters, n_max, n_cells, rs.bsum);
        run_scan<<<1, kB, 0, s>>>(rs.bsum, rs.nb, rs.n_runs);
        run_compact<<<(unsigned)rs.nb, kB, 0, s>>>(p2c, counters, n_max, n_cells, rs.bsum, rs.heads);
        cell_runs<<<grid_for(n_max, kB), kB, 0, s>>>(p2c, lc, stride, counters, n_max, rs.heads, rs.n_runs,
                                                     cell_partials);
        OPPX_LAUNCH_CHECK("cell_runs");
    }
    node_gather<<<grid_for(n_nodes, kB), kB, 0, s>>>(node_start, node_end, node_ref, n_nodes, cell_partials,
                                                     node_charge, node_charge_f32);
    OPPX_LAUNCH_CHECK("node_gather");
    return OPPX_OK;
}

}  // extern "C"
