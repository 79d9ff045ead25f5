// OPP_INC charge deposit in FP32 with compensated (Kahan) sums over cell-sorted particles.
//
// Reference semantics: deposit_charge_on_nodes (apps/app_fempic/kernels.h:164-175) through the
// double-indirect c2n[p2c] map; the reference's segmented-reduction variant sorts (node, value)
// pairs and reduces by key (apps_codegen/app_fempic_cg/hip/deposit_charge_on_nodes_kernel_loop.hpp:98-158,
// opp_lib/include/device_kernels/hip_inline_kernels.h:396-535). Here particles are already ordered
// by cell (oppx_sort_by_cell), so:
//   pass 1: one lane per run of equal cells sums lc[0..3] of the run in index order (FP32 Kahan);
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

__global__ void __launch_bounds__(kB) cell_runs(const int32_t* __restrict__ p2c, const double* __restrict__ lc,
                                                int64_t stride, const oppx_move_counters* ctr, int64_t n_max,
                                                int64_t n_cells, float* __restrict__ part) {
    int64_t n = ctr->size;
    n = n < 0 ? 0 : (n > n_max ? n_max : n);
    const int64_t j = (int64_t)blockIdx.x * kB + threadIdx.x;
    if (j >= n) return;
    const int32_t c = p2c[j];
    if (!((unsigned)c < (unsigned)n_cells)) return;
    if (j > 0 && p2c[j - 1] == c) return;  // not the head of its run
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

int oppx_fempic_deposit_sorted_f32(int64_t n_max, int64_t stride, oppx_move_counters* counters, const int32_t* p2c,
                                   const double* lc, int64_t n_cells, const int32_t* node_start, const int32_t* node_end,
                                   const int32_t* node_ref, int64_t n_nodes, float* cell_partials, double* node_charge,
                                   float* node_charge_f32, void* stream) {
    OPPX_ARG_CHECK(counters && p2c && lc && node_start && node_ref && cell_partials,
                   "oppx_fempic_deposit_sorted_f32: null pointer");
    OPPX_ARG_CHECK(node_charge || node_charge_f32, "oppx_fempic_deposit_sorted_f32: no output");
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max && n_cells > 0 && n_cells < kMaxCell && n_nodes > 0,
                   "oppx_fempic_deposit_sorted_f32: bad sizes");
    hipStream_t s = as_stream(stream);
    const int64_t np = 4 * n_cells;
    int64_t gz = grid_for(np, kB);
    if (gz > 4096) gz = 4096;
    zero_partials<<<(unsigned)gz, kB, 0, s>>>(cell_partials, np);
    OPPX_LAUNCH_CHECK("zero_partials");
    if (n_max > 0) {
        cell_runs<<<grid_for(n_max, kB), kB, 0, s>>>(p2c, lc, stride, counters, n_max, n_cells, cell_partials);
        OPPX_LAUNCH_CHECK("cell_runs");
    }
    node_gather<<<grid_for(n_nodes, kB), kB, 0, s>>>(node_start, node_end, node_ref, n_nodes, cell_partials,
                                                     node_charge, node_charge_f32);
    OPPX_LAUNCH_CHECK("node_gather");
    return OPPX_OK;
}

}  // extern "C"
