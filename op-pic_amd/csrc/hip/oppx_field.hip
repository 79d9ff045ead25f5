// FEM-PIC loops on either side of the field solve, and ion injection, on gfx950: inject_ions,
// compute_node_charge_density, compute_electric_field, init_boundary_pot and the two max
// reductions of the main loop (apps/app_fempic/kernels.h:48-90,178-224, loops
// apps_codegen/app_fempic_cg/{seq,hip}/*_kernel_loop.hpp). They keep E and the charge density
// resident on the device around the host field solve.
//
// The reference's GCC -O3 build contracts `a*u + b*v + n` into fma(a, u, b*v) + n, the injected
// velocity `n*iv - (q/m)*E*(dt/2)` into fma(-((q/m)*E), dt/2, n*iv) (n*iv also feeds a store, so it
// stays a product), and the four-term sum of compute_electric_field into
// fma(sd3, p3, fma(sd2, p2, fma(sd0, p0, sd1*p1))); the kernels write those FMAs explicitly and the
// file is built with -ffp-contract=off.

#include "oppx_internal.h"

namespace oppx {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxItems = 8;  // consecutive values one lane folds in the max reduction

__device__ __forceinline__ void count_fault(oppx_move_counters* ctr) {
    if (ctr) atomicAdd(&ctr->faults, 1);
}

// inject_ions_kernel (apps/app_fempic/kernels.h:63-90) over the injected particles
// [first, first + count): their p2c holds the inlet face (opp_inc_part_count_with_distribution);
// random pair k is the k-th injected particle's (the dat is read from its beginning).
__global__ void inject_kernel(int64_t first, int64_t count, int64_t stride, double* __restrict__ pos,
                              double* __restrict__ vel, int32_t* __restrict__ p2c, oppx_inlet in, int64_t n_cells,
                              const double* __restrict__ ef, const double* __restrict__ rnd, int64_t rstride,
                              double iv, double qm, double hdt, oppx_move_counters* ctr) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int64_t i = first + k;
    const int face = p2c[i];
    if (!((uint64_t)face < (uint64_t)in.n_faces)) {
        count_fault(ctr);
        return;
    }
    const int cell = in.face_cell[face];
    if (!((uint64_t)cell < (uint64_t)n_cells)) {
        count_fault(ctr);
        return;
    }
    double a = rnd[k], b = rnd[rstride + k];
    if ((a + b) > 1) {
        a = (1 - a);
        b = (1 - b);
    }
    const double* u = in.u + 3 * (int64_t)face;
    const double* v = in.v + 3 * (int64_t)face;
    const double* nrm = in.normal + 3 * (int64_t)face;
    const double* o = in.node_pos + 9 * (int64_t)face;
    const double* e = ef + 3 * (int64_t)cell;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        pos[d * stride + i] = fma(a, u[d], b * v[d]) + o[d];
        vel[d * stride + i] = fma(-(qm * e[d]), hdt, nrm[d] * iv);
    }
    p2c[i] = cell;
}

// compute_node_charge_density_kernel (kernels.h:178-183)
__global__ void node_density_kernel(int64_t n, double* __restrict__ q, const double* __restrict__ vol, double spwt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = q[i] * (spwt / vol[i]);
}

// compute_electric_field_kernel (kernels.h:186-202): E_c -= sum_k sd[k] * phi(c2n[c][k]) per axis
__global__ void efield_kernel(int64_t n_cells, double* __restrict__ ef, const double* __restrict__ sd,
                              const int32_t* __restrict__ c2n, const double* __restrict__ pot, int64_t n_nodes,
                              oppx_move_counters* ctr) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cells) return;
    const int4 m = *reinterpret_cast<const int4*>(c2n + 4 * c);
    if (!((uint64_t)m.x < (uint64_t)n_nodes && (uint64_t)m.y < (uint64_t)n_nodes &&
          (uint64_t)m.z < (uint64_t)n_nodes && (uint64_t)m.w < (uint64_t)n_nodes)) {
        count_fault(ctr);
        return;
    }
    const double p0 = pot[m.x], p1 = pot[m.y], p2 = pot[m.z], p3 = pot[m.w];
    const double* s = sd + 12 * c;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double t = fma(s[d], p0, s[3 + d] * p1);
        t = fma(s[6 + d], p2, t);
        t = fma(s[9 + d], p3, t);
        ef[3 * c + d] = ef[3 * c + d] - t;
    }
}

// init_boundary_pot_kernel (kernels.h:48-59): FIXED (3) nodes hold -1 * wall potential, the rest 0
__global__ void boundary_pot_kernel(int64_t n, const int32_t* __restrict__ type, double wall, double* __restrict__ bnd) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bnd[i] = type[i] == 3 ? -1 * wall : 0;
}

// Max of the reference's OPP_MAX global, folded in index order from the initial value:
// m = init; for v: if (v > m) m = v (the MAX macro, opp_lib/include/opp_util.h:45). The result is
// the first value in that order that attains the maximum (a NaN never replaces m), so ranges of
// consecutive values combine exactly: the later range's candidate wins only when strictly greater.
struct MaxCand {
    double v;
    int valid;  // 0: the range held only NaNs
};

__device__ __forceinline__ MaxCand max_take(MaxCand a, MaxCand b) {  // a's range precedes b's
    if (!b.valid) return a;
    if (!a.valid) return b;
    return b.v > a.v ? b : a;
}

__device__ __forceinline__ MaxCand max_from_lane(MaxCand c, int src) {
    MaxCand o;
    o.v = __shfl(c.v, src);
    o.valid = __shfl(c.valid, src);
    return o;
}

// Block b folds values [b * BS * kMaxItems, (b + 1) * BS * kMaxItems): lane t its kMaxItems
// consecutive values, then lanes and waves in index order.
__global__ void __launch_bounds__(kBlock) max_partial_kernel(int64_t n, const double* __restrict__ x, int abs_values,
                                                             MaxCand* __restrict__ part) {
    const int64_t base = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * kMaxItems;
    MaxCand c{0.0, 0};
#pragma unroll
    for (int q = 0; q < kMaxItems; q++) {
        const int64_t i = base + q;
        if (i < n) {
            double v = x[i];
            if (abs_values) v = fabs(v);
            if (v == v) c = max_take(c, MaxCand{v, 1});
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    for (int off = 1; off < kWave; off <<= 1) {  // lane l takes lane l + off's later range
        const MaxCand o = max_from_lane(c, (lane + off) & (kWave - 1));
        if ((lane & (2 * off - 1)) == 0) c = max_take(c, o);
    }
    __shared__ MaxCand s[kBlock / kWave];
    if (lane == 0) s[threadIdx.x / kWave] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        MaxCand r = s[0];
        for (int w = 1; w < kBlock / kWave; w++) r = max_take(r, s[w]);
        part[blockIdx.x] = r;
    }
}

// One block folds the partials in order, then the initial value goes first.
__global__ void __launch_bounds__(kBlock) max_final_kernel(int64_t n_part, const MaxCand* __restrict__ part,
                                                           double init, double* __restrict__ out) {
    __shared__ MaxCand s[kBlock];
    const int64_t per = (n_part + kBlock - 1) / kBlock;
    MaxCand c{0.0, 0};
    for (int64_t q = 0; q < per; q++) {
        const int64_t i = (int64_t)threadIdx.x * per + q;
        if (i < n_part) c = max_take(c, part[i]);
    }
    s[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        MaxCand r{init, 1};
        for (int t = 0; t < kBlock; t++) r = max_take(r, s[t]);
        *out = r.v;
    }
}

int64_t max_blocks(int64_t n) { return (n + (int64_t)kBlock * kMaxItems - 1) / ((int64_t)kBlock * kMaxItems); }

}  // namespace
}  // namespace oppx

using namespace oppx;

extern "C" {

int oppx_fempic_inject(int64_t first, int64_t count, const oppx_fempic_parts* parts, const oppx_inlet* inlet,
                       int64_t n_cells, const double* c_ef, const double* rand, int64_t rand_stride,
                       double ion_velocity, double charge, double mass, double dt, oppx_move_counters* counters,
                       void* stream) {
    OPPX_ARG_CHECK(parts && inlet && parts->pos && parts->vel && parts->p2c, "oppx_fempic_inject: null pointer");
    OPPX_ARG_CHECK(first >= 0 && count >= 0 && first + count <= parts->stride,
                   "oppx_fempic_inject: [first, first + count) beyond capacity");
    if (count == 0) return OPPX_OK;
    OPPX_ARG_CHECK(inlet->n_faces > 0 && inlet->face_cell && inlet->u && inlet->v && inlet->normal && inlet->node_pos,
                   "oppx_fempic_inject: incomplete inlet");
    OPPX_ARG_CHECK(c_ef && rand && rand_stride >= count && n_cells > 0, "oppx_fempic_inject: bad field or random dat");
    const unsigned g = (unsigned)((count + kBlock - 1) / kBlock);
    inject_kernel<<<g, kBlock, 0, as_stream(stream)>>>(first, count, parts->stride, parts->pos, parts->vel, parts->p2c,
                                                       *inlet, n_cells, c_ef, rand, rand_stride, ion_velocity,
                                                       charge / mass, 0.5 * dt, counters);
    OPPX_LAUNCH_CHECK("fempic_inject");
    return OPPX_OK;
}

int oppx_fempic_node_charge_density(int64_t n_nodes, double* node_charge, const double* node_volume, double spwt,
                                    void* stream) {
    OPPX_ARG_CHECK(n_nodes >= 0 && (n_nodes == 0 || (node_charge && node_volume)),
                   "oppx_fempic_node_charge_density: bad arguments");
    if (n_nodes == 0) return OPPX_OK;
    node_density_kernel<<<grid_for(n_nodes, kBlock), kBlock, 0, as_stream(stream)>>>(n_nodes, node_charge, node_volume,
                                                                                    spwt);
    OPPX_LAUNCH_CHECK("fempic_node_charge_density");
    return OPPX_OK;
}

int oppx_fempic_electric_field(int64_t n_cells, double* c_ef, const double* c_sd, const int32_t* c2n,
                               const double* node_potential, int64_t n_nodes, oppx_move_counters* counters,
                               void* stream) {
    OPPX_ARG_CHECK(n_cells >= 0 && n_nodes > 0 && (n_cells == 0 || (c_ef && c_sd && c2n && node_potential)),
                   "oppx_fempic_electric_field: bad arguments");
    if (n_cells == 0) return OPPX_OK;
    efield_kernel<<<grid_for(n_cells, kBlock), kBlock, 0, as_stream(stream)>>>(n_cells, c_ef, c_sd, c2n,
                                                                              node_potential, n_nodes, counters);
    OPPX_LAUNCH_CHECK("fempic_electric_field");
    return OPPX_OK;
}

int oppx_fempic_boundary_potential(int64_t n_nodes, const int32_t* node_type, double wall_potential,
                                   double* node_bnd_pot, void* stream) {
    OPPX_ARG_CHECK(n_nodes >= 0 && (n_nodes == 0 || (node_type && node_bnd_pot)),
                   "oppx_fempic_boundary_potential: bad arguments");
    if (n_nodes == 0) return OPPX_OK;
    boundary_pot_kernel<<<grid_for(n_nodes, kBlock), kBlock, 0, as_stream(stream)>>>(n_nodes, node_type,
                                                                                    wall_potential, node_bnd_pot);
    OPPX_LAUNCH_CHECK("fempic_boundary_potential");
    return OPPX_OK;
}

size_t oppx_max_scratch_bytes(int64_t n) { return (size_t)(max_blocks(n < 1 ? 1 : n)) * sizeof(MaxCand); }

int oppx_max_reduce(int64_t n, const double* values, int abs_values, double init, double* result, void* scratch,
                    size_t scratch_bytes, void* stream) {
    OPPX_ARG_CHECK(n >= 0 && result && (n == 0 || values), "oppx_max_reduce: bad arguments");
    OPPX_ARG_CHECK(scratch && scratch_bytes >= oppx_max_scratch_bytes(n), "oppx_max_reduce: scratch too small");
    hipStream_t s = as_stream(stream);
    const int64_t nb = n == 0 ? 0 : max_blocks(n);
    MaxCand* part = reinterpret_cast<MaxCand*>(scratch);
    if (nb > 0) max_partial_kernel<<<(unsigned)nb, kBlock, 0, s>>>(n, values, abs_values ? 1 : 0, part);
    max_final_kernel<<<1, kBlock, 0, s>>>(nb, part, init, result);
    OPPX_LAUNCH_CHECK("max_reduce");
    return OPPX_OK;
}

}  // extern "C"
