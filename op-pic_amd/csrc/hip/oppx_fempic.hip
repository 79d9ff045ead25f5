// FEM-PIC particle hot path on gfx950: push, direct-hop relocation, multi-hop barycentric walk,
// FP64 atomic charge deposit, and the fused single-pass step.
//
// Elemental math: apps/app_fempic/kernels.h:93-175. Loop semantics: the generated seq/omp/hip
// loops under apps_codegen/app_fempic_cg/{seq,omp,hip}/. The reference's GCC -O3 build contracts
// the barycentric and push expressions into FMAs; they are written out here with fma() and this
// file is compiled with -ffp-contract=off so cell assignment and hop counts are bit-exact.
//
// One lane per particle; particles are SoA (coalesced), cell records are AoS (one 128 B line of
// c_det + a 32 B aux record per hop). Blocks are remapped so each XCD walks one contiguous
// eighth of the particle range.

#include "oppx_internal.h"

namespace oppx {
namespace {

constexpr int kBlock = 256;

struct StepArgs {
    int64_t n_cells;
    const double* __restrict__ det;
    const oppx_tet_aux* __restrict__ aux;
    const int32_t* __restrict__ c2n;
    const double* __restrict__ ef;
    int64_t stride;
    double* __restrict__ pos;
    double* __restrict__ vel;
    double* __restrict__ lc;
    int32_t* __restrict__ p2c;
    const int32_t* __restrict__ dh_cell;
    int64_t dx, dy, dz;
    double mx, my, mz, inv_gs;
    double k, dt;
    int64_t n_max;
    oppx_move_counters* ctr;
    uint64_t* __restrict__ hole_bits;
    int32_t* __restrict__ hops;
    double* __restrict__ charge;
};

// findStructuredCellIndex3D (opp_lib/include/opp_direct_hop_core.h:410-426) with the x86
// size_t conversion made explicit: an axis is inside iff -1 < t < dims (t in (-1,0) truncates
// to 0); then the overlay owner (kMaxCell = no owner -> remove, opp_lib/include/opp_omp.h:206-212).
__device__ __forceinline__ int dh_lookup(const StepArgs& P, double x, double y, double z) {
    const double tx = (x - P.mx) * P.inv_gs;
    const double ty = (y - P.my) * P.inv_gs;
    const double tz = (z - P.mz) * P.inv_gs;
    if (!(tx > -1.0 && tx < (double)P.dx && ty > -1.0 && ty < (double)P.dy && tz > -1.0 && tz < (double)P.dz))
        return kMaxCell;
    const int64_t ix = (int64_t)tx, iy = (int64_t)ty, iz = (int64_t)tz;
    return P.dh_cell[ix + P.dx * (iy + P.dy * iz)];
}

enum WalkEnd { kWalkDone = 0, kWalkRemoved = 1, kWalkFault = 2 };

// Multi-hop walk from `cell` (apps/app_fempic/kernels.h:106-161 per hop, do/while of
// apps_codegen/app_fempic_cg/hip/move_kernel_loop.hpp:162-183). lc = c2*fnma(d3,z,fma(d2,y,fnma(d1,x,d0))).
__device__ __forceinline__ int walk(const StepArgs& P, double x, double y, double z, int& cell, double lc[4],
                                    int& h) {
    for (;;) {
        const double2* __restrict__ d = reinterpret_cast<const double2*>(P.det + 16 * (int64_t)cell);
        const double2 a0 = d[0], a1 = d[1], a2 = d[2], a3 = d[3], a4 = d[4], a5 = d[5], a6 = d[6], a7 = d[7];
        const oppx_tet_aux au = P.aux[cell];
        const double c2 = au.c2;
        lc[0] = c2 * fma(-a1.y, z, fma(a1.x, y, fma(-a0.y, x, a0.x)));
        lc[1] = c2 * fma(-a3.y, z, fma(a3.x, y, fma(-a2.y, x, a2.x)));
        lc[2] = c2 * fma(-a5.y, z, fma(a5.x, y, fma(-a4.y, x, a4.x)));
        lc[3] = c2 * fma(-a7.y, z, fma(a7.x, y, fma(-a6.y, x, a6.x)));
        h++;
        if ((lc[0] > -1e-10) && ((lc[0] - 1.0) < 1e-10) && (lc[1] > -1e-10) && ((lc[1] - 1.0) < 1e-10) &&
            (lc[2] > -1e-10) && ((lc[2] - 1.0) < 1e-10) && (lc[3] > -1e-10) && ((lc[3] - 1.0) < 1e-10))
            return kWalkDone;
        int mi = 0;
        double ml = lc[0];
        if (lc[1] < ml) { ml = lc[1]; mi = 1; }
        if (lc[2] < ml) { ml = lc[2]; mi = 2; }
        if (lc[3] < ml) { ml = lc[3]; mi = 3; }
        const int nb = au.c2c[mi];
        if (nb < 0) {
            cell = kMaxCell;
            return kWalkRemoved;
        }
        if (nb >= P.n_cells || h >= OPPX_MAX_HOPS) return kWalkFault;
        cell = nb;
    }
}

template <bool PUSH, bool MOVE, bool DH, bool DEPOSIT, bool STATS>
__global__ void __launch_bounds__(kBlock) fempic_pass_kernel(StepArgs P, uint32_t nblocks) {
    const uint32_t b = xcd_block(blockIdx.x, nblocks);
    const int64_t i = (int64_t)b * kBlock + threadIdx.x;
    int64_t n = P.ctr->size;
    n = n < 0 ? 0 : (n > P.n_max ? P.n_max : n);  // never beyond the host's bound
    const bool live = i < n;
    const int64_t S = P.stride;

    int cell = live ? P.p2c[i] : kMaxCell;
    bool fault = live && !((unsigned)cell < (unsigned)P.n_cells);
    bool removed = false;
    int h = 0;
    double lc[4];

    if (live && !fault && (PUSH || MOVE)) {
        double x = P.pos[i], y = P.pos[S + i], z = P.pos[2 * S + i];
        if (PUSH) {
            // calculate_new_pos_vel_kernel: v = fma(k, E, v); x = fma(v, dt, x)
            const double* e = P.ef + 3 * (int64_t)cell;
            const double ex = e[0], ey = e[1], ez = e[2];
            double vx = P.vel[i], vy = P.vel[S + i], vz = P.vel[2 * S + i];
            vx = fma(P.k, ex, vx);
            x = fma(vx, P.dt, x);
            vy = fma(P.k, ey, vy);
            y = fma(vy, P.dt, y);
            vz = fma(P.k, ez, vz);
            z = fma(vz, P.dt, z);
            P.vel[i] = vx; P.vel[S + i] = vy; P.vel[2 * S + i] = vz;
            P.pos[i] = x; P.pos[S + i] = y; P.pos[2 * S + i] = z;
        }
        if (MOVE) {
            if (DH) {
                cell = dh_lookup(P, x, y, z);
                if (cell == kMaxCell) removed = true;
                else if (!((unsigned)cell < (unsigned)P.n_cells)) fault = true;
            }
            if (!removed && !fault) {
                const int r = walk(P, x, y, z, cell, lc, h);
                removed = (r == kWalkRemoved);
                fault = (r == kWalkFault);
                P.lc[i] = lc[0]; P.lc[S + i] = lc[1]; P.lc[2 * S + i] = lc[2]; P.lc[3 * S + i] = lc[3];
            }
            if (!fault) P.p2c[i] = removed ? kMaxCell : cell;
            if (P.hops) P.hops[i] = h;
        }
    }
    if (DEPOSIT && live && !removed && !fault) {
        if (!MOVE) {
            lc[0] = P.lc[i]; lc[1] = P.lc[S + i]; lc[2] = P.lc[2 * S + i]; lc[3] = P.lc[3 * S + i];
        }
        const int4 m = *reinterpret_cast<const int4*>(P.c2n + 4 * (int64_t)cell);
        atomicAdd(P.charge + m.x, lc[0]);
        atomicAdd(P.charge + m.y, lc[1]);
        atomicAdd(P.charge + m.z, lc[2]);
        atomicAdd(P.charge + m.w, lc[3]);
    }

    const int lane = threadIdx.x & (kWave - 1);
    if (MOVE) {
        // removal bitmap: one 64-bit word per wave, written for every word of [0, n_max)
        const unsigned long long rmask = __ballot(live && removed);
        if (lane == 0) P.hole_bits[i >> 6] = rmask;
    }
    const unsigned long long fmask = __ballot(fault);
    if (lane == 0 && fmask) atomicAdd(&P.ctr->faults, (int)__popcll(fmask));
    if (STATS) {
        __shared__ long long s_sum[kBlock / kWave];
        __shared__ int s_max[kBlock / kWave];
        const long long hs = wave_sum_i64(h);
        const int hm = wave_max_i32(h);
        const int w = threadIdx.x / kWave;
        if (lane == 0) { s_sum[w] = hs; s_max[w] = hm; }
        __syncthreads();
        if (threadIdx.x == 0) {
            long long t = 0;
            int m = 0;
            for (int q = 0; q < kBlock / kWave; q++) { t += s_sum[q]; m = m > s_max[q] ? m : s_max[q]; }
            if (t) atomicAdd((unsigned long long*)&P.ctr->total_hops, (unsigned long long)t);
            if (m) atomicMax(&P.ctr->max_hops, m);
        }
    }
}

// Per-move reset of the counters (a kernel, not a memset, so the step can be graph-captured).
__global__ void reset_counters_kernel(oppx_move_counters* ctr) {
    ctr->removed = 0;
    ctr->holes_low = 0;
    ctr->max_hops = 0;
    ctr->total_hops = 0;
}

__global__ void tet_aux_kernel(const double* __restrict__ vol, const int32_t* __restrict__ c2c, int64_t n,
                               oppx_tet_aux* __restrict__ aux) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    oppx_tet_aux a;
    a.c2 = (1.0 / 6.0) / vol[c];  // coefficient2 of apps/app_fempic/kernels.h:110, IEEE division
    for (int q = 0; q < 4; q++) {
        const int v = c2c[4 * c + q];
        a.c2c[q] = (v >= -1 && v < n) ? v : -2;  // -2: invalid entry, rejected by the walk as a fault
    }
    a.pad[0] = a.pad[1] = 0;
    aux[c] = a;
}

int fill_args(StepArgs& P, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh, const oppx_dh_grid* dh,
              oppx_move_counters* ctr) {
    OPPX_ARG_CHECK(parts && mesh && ctr, "fempic: null descriptor");
    OPPX_ARG_CHECK(mesh->n_cells > 0 && mesh->n_cells < kMaxCell, "fempic: n_cells out of range");
    OPPX_ARG_CHECK(parts->stride >= 0, "fempic: negative stride");
    P = StepArgs{};
    P.n_cells = mesh->n_cells;
    P.det = mesh->det;
    P.aux = mesh->aux;
    P.c2n = mesh->c2n;
    P.ef = mesh->ef;
    P.stride = parts->stride;
    P.pos = parts->pos;
    P.vel = parts->vel;
    P.lc = parts->lc;
    P.p2c = parts->p2c;
    P.ctr = ctr;
    if (dh) {
        OPPX_ARG_CHECK(dh->inv_gs > 0, "fempic: overlay grid spacing must be positive");
        OPPX_ARG_CHECK(dh->cell && dh->dims[0] > 0 && dh->dims[1] > 0 && dh->dims[2] > 0, "fempic: bad overlay");
        P.dh_cell = dh->cell;
        P.dx = dh->dims[0]; P.dy = dh->dims[1]; P.dz = dh->dims[2];
        P.mx = dh->min[0]; P.my = dh->min[1]; P.mz = dh->min[2];
        P.inv_gs = dh->inv_gs;
    }
    return OPPX_OK;
}

template <bool PUSH, bool MOVE, bool DEPOSIT>
void launch_pass(const StepArgs& P, bool dh, bool stats, unsigned g, hipStream_t s) {
#define OPPX_PASS(DHV, STV) fempic_pass_kernel<PUSH, MOVE, DHV, DEPOSIT, STV><<<g, kBlock, 0, s>>>(P, g)
    if (dh) {
        if (stats) OPPX_PASS(true, true); else OPPX_PASS(true, false);
    } else {
        if (stats) OPPX_PASS(false, true); else OPPX_PASS(false, false);
    }
#undef OPPX_PASS
}

}  // namespace
}  // namespace oppx

using namespace oppx;

extern "C" {

int oppx_tet_aux_build(const double* c_vol, const int32_t* c2c, int64_t n_cells, oppx_tet_aux* aux, void* stream) {
    OPPX_ARG_CHECK(c_vol && c2c && aux && n_cells > 0, "oppx_tet_aux_build: bad arguments");
    tet_aux_kernel<<<grid_for(n_cells, 256), 256, 0, as_stream(stream)>>>(c_vol, c2c, n_cells, aux);
    OPPX_LAUNCH_CHECK("tet_aux");
    return OPPX_OK;
}

int oppx_fempic_push(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh, double k,
                     double dt, oppx_move_counters* counters, void* stream) {
    StepArgs P;
    int rc = fill_args(P, parts, mesh, nullptr, counters);
    if (rc) return rc;
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, "oppx_fempic_push: n_max beyond capacity");
    if (n_max == 0) return OPPX_OK;
    P.n_max = n_max;
    P.k = k;
    P.dt = dt;
    launch_pass<true, false, false>(P, false, false, grid_for(n_max, kBlock), as_stream(stream));
    OPPX_LAUNCH_CHECK("fempic_push");
    return OPPX_OK;
}

int oppx_fempic_move(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                     const oppx_dh_grid* dh, oppx_move_counters* counters, uint64_t* hole_bits, int32_t* hops,
                     void* stream) {
    StepArgs P;
    int rc = fill_args(P, parts, mesh, dh, counters);
    if (rc) return rc;
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, "oppx_fempic_move: n_max beyond capacity");
    OPPX_ARG_CHECK(hole_bits != nullptr, "oppx_fempic_move: hole_bits required");
    hipStream_t s = as_stream(stream);
    reset_counters_kernel<<<1, 1, 0, s>>>(counters);
    OPPX_LAUNCH_CHECK("reset_counters");
    if (n_max == 0) return OPPX_OK;
    P.n_max = n_max;
    P.hole_bits = hole_bits;
    P.hops = hops;
    launch_pass<false, true, false>(P, dh != nullptr, true, grid_for(n_max, kBlock), s);
    OPPX_LAUNCH_CHECK("fempic_move");
    return OPPX_OK;
}

int oppx_fempic_deposit_atomic(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                               double* node_charge, oppx_move_counters* counters, void* stream) {
    StepArgs P;
    int rc = fill_args(P, parts, mesh, nullptr, counters);
    if (rc) return rc;
    OPPX_ARG_CHECK(node_charge && mesh->c2n, "oppx_fempic_deposit_atomic: null node array or c2n");
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, "oppx_fempic_deposit_atomic: n_max beyond capacity");
    if (n_max == 0) return OPPX_OK;
    P.n_max = n_max;
    P.charge = node_charge;
    launch_pass<false, false, true>(P, false, false, grid_for(n_max, kBlock), as_stream(stream));
    OPPX_LAUNCH_CHECK("fempic_deposit");
    return OPPX_OK;
}

int oppx_fempic_step_fused(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                           const oppx_dh_grid* dh, double k, double dt, oppx_move_counters* counters,
                           uint64_t* hole_bits, int32_t* hops, double* node_charge, void* stream) {
    StepArgs P;
    int rc = fill_args(P, parts, mesh, dh, counters);
    if (rc) return rc;
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, "oppx_fempic_step_fused: n_max beyond capacity");
    OPPX_ARG_CHECK(hole_bits != nullptr, "oppx_fempic_step_fused: hole_bits required");
    hipStream_t s = as_stream(stream);
    reset_counters_kernel<<<1, 1, 0, s>>>(counters);
    OPPX_LAUNCH_CHECK("reset_counters");
    if (n_max == 0) return OPPX_OK;
    P.n_max = n_max;
    P.k = k;
    P.dt = dt;
    P.hole_bits = hole_bits;
    P.hops = hops;
    P.charge = node_charge;
    const unsigned g = grid_for(n_max, kBlock);
    if (node_charge) launch_pass<true, true, true>(P, dh != nullptr, true, g, s);
    else launch_pass<true, true, false>(P, dh != nullptr, true, g, s);
    OPPX_LAUNCH_CHECK("fempic_step_fused");
    return OPPX_OK;
}

}  // extern "C"
