// NESO-advection hot path on gfx950: periodic drift + axis-wise multi-hop walk over a 2-D quad
// mesh, one lane per particle.
//
// Elemental math: apps/app_neso_advection/kernels.h:41-113; loops:
// apps_codegen/app_neso_advection_mdir_cg/{seq,hip}/{update_pos,move}_kernel_loop.hpp. The reference's
// GCC build contracts pos += vel*dt, pos + n*extent and FMOD's a - b*trunc(a/b) into FMAs; they
// are explicit fma() here (file built with -ffp-contract=off). OPP_DO_ONCE holds on the first
// hop only.

#include "oppx_internal.h"

namespace oppx {
namespace {

constexpr int kBlock = 256;

struct NesoArgs {
    int64_t n_cells;
    const oppx_quad_cell* __restrict__ cells;
    int64_t stride;
    double* __restrict__ pos;
    const double* __restrict__ vel;
    int32_t* __restrict__ mdir;
    int32_t* __restrict__ p2c;
    double dt, cw, ext_x, ext_y;
    int64_t n_max;
    oppx_move_counters* ctr;
    int32_t* __restrict__ hops;
    int gnx, gny;  // grid mode: the mesh is the nx x ny periodic grid, its records computed in registers
    double inv_gnx;
};

// Record of quad (ix, iy) of the periodic grid, as apps/app_neso_advection/advec_misc.h:115-143 builds
// it: lower-left corner (x * cell_width, y * cell_width), neighbours wrapped on both axes.
__device__ __forceinline__ int grid_wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ __forceinline__ double periodic_update(double p, double v, double dt, double ext) {
    p = fma(v, dt, p);
    const int off = (int)(fabs(p) + 2.0);
    const double t = fma((double)off, ext, p);
    return fma(-ext, trunc(t / ext), t);
}

// Particle state as one grid-stride iteration loads it.
struct NesoPart {
    double x, y, vx, vy;
    int cell, mx, my;
};

__device__ __forceinline__ NesoPart neso_load(const NesoArgs& P, int flags, int64_t i) {
    const int64_t S = P.stride;
    NesoPart q;
    q.x = __builtin_nontemporal_load(P.pos + i);
    q.y = __builtin_nontemporal_load(P.pos + S + i);
    q.vx = q.vy = 0.0;
    if (flags & OPPX_NESO_UPDATE) {
        q.vx = __builtin_nontemporal_load(P.vel + i);
        q.vy = __builtin_nontemporal_load(P.vel + S + i);
    }
    q.cell = kMaxCell;
    q.mx = q.my = 0;
    if (flags & OPPX_NESO_MOVE) {
        q.cell = __builtin_nontemporal_load(P.p2c + i);
        q.mx = __builtin_nontemporal_load(P.mdir + i);
        q.my = __builtin_nontemporal_load(P.mdir + S + i);
    }
    return q;
}

// update_pos_kernel then move_kernel for particle i; returns the evaluations made.
template <bool GRID>
__device__ __forceinline__ int neso_particle(const NesoArgs& P, int flags, int64_t i, NesoPart q, bool& fault) {
    const int64_t S = P.stride;
    int h = 0;
    double x = q.x, y = q.y;
    if (flags & OPPX_NESO_UPDATE) {
        x = periodic_update(x, q.vx, P.dt, P.ext_x);
        y = periodic_update(y, q.vy, P.dt, P.ext_y);
        __builtin_nontemporal_store(x, P.pos + i);
        __builtin_nontemporal_store(y, P.pos + S + i);
    }
    int cell = q.cell;
    if (cell == kMaxCell) return 0;
    if (!((unsigned)cell < (unsigned)P.n_cells)) {
        fault = true;
        return 0;
    }
    int mx = q.mx, my = q.my;
    const double cw = P.cw;
    int ix = 0, iy = 0;
    if (GRID) {
        // cell = ix + nx * iy: the quotient from one FP64 product, then corrected by one step
        // (the product's error stays below one for cells under 2^31)
        iy = (int)((double)cell * P.inv_gnx);
        ix = cell - iy * P.gnx;
        if (ix < 0) { iy--; ix += P.gnx; }
        else if (ix >= P.gnx) { iy++; ix -= P.gnx; }
    }
    for (;;) {
        double llx, lly;
        int4 nb = make_int4(0, 0, 0, 0);
        if (GRID) {
            llx = (double)ix * cw;
            lly = (double)iy * cw;
        } else {
            const oppx_quad_cell c = P.cells[cell];
            llx = c.llx;
            lly = c.lly;
            nb = make_int4(c.c2c[0], c.c2c[1], c.c2c[2], c.c2c[3]);
        }
        const double dx = x - llx, dy = y - lly;
        h++;
        if (h == 1) {
            const bool xp = ((dx > cw) && (P.ext_x > 2 * dx - cw)) || ((dx < 0) && (P.ext_x < cw - 2 * dx));
            const bool yp = ((dy > cw) && (P.ext_y > 2 * dy - cw)) || ((dy < 0) && (P.ext_y < cw - 2 * dy));
            mx = xp ? 1 : -1;
            my = yp ? 1 : -1;
        }
        int next = 0;
        bool hop = false;
        if ((dx >= 0.0) && (dx <= cw)) mx = 0;
        else if (mx > 0) { next = nb.y; hop = true; if (GRID) ix = grid_wrap(ix + 1, P.gnx); }
        else if (mx < 0) { next = nb.x; hop = true; if (GRID) ix = grid_wrap(ix - 1, P.gnx); }
        if (!hop) {
            if ((dy >= 0.0) && (dy <= cw)) my = 0;
            else if (my > 0) { next = nb.w; hop = true; if (GRID) iy = grid_wrap(iy + 1, P.gny); }
            else if (my < 0) { next = nb.z; hop = true; if (GRID) iy = grid_wrap(iy - 1, P.gny); }
        }
        if (!hop) break;  // MOVE_DONE
        if (GRID) next = ix + P.gnx * iy;
        if (!((unsigned)next < (unsigned)P.n_cells) || h >= OPPX_MAX_HOPS) {
            fault = true;
            break;
        }
        cell = next;
    }
    __builtin_nontemporal_store(mx, P.mdir + i);
    __builtin_nontemporal_store(my, P.mdir + S + i);
    if (!fault) __builtin_nontemporal_store(cell, P.p2c + i);
    return h;
}

// Resident grid, grid-stride over the particles, kNesoPerLane particles per lane per iteration (all
// loaded before any is walked, so each lane keeps that many particles' loads in flight); the
// per-block statistics then reach the counters from a few thousand blocks instead of one block per
// 256 particles (same-address atomics serialise).
constexpr int kNesoPerLane = 4;

template <bool GRID>
__global__ void __launch_bounds__(kBlock) neso_step_kernel(NesoArgs P, int flags) {
    int64_t n = P.ctr->size;
    n = n < 0 ? 0 : (n > P.n_max ? P.n_max : n);
    long long hsum = 0;
    int hmax = 0, nfault = 0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i0 = (int64_t)blockIdx.x * kBlock + threadIdx.x; i0 < n; i0 += kNesoPerLane * stride) {
        NesoPart part[kNesoPerLane];
#pragma unroll
        for (int u = 0; u < kNesoPerLane; u++) {
            const int64_t i = i0 + u * stride;
            part[u] = i < n ? neso_load(P, flags, i) : NesoPart{};
        }
#pragma unroll
        for (int u = 0; u < kNesoPerLane; u++) {
            const int64_t i = i0 + u * stride;
            if (i < n) {
                bool f = false;
                const int h = neso_particle<GRID>(P, flags, i, part[u], f);
                if (P.hops && (flags & OPPX_NESO_MOVE)) P.hops[i] = h;
                hsum += h;
                hmax = hmax > h ? hmax : h;
                nfault += f ? 1 : 0;
            }
        }
    }
    __shared__ long long s_sum[kBlock / kWave];
    __shared__ int s_max[kBlock / kWave], s_flt[kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const long long hs = wave_sum_i64(hsum);
    const int hm = wave_max_i32(hmax), nf = wave_sum_i32(nfault);
    if (lane == 0) { s_sum[w] = hs; s_max[w] = hm; s_flt[w] = nf; }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        int m = 0, f = 0;
        for (int q = 0; q < kBlock / kWave; q++) { t += s_sum[q]; m = m > s_max[q] ? m : s_max[q]; f += s_flt[q]; }
        if (t) atomicAdd((unsigned long long*)&P.ctr->total_hops, (unsigned long long)t);
        if (m) atomicMax(&P.ctr->max_hops, m);
        if (f) atomicAdd(&P.ctr->faults, f);
    }
}

// Counts the records that differ, bit for bit, from the nx x ny periodic grid's.
__global__ void quad_grid_check_kernel(const oppx_quad_cell* __restrict__ cells, int64_t n, int nx, int ny, double cw,
                                       int32_t* __restrict__ mismatches) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (c < n) {
        const int ix = (int)(c % nx), iy = (int)(c / nx);
        const oppx_quad_cell q = cells[c];
        bad = __double_as_longlong(q.llx) != __double_as_longlong((double)ix * cw) ||
              __double_as_longlong(q.lly) != __double_as_longlong((double)iy * cw) ||
              q.c2c[0] != grid_wrap(ix - 1, nx) + nx * iy || q.c2c[1] != grid_wrap(ix + 1, nx) + nx * iy ||
              q.c2c[2] != ix + nx * grid_wrap(iy - 1, ny) || q.c2c[3] != ix + nx * grid_wrap(iy + 1, ny);
    }
    const unsigned long long m = __ballot(bad);
    if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicAdd(mismatches, (int)__popcll(m));
}

__global__ void neso_reset_kernel(oppx_move_counters* ctr) {
    ctr->removed = 0;
    ctr->holes_low = 0;
    ctr->max_hops = 0;
    ctr->total_hops = 0;
}

__global__ void quad_cells_kernel(const double* __restrict__ ll, const int32_t* __restrict__ c2c, int64_t n,
                                  oppx_quad_cell* __restrict__ out) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    oppx_quad_cell q;
    q.llx = ll[2 * c];
    q.lly = ll[2 * c + 1];
    for (int k = 0; k < 4; k++) {
        const int v = c2c[4 * c + k];
        q.c2c[k] = (v >= 0 && v < n) ? v : -2;  // invalid neighbours are rejected by the walk as faults
    }
    out[c] = q;
}

__global__ void zero_i32_kernel(int32_t* p) { *p = 0; }

int neso_step_impl(int64_t n_max, const oppx_neso_parts* parts, const oppx_quad_cell* cells, int64_t n_cells,
                   int gnx, int gny, double dt, double cell_width, double ext_x, double ext_y, int flags,
                   oppx_move_counters* counters, int32_t* hops, void* stream) {
    const bool grid = gnx > 0;
    OPPX_ARG_CHECK(parts && counters, "oppx_neso_step: null pointer");
    OPPX_ARG_CHECK(flags >= 1 && flags <= 3, "oppx_neso_step: flags must be OPPX_NESO_UPDATE and/or OPPX_NESO_MOVE");
    OPPX_ARG_CHECK(!(flags & OPPX_NESO_MOVE) || ((cells || grid) && parts->mdir && parts->p2c),
                   "oppx_neso_step: move needs cells, mdir, p2c");
    OPPX_ARG_CHECK(!(flags & OPPX_NESO_UPDATE) || (parts->pos && parts->vel), "oppx_neso_step: update needs pos, vel");
    OPPX_ARG_CHECK(n_cells > 0 && n_cells < kMaxCell, "oppx_neso_step: n_cells out of range");
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, "oppx_neso_step: n_max beyond capacity");
    OPPX_ARG_CHECK(cell_width > 0 && ext_x > 0 && ext_y > 0, "oppx_neso_step: widths must be positive");
    hipStream_t s = as_stream(stream);
    neso_reset_kernel<<<1, 1, 0, s>>>(counters);
    OPPX_LAUNCH_CHECK("neso_reset");
    if (n_max == 0) return OPPX_OK;
    NesoArgs P{};
    P.n_cells = n_cells;
    P.cells = cells;
    P.stride = parts->stride;
    P.pos = parts->pos;
    P.vel = parts->vel;
    P.mdir = parts->mdir;
    P.p2c = parts->p2c;
    P.dt = dt;
    P.cw = cell_width;
    P.ext_x = ext_x;
    P.ext_y = ext_y;
    P.n_max = n_max;
    P.ctr = counters;
    P.hops = hops;
    P.gnx = gnx;
    P.gny = gny;
    P.inv_gnx = grid ? 1.0 / gnx : 0.0;
    auto kern = grid ? neso_step_kernel<true> : neso_step_kernel<false>;
    unsigned g = resident_blocks(reinterpret_cast<const void*>(kern), kBlock);
    const unsigned need = grid_for(n_max, kBlock);
    if (g > need) g = need;
    kern<<<g, kBlock, 0, s>>>(P, flags);
    OPPX_LAUNCH_CHECK("neso_step");
    return OPPX_OK;
}
// verify_kernel (apps/app_neso_advection/kernels.h:115-150) over the counters->size particles: a
// particle counts as incorrect when its cell's global index c_gbl_idx[p2c] splits into a negative
// (ix, iy) = (g mod nx, g div nx), or when a coordinate lies outside [ll, ll + cell_width] with
// ll = ix * cell_width (a product then a sum: the reference build forms them apart, no FMA). Block
// sums of the count are added to *count.
__global__ void __launch_bounds__(kBlock) neso_verify_kernel(int64_t n_max, int64_t stride, const double* __restrict__ pos,
                                                             const int32_t* __restrict__ p2c,
                                                             const int32_t* __restrict__ c_gbl_idx, int64_t n_cells,
                                                             int nx, double cw, const oppx_move_counters* __restrict__ ctr,
                                                             int32_t* __restrict__ count, int32_t* __restrict__ faults) {
    int64_t n = ctr->size;
    n = n < 0 ? 0 : (n > n_max ? n_max : n);
    int bad = 0, flt = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const int c = p2c[i];
        if (!((unsigned)c < (unsigned)n_cells)) {  // the reference would read outside c_idx
            flt++;
            continue;
        }
        const int g = c_gbl_idx[c];
        const int iy = g / nx, ix = g - iy * nx;
        if (ix < 0 || iy < 0) {
            bad++;
            continue;
        }
        const double llx = (double)ix * cw, x = pos[i];
        if (x < llx || x > llx + cw) {
            bad++;
            continue;
        }
        const double lly = (double)iy * cw, y = pos[stride + i];
        if (y < lly || y > lly + cw) bad++;
    }
    __shared__ int s_bad[kBlock / kWave], s_flt[kBlock / kWave];
    bad = wave_sum_i32(bad);
    flt = wave_sum_i32(flt);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) { s_bad[w] = bad; s_flt[w] = flt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        int b = 0, f = 0;
        for (int q = 0; q < kBlock / kWave; q++) { b += s_bad[q]; f += s_flt[q]; }
        if (b) atomicAdd(count, b);
        if (f && faults) atomicAdd(faults, f);
    }
}

}  // namespace
}  // namespace oppx

using namespace oppx;

extern "C" {

int oppx_quad_cells_build(const double* c_pos_ll, const int32_t* c2c, int64_t n_cells, oppx_quad_cell* cells,
                          void* stream) {
    OPPX_ARG_CHECK(c_pos_ll && c2c && cells && n_cells > 0 && n_cells < kMaxCell, "oppx_quad_cells_build: bad args");
    quad_cells_kernel<<<grid_for(n_cells, 256), 256, 0, as_stream(stream)>>>(c_pos_ll, c2c, n_cells, cells);
    OPPX_LAUNCH_CHECK("quad_cells");
    return OPPX_OK;
}

int oppx_quad_grid_check(const oppx_quad_cell* cells, int64_t n_cells, int nx, int ny, double cell_width,
                         int32_t* mismatches, void* stream) {
    OPPX_ARG_CHECK(cells && mismatches && nx > 0 && ny > 0 && (int64_t)nx * ny == n_cells && n_cells < kMaxCell,
                   "oppx_quad_grid_check: bad arguments (n_cells must be nx * ny)");
    hipStream_t s = as_stream(stream);
    zero_i32_kernel<<<1, 1, 0, s>>>(mismatches);
    quad_grid_check_kernel<<<grid_for(n_cells, 256), 256, 0, s>>>(cells, n_cells, nx, ny, cell_width, mismatches);
    OPPX_LAUNCH_CHECK("quad_grid_check");
    return OPPX_OK;
}


int oppx_neso_step(int64_t n_max, const oppx_neso_parts* parts, const oppx_quad_cell* cells, int64_t n_cells,
                   double dt, double cell_width, double ext_x, double ext_y, int flags,
                   oppx_move_counters* counters, int32_t* hops, void* stream) {
    return neso_step_impl(n_max, parts, cells, n_cells, 0, 0, dt, cell_width, ext_x, ext_y, flags, counters, hops,
                          stream);
}

int oppx_neso_step_grid(int64_t n_max, const oppx_neso_parts* parts, int nx, int ny, double dt, double cell_width,
                        double ext_x, double ext_y, int flags, oppx_move_counters* counters, int32_t* hops,
                        void* stream) {
    OPPX_ARG_CHECK(nx > 0 && ny > 0 && (int64_t)nx * ny < kMaxCell, "oppx_neso_step_grid: nx, ny out of range");
    return neso_step_impl(n_max, parts, nullptr, (int64_t)nx * ny, nx, ny, dt, cell_width, ext_x, ext_y, flags,
                          counters, hops, stream);
}

int oppx_neso_verify(int64_t n_max, int64_t stride, const double* pos, const int32_t* p2c, const int32_t* c_gbl_idx,
                     int64_t n_cells, int nx, double cell_width, oppx_move_counters* counters, int32_t* count,
                     void* stream) {
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max, "oppx_neso_verify: need 0 <= n_max <= stride");
    OPPX_ARG_CHECK(counters && count, "oppx_neso_verify: null counters or count");
    OPPX_ARG_CHECK(nx > 0 && n_cells > 0 && n_cells < kMaxCell, "oppx_neso_verify: nx or n_cells out of range");
    if (n_max == 0) return OPPX_OK;
    OPPX_ARG_CHECK(pos && p2c && c_gbl_idx, "oppx_neso_verify: null array");
    unsigned g = grid_for(n_max, kBlock);
    if (g > 4096) g = 4096;
    neso_verify_kernel<<<g, kBlock, 0, as_stream(stream)>>>(n_max, stride, pos, p2c, c_gbl_idx, n_cells, nx, cell_width,
                                                            counters, count, &counters->faults);
    OPPX_LAUNCH_CHECK("neso_verify");
    return OPPX_OK;
}

}  // extern "C"
