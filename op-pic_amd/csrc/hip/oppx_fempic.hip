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
    int tile;        // particles per wave tile (multiple of 64, <= kMaxTile)
    int refill_min;  // free lanes that trigger a refill
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
__device__ __forceinline__ int walk(const StepArgs& P, double x, double y, double z, int& cell, double& l0,
                                    double& l1, double& l2, double& l3, int& h) {
    for (;;) {
        const double2* __restrict__ d = reinterpret_cast<const double2*>(P.det + 16 * (int64_t)cell);
        const double2 a0 = d[0], a1 = d[1], a2 = d[2], a3 = d[3], a4 = d[4], a5 = d[5], a6 = d[6], a7 = d[7];
        const double4 au = *reinterpret_cast<const double4*>(P.aux + cell);  // {c2, c2c[0..1], c2c[2..3], pad}
        const double c2 = au.x;
        l0 = c2 * fma(-a1.y, z, fma(a1.x, y, fma(-a0.y, x, a0.x)));
        l1 = c2 * fma(-a3.y, z, fma(a3.x, y, fma(-a2.y, x, a2.x)));
        l2 = c2 * fma(-a5.y, z, fma(a5.x, y, fma(-a4.y, x, a4.x)));
        l3 = c2 * fma(-a7.y, z, fma(a7.x, y, fma(-a6.y, x, a6.x)));
        h++;
        if ((l0 > -1e-10) && ((l0 - 1.0) < 1e-10) && (l1 > -1e-10) && ((l1 - 1.0) < 1e-10) && (l2 > -1e-10) &&
            ((l2 - 1.0) < 1e-10) && (l3 > -1e-10) && ((l3 - 1.0) < 1e-10))
            return kWalkDone;
        // most negative weight, first index wins ties (strict <); neighbour picked without
        // a dynamically indexed array so nothing spills to scratch or LDS
        const int2 n01 = __builtin_bit_cast(int2, au.y), n23 = __builtin_bit_cast(int2, au.z);
        int nb = n01.x;
        double ml = l0;
        if (l1 < ml) { ml = l1; nb = n01.y; }
        if (l2 < ml) { ml = l2; nb = n23.x; }
        if (l3 < ml) { ml = l3; nb = n23.y; }
        if (nb < 0) {
            cell = kMaxCell;
            return kWalkRemoved;
        }
        if (nb >= P.n_cells || h >= OPPX_MAX_HOPS) return kWalkFault;
        cell = nb;
    }
}

// Workgroup-level deposit aggregation: (node, value) pairs go into an LDS open-addressing table
// (ds_add_f64), then one device-scope atomic per distinct node leaves the workgroup. Cell-ordered
// particles share most of their nodes, so this removes most of the global atomics. A pair that
// finds no slot within kProbe probes goes straight to global memory.
constexpr int kTable = 1024;
constexpr int kProbe = 16;

struct DepositTable {
    int key[kTable];
    double val[kTable];
};

__device__ __forceinline__ void table_clear(DepositTable& T) {
    for (int q = threadIdx.x; q < kTable; q += kBlock) {
        T.key[q] = -1;
        T.val[q] = 0.0;
    }
}

__device__ __forceinline__ void table_add(DepositTable& T, double* __restrict__ charge, int node, double v) {
    uint32_t h = ((uint32_t)node * 2654435761u) >> 22;  // 10 bits
    for (int p = 0; p < kProbe; p++) {
        const int k = T.key[h];
        if (k == node) {
            atomicAdd(&T.val[h], v);
            return;
        }
        if (k == -1) {
            const int prev = atomicCAS(&T.key[h], -1, node);
            if (prev == -1 || prev == node) {
                atomicAdd(&T.val[h], v);
                return;
            }
        }
        h = (h + 1) & (kTable - 1);
    }
    atomicAdd(charge + node, v);
}

__device__ __forceinline__ void table_flush(DepositTable& T, double* __restrict__ charge) {
    for (int q = threadIdx.x; q < kTable; q += kBlock) {
        const int k = T.key[q];
        if (k >= 0) atomicAdd(charge + k, T.val[q]);
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
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
    __shared__ DepositTable s_table;
    if (DEPOSIT) table_clear(s_table);

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
        if (!MOVE && DH) {  // relocation only: the walk runs in a second pass
            const int c = dh_lookup(P, x, y, z);
            P.p2c[i] = c;
            removed = (c == kMaxCell);
            if (!removed && !((unsigned)c < (unsigned)P.n_cells)) fault = true;
        }
        if (MOVE) {
            if (DH) {
                cell = dh_lookup(P, x, y, z);
                if (cell == kMaxCell) removed = true;
                else if (!((unsigned)cell < (unsigned)P.n_cells)) fault = true;
            }
            if (!removed && !fault) {
                const int r = walk(P, x, y, z, cell, l0, l1, l2, l3, h);
                removed = (r == kWalkRemoved);
                fault = (r == kWalkFault);
                P.lc[i] = l0; P.lc[S + i] = l1; P.lc[2 * S + i] = l2; P.lc[3 * S + i] = l3;
            }
            if (!fault) P.p2c[i] = removed ? kMaxCell : cell;
            if (P.hops) P.hops[i] = h;
        }
    }
    if (DEPOSIT) {
        __syncthreads();  // table cleared
        if (live && !removed && !fault) {
            if (!MOVE) {
                l0 = P.lc[i]; l1 = P.lc[S + i]; l2 = P.lc[2 * S + i]; l3 = P.lc[3 * S + i];
            }
            const int4 m = *reinterpret_cast<const int4*>(P.c2n + 4 * (int64_t)cell);
            table_add(s_table, P.charge, m.x, l0);
            table_add(s_table, P.charge, m.y, l1);
            table_add(s_table, P.charge, m.z, l2);
            table_add(s_table, P.charge, m.w, l3);
        }
        __syncthreads();
        table_flush(s_table, P.charge);
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

// ---------------------------------------------------------------------------------------------
// Wavefront-cooperative walk. Each wave owns a tile of kTile consecutive particles; a lane whose
// particle has finished its walk (or was removed) retires it and, once enough lanes are free,
// the free lanes take the next particles of the tile (ranks from a ballot, no atomics). Lanes
// stay busy while the longest walks of the tile finish, instead of idling until the slowest
// lane of a fixed 64-particle group is done. Removal bits gather in LDS per tile; the deposit
// is aggregated in the workgroup's LDS table.
constexpr int kMaxTile = 512;
constexpr int kMaxTileWords = kMaxTile / kWave;
constexpr int kWavesPerBlock = kBlock / kWave;

template <bool PUSH, bool DH, bool DEPOSIT, bool STATS, bool PRE>
__global__ void __launch_bounds__(kBlock) fempic_tile_kernel(StepArgs P, uint32_t nblocks) {
    const uint32_t b = xcd_block(blockIdx.x, nblocks);
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    int64_t n = P.ctr->size;
    n = n < 0 ? 0 : (n > P.n_max ? P.n_max : n);
    const int tile = P.tile, tile_words = P.tile / kWave;
    const int64_t tile0 = ((int64_t)b * kWavesPerBlock + wave) * tile;
    const int64_t tile_end = (tile0 + tile < n) ? tile0 + tile : n;
    const int64_t S = P.stride;

    __shared__ unsigned long long s_holes[kWavesPerBlock][kMaxTileWords];
    __shared__ DepositTable s_table;
    if (lane < tile_words) s_holes[wave][lane] = 0ull;
    if (DEPOSIT) table_clear(s_table);
    __syncthreads();

    const unsigned long long below = (1ull << lane) - 1ull;
    int64_t next = tile0 + kWave;  // wave-uniform: first particle no lane has taken yet
    int64_t p = tile0 + lane;
    bool have = false, walking = false, removed = false, fault = false;
    int cell = 0, h = 0;
    double x = 0.0, y = 0.0, z = 0.0, l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
    long long hsum = 0;
    int hmax = 0, nfault = 0;

    auto begin = [&](int64_t q) {
        have = true;
        walking = false;
        removed = false;
        fault = false;
        h = 0;
        cell = P.p2c[q];
        if (PRE && cell == kMaxCell) {  // removed by the relocation pass
            removed = true;
            return;
        }
        if (!((unsigned)cell < (unsigned)P.n_cells)) {
            fault = true;
            return;
        }
        x = P.pos[q];
        y = P.pos[S + q];
        z = P.pos[2 * S + q];
        if (PUSH) {
            const double* e = P.ef + 3 * (int64_t)cell;
            const double ex = e[0], ey = e[1], ez = e[2];
            double vx = P.vel[q], vy = P.vel[S + q], vz = P.vel[2 * S + q];
            vx = fma(P.k, ex, vx);
            x = fma(vx, P.dt, x);
            vy = fma(P.k, ey, vy);
            y = fma(vy, P.dt, y);
            vz = fma(P.k, ez, vz);
            z = fma(vz, P.dt, z);
            P.vel[q] = vx; P.vel[S + q] = vy; P.vel[2 * S + q] = vz;
            P.pos[q] = x; P.pos[S + q] = y; P.pos[2 * S + q] = z;
        }
        if (DH) {
            cell = dh_lookup(P, x, y, z);
            if (cell == kMaxCell) {
                removed = true;
                return;
            }
            if (!((unsigned)cell < (unsigned)P.n_cells)) {
                fault = true;
                return;
            }
        }
        walking = true;
    };

    if (p < tile_end) begin(p);

    for (;;) {
        if (walking) {
            // one evaluation of the elemental move kernel (apps/app_fempic/kernels.h:106-161)
            const double2* __restrict__ d = reinterpret_cast<const double2*>(P.det + 16 * (int64_t)cell);
            const double2 a0 = d[0], a1 = d[1], a2 = d[2], a3 = d[3], a4 = d[4], a5 = d[5], a6 = d[6], a7 = d[7];
            const double4 au = *reinterpret_cast<const double4*>(P.aux + cell);
            const double c2 = au.x;
            l0 = c2 * fma(-a1.y, z, fma(a1.x, y, fma(-a0.y, x, a0.x)));
            l1 = c2 * fma(-a3.y, z, fma(a3.x, y, fma(-a2.y, x, a2.x)));
            l2 = c2 * fma(-a5.y, z, fma(a5.x, y, fma(-a4.y, x, a4.x)));
            l3 = c2 * fma(-a7.y, z, fma(a7.x, y, fma(-a6.y, x, a6.x)));
            h++;
            if ((l0 > -1e-10) && ((l0 - 1.0) < 1e-10) && (l1 > -1e-10) && ((l1 - 1.0) < 1e-10) && (l2 > -1e-10) &&
                ((l2 - 1.0) < 1e-10) && (l3 > -1e-10) && ((l3 - 1.0) < 1e-10)) {
                walking = false;
            } else {
                const int2 n01 = __builtin_bit_cast(int2, au.y), n23 = __builtin_bit_cast(int2, au.z);
                int nb = n01.x;
                double ml = l0;
                if (l1 < ml) { ml = l1; nb = n01.y; }
                if (l2 < ml) { ml = l2; nb = n23.x; }
                if (l3 < ml) { ml = l3; nb = n23.y; }
                if (nb < 0) {
                    removed = true;
                    cell = kMaxCell;
                    walking = false;
                } else if (nb >= P.n_cells || h >= OPPX_MAX_HOPS) {
                    fault = true;
                    walking = false;
                } else {
                    cell = nb;
                }
            }
        }
        if (have && !walking) {  // retire
            if (!fault) {
                if (h > 0) {
                    P.lc[p] = l0; P.lc[S + p] = l1; P.lc[2 * S + p] = l2; P.lc[3 * S + p] = l3;
                }
                P.p2c[p] = removed ? kMaxCell : cell;
                if (removed) {
                    const int64_t r = p - tile0;
                    atomicOr(&s_holes[wave][r >> 6], 1ull << (r & 63));
                } else if (DEPOSIT) {
                    const int4 m = *reinterpret_cast<const int4*>(P.c2n + 4 * (int64_t)cell);
                    table_add(s_table, P.charge, m.x, l0);
                    table_add(s_table, P.charge, m.y, l1);
                    table_add(s_table, P.charge, m.z, l2);
                    table_add(s_table, P.charge, m.w, l3);
                }
            } else {
                nfault++;
            }
            if (P.hops) P.hops[p] = h;
            if (STATS) {
                hsum += h;
                hmax = hmax > h ? hmax : h;
            }
            have = false;
        }
        const unsigned long long free_mask = __ballot(!have);
        const unsigned long long walk_mask = __ballot(walking);
        const int nfree = __popcll(free_mask);
        if (next < tile_end && (nfree >= P.refill_min || walk_mask == 0ull)) {
            if (!have) {
                const int64_t q = next + __popcll(free_mask & below);
                if (q < tile_end) {
                    p = q;
                    begin(q);
                }
            }
            next += nfree;
        } else if (walk_mask == 0ull && __ballot(have) == 0ull) {
            break;
        }
    }

    __syncthreads();
    int nrem = 0;
    if (threadIdx.x < kWavesPerBlock * tile_words) {
        const int w = threadIdx.x / tile_words, k = threadIdx.x % tile_words;
        const unsigned long long bits = s_holes[w][k];
        P.hole_bits[((int64_t)b * kWavesPerBlock + w) * tile_words + k] = bits;
        nrem = __popcll(bits);
    }
    if (threadIdx.x < kWave) {  // the hole words sit in the first wave's lanes
        const int tot = wave_sum_i32(nrem);
        if (threadIdx.x == 0 && tot) atomicAdd(&P.ctr->removed, tot);
    }
    if (DEPOSIT) table_flush(s_table, P.charge);
    const unsigned long long fm = __ballot(nfault > 0);
    if (fm) {
        const int nf = wave_sum_i32(nfault);
        if (lane == 0) atomicAdd(&P.ctr->faults, nf);
    }
    if (STATS) {
        __shared__ long long s_sum[kWavesPerBlock];
        __shared__ int s_max[kWavesPerBlock];
        const long long hs = wave_sum_i64(hsum);
        const int hm = wave_max_i32(hmax);
        if (lane == 0) { s_sum[wave] = hs; s_max[wave] = hm; }
        __syncthreads();
        if (threadIdx.x == 0) {
            long long t = 0;
            int m = 0;
            for (int q = 0; q < kWavesPerBlock; q++) { t += s_sum[q]; m = m > s_max[q] ? m : s_max[q]; }
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

// Start of a fused step: zero the deposit target and the per-step counters in one launch.
__global__ void step_begin_kernel(oppx_move_counters* __restrict__ ctr, double* __restrict__ charge, int64_t n_nodes) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) {
        ctr->removed = 0;
        ctr->holes_low = 0;
        ctr->max_hops = 0;
        ctr->total_hops = 0;
    }
    if (charge)
        for (int64_t q = j; q < n_nodes; q += (int64_t)gridDim.x * blockDim.x) charge[q] = 0.0;
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

oppx_tuning g_tuning = {256, 16, {0, 0, 0, 0, 0, 0}};

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
    P.tile = g_tuning.tile;
    P.refill_min = g_tuning.refill_min;
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
void launch_pass(const StepArgs& P, bool dh, bool stats, int64_t n_max, hipStream_t s) {
    if constexpr (MOVE) {
        const int64_t per_block = (int64_t)P.tile * kWavesPerBlock;
        const unsigned g = (unsigned)((n_max + per_block - 1) / per_block);
#define OPPX_TILE(DHV, STV) fempic_tile_kernel<PUSH, DHV, DEPOSIT, STV, false><<<g, kBlock, 0, s>>>(P, g)
        if (dh) {
            if (stats) OPPX_TILE(true, true); else OPPX_TILE(true, false);
        } else {
            if (stats) OPPX_TILE(false, true); else OPPX_TILE(false, false);
        }
#undef OPPX_TILE
    } else {
    const unsigned g = grid_for(n_max, kBlock);
#define OPPX_PASS(DHV, STV) fempic_pass_kernel<PUSH, MOVE, DHV, DEPOSIT, STV><<<g, kBlock, 0, s>>>(P, g)
    if (dh) {
        if (stats) OPPX_PASS(true, true); else OPPX_PASS(true, false);
    } else {
        if (stats) OPPX_PASS(false, true); else OPPX_PASS(false, false);
    }
#undef OPPX_PASS
    }
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
    launch_pass<true, false, false>(P, false, false, n_max, as_stream(stream));
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
    launch_pass<false, true, false>(P, dh != nullptr, true, n_max, s);
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
    launch_pass<false, false, true>(P, false, false, n_max, as_stream(stream));
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
    if (n_max == 0) return OPPX_OK;
    P.n_max = n_max;
    P.k = k;
    P.dt = dt;
    P.hole_bits = hole_bits;
    P.hops = hops;
    P.charge = node_charge;
    if (g_tuning.reserved[0] == 1 && dh != nullptr) {
        // two passes: coalesced push + overlay relocation, then the cooperative walk (+ deposit)
        fempic_pass_kernel<true, false, true, false, false><<<grid_for(n_max, kBlock), kBlock, 0, s>>>(
            P, grid_for(n_max, kBlock));
        OPPX_LAUNCH_CHECK("fempic_push_relocate");
        const int64_t per_block = (int64_t)P.tile * kWavesPerBlock;
        const unsigned g = (unsigned)((n_max + per_block - 1) / per_block);
        if (node_charge) fempic_tile_kernel<false, false, true, true, true><<<g, kBlock, 0, s>>>(P, g);
        else fempic_tile_kernel<false, false, false, true, true><<<g, kBlock, 0, s>>>(P, g);
    } else if (node_charge) {
        launch_pass<true, true, true>(P, dh != nullptr, true, n_max, s);
    } else {
        launch_pass<true, true, false>(P, dh != nullptr, true, n_max, s);
    }
    OPPX_LAUNCH_CHECK("fempic_step_fused");
    return OPPX_OK;
}

int oppx_set_tuning(const oppx_tuning* t) {
    OPPX_ARG_CHECK(t != nullptr, "oppx_set_tuning: null");
    OPPX_ARG_CHECK(t->tile >= kWave && t->tile <= kMaxTile && t->tile % kWave == 0,
                   "oppx_set_tuning: tile must be a multiple of 64 in [64, 512]");
    OPPX_ARG_CHECK(t->refill_min >= 1 && t->refill_min <= kWave, "oppx_set_tuning: refill_min in [1, 64]");
    g_tuning = *t;
    return OPPX_OK;
}

int oppx_get_tuning(oppx_tuning* t) {
    OPPX_ARG_CHECK(t != nullptr, "oppx_get_tuning: null");
    *t = g_tuning;
    return OPPX_OK;
}

int oppx_fempic_step_begin(oppx_move_counters* counters, double* node_charge, int64_t n_nodes, void* stream) {
    OPPX_ARG_CHECK(counters != nullptr && n_nodes >= 0, "oppx_fempic_step_begin: bad arguments");
    int64_t g = node_charge ? grid_for(n_nodes, 256) : 1;
    if (g < 1) g = 1;
    if (g > 2048) g = 2048;
    step_begin_kernel<<<(unsigned)g, 256, 0, as_stream(stream)>>>(counters, node_charge, n_nodes);
    OPPX_LAUNCH_CHECK("step_begin");
    return OPPX_OK;
}

}  // extern "C"
