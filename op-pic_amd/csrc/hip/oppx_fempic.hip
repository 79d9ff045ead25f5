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

#include <type_traits>

#include "oppx_internal.h"

namespace oppx {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlock = 1024;  // largest block of the first-hop pass

// Direct-hop overlay as a kernel sees it (one per instance in batched mode).
struct DhView {
    const int32_t* __restrict__ cell;
    int64_t dx, dy, dz;
    double mx, my, mz, inv_gs;
};

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
    DhView dh;
    // gather mode: particle i of the pass reads its state at perm[i] of the src arrays (a pending
    // sort permutation applied on the fly) and writes every field at i of pos/vel/lc/p2c
    const int32_t* __restrict__ perm;
    const double* __restrict__ src_pos;
    const double* __restrict__ src_vel;
    const double* __restrict__ src_lc;
    const int32_t* __restrict__ src_p2c;
    // 48-byte position/velocity records {x, y, z, vx, vy, vz}: src_rec (gather mode) replaces
    // src_pos/src_vel as the input, read at perm[i]; rec_out receives the pass's new position and
    // velocity of particle i at 6i instead of pos/vel (a step before a lazy sort)
    const double* __restrict__ src_rec;
    double* __restrict__ rec_out;
    double k, dt;
    int64_t n_max;
    int tile;        // particles per wave tile (multiple of 64, <= kMaxTile)
    int refill_min;  // free lanes that trigger a refill
    int persistent;  // 1: resident grid taking block tiles from the per-XCD queues
    int n_insts;     // batched mode: instance table size
    const int32_t* __restrict__ cell_inst;       // batched mode: instance of every cell
    const oppx_batch_inst* __restrict__ insts;   // batched mode: per-instance overlay, k, dt
    oppx_move_counters* ctr;
    struct StatSlot* slots;  // nullable: per-block statistics spread over kSlots lines, reduced after the pass
    int32_t* qcount;         // two-pass move: worklist lengths, one per queue
    uint64_t* __restrict__ hole_bits;
    int32_t* __restrict__ hops;
    double* __restrict__ charge;
};

// findStructuredCellIndex3D (opp_lib/include/opp_direct_hop_core.h:410-426) with the x86
// size_t conversion made explicit: an axis is inside iff -1 < t < dims (t in (-1,0) truncates
// to 0); then the overlay owner (kMaxCell = no owner -> remove, opp_lib/include/opp_omp.h:206-212).
__device__ __forceinline__ int dh_lookup(const DhView& D, double x, double y, double z) {
    const double tx = (x - D.mx) * D.inv_gs;
    const double ty = (y - D.my) * D.inv_gs;
    const double tz = (z - D.mz) * D.inv_gs;
    if (!(tx > -1.0 && tx < (double)D.dx && ty > -1.0 && ty < (double)D.dy && tz > -1.0 && tz < (double)D.dz))
        return kMaxCell;
    const int64_t ix = (int64_t)tx, iy = (int64_t)ty, iz = (int64_t)tz;
    return D.cell[ix + D.dx * (iy + D.dy * iz)];
}

// Particle dats are touched once per pass: non-temporal loads and stores keep them from evicting the
// cell records the walk re-reads from L2.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) { __builtin_nontemporal_store(v, p); }
// a 48-byte position/velocity record {x, y, z, vx, vy, vz} as three 16-byte stores; ordinary (cached)
// stores, so that L2 merges a wave's three partial-line writes into whole lines before they leave
typedef double f64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void st_record(double* rec, double x, double y, double z, double vx, double vy, double vz) {
    f64x2* r = reinterpret_cast<f64x2*>(rec);
    r[0] = f64x2{x, y};
    r[1] = f64x2{z, vx};
    r[2] = f64x2{vy, vz};
}

// Per-block statistics go to one of kSlots slots (blockIdx modulo kSlots) instead of the single
// counters record: thousands of blocks adding to one address serialise in one L2 channel. The
// slots are zeroed before the pass and folded into the counters by slots_reduce_kernel after it.
constexpr int kSlots = 256;
struct StatSlot {
    unsigned long long hops;
    int max_hops, removed, faults, pad[3];
};

__device__ __forceinline__ void block_stats(const StepArgs& P, long long hsum, int hmax, int nrem, int nfault) {
    __shared__ long long s_sum[kMaxBlock / kWave];
    __shared__ int s_max[kMaxBlock / kWave], s_rem[kMaxBlock / kWave], s_flt[kMaxBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const long long hs = wave_sum_i64(hsum);
    const int hm = wave_max_i32(hmax), nr = wave_sum_i32(nrem), nf = wave_sum_i32(nfault);
    if (lane == 0) { s_sum[w] = hs; s_max[w] = hm; s_rem[w] = nr; s_flt[w] = nf; }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        int m = 0, r = 0, f = 0;
        for (int q = 0; q < (int)(blockDim.x / kWave); q++) {
            t += s_sum[q];
            m = m > s_max[q] ? m : s_max[q];
            r += s_rem[q];
            f += s_flt[q];
        }
        if (P.slots) {
            StatSlot* sl = P.slots + (blockIdx.x % kSlots);
            if (t) atomicAdd(&sl->hops, (unsigned long long)t);
            if (m) atomicMax(&sl->max_hops, m);
            if (r) atomicAdd(&sl->removed, r);
            if (f) atomicAdd(&sl->faults, f);
        } else {
            if (t) atomicAdd((unsigned long long*)&P.ctr->total_hops, (unsigned long long)t);
            if (m) atomicMax(&P.ctr->max_hops, m);
            if (r) atomicAdd(&P.ctr->removed, r);
            if (f) atomicAdd(&P.ctr->faults, f);
        }
    }
}

__global__ void slots_reset_kernel(StatSlot* slots, int32_t* qcount) {
    const int t = threadIdx.x;  // kSlots threads
    slots[t].hops = 0ull;
    slots[t].max_hops = 0;
    slots[t].removed = 0;
    slots[t].faults = 0;
    if (qcount) qcount[t] = 0;
}

__global__ void slots_reduce_kernel(const StatSlot* slots, oppx_move_counters* ctr) {
    const int t = threadIdx.x;  // kSlots threads
    const StatSlot sl = slots[t];
    const long long hs = wave_sum_i64((long long)sl.hops);
    const int hm = wave_max_i32(sl.max_hops), nr = wave_sum_i32(sl.removed), nf = wave_sum_i32(sl.faults);
    if ((t & (kWave - 1)) == 0) {
        if (hs) atomicAdd((unsigned long long*)&ctr->total_hops, (unsigned long long)hs);
        if (hm) atomicMax(&ctr->max_hops, hm);
        if (nr) atomicAdd(&ctr->removed, nr);
        if (nf) atomicAdd(&ctr->faults, nf);
    }
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
    for (int q = threadIdx.x; q < kTable; q += blockDim.x) {
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
    for (int q = threadIdx.x; q < kTable; q += blockDim.x) {
        const int k = T.key[q];
        if (k >= 0) atomicAdd(charge + k, T.val[q]);
    }
}

template <bool PUSH, bool MOVE, bool DH, bool DEPOSIT, bool STATS>
__global__ void __launch_bounds__(kBlock) fempic_pass_kernel(StepArgs P, uint32_t nblocks) {
    int64_t n = P.ctr->size;
    n = n < 0 ? 0 : (n > P.n_max ? P.n_max : n);  // never beyond the host's bound
    const uint32_t nlive = live_blocks(n, kBlock, nblocks);
    if (blockIdx.x >= nlive) return;  // no block of live particles left for this workgroup
    const uint32_t b = xcd_block(blockIdx.x, nlive);
    const int64_t i = (int64_t)b * kBlock + threadIdx.x;
    const bool live = i < n;
    const int64_t S = P.stride;

    int cell = live ? P.p2c[i] : kMaxCell;
    // a particle already holding MAX_CELL_INDEX is skipped, as the reference loops do
    // (apps_codegen/app_fempic_cg/seq/move_kernel_loop.hpp:92-94)
    const bool skip = live && cell == kMaxCell;
    bool fault = live && !skip && !((unsigned)cell < (unsigned)P.n_cells);
    bool removed = false;
    int h = 0;
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
    __shared__ DepositTable s_table;
    if (DEPOSIT) table_clear(s_table);

    if (live && !skip && !fault && (PUSH || MOVE)) {
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
            const int c = dh_lookup(P.dh, x, y, z);
            P.p2c[i] = c;
            removed = (c == kMaxCell);
            if (!removed && !((unsigned)c < (unsigned)P.n_cells)) fault = true;
        }
        if (MOVE) {
            if (DH) {
                cell = dh_lookup(P.dh, x, y, z);
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
        if (live && !skip && !removed && !fault) {
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
    block_stats(P, STATS ? h : 0, STATS ? h : 0, (MOVE && live && removed) ? 1 : 0, fault ? 1 : 0);
}

// ---------------------------------------------------------------------------------------------
// Two-pass move. The first pass is one lane per particle over the whole set (full occupancy, every
// load independent across lanes): push, overlay lookup and up to `inline_hops` evaluations of the
// elemental move kernel. A particle still moving after that is appended to its XCD's worklist
// (wave-aggregated append) with its current cell in p2c; the second pass walks only those, on a
// resident grid, to completion. Per-particle results are those of the single walk.

enum HopResult { kHopInside = 0, kHopMove = 1, kHopExit = 2, kHopBad = 3 };

// One evaluation of apps/app_fempic/kernels.h:106-161 at `cell`: inside -> kHopInside; otherwise
// cell becomes the face neighbour of the most negative weight (first index wins ties), or
// kMaxCell when that face is on the boundary. In two steps, the cell's records (its 128 B c_det line
// and 32 B aux record) and the evaluation, so that the records can also come from LDS.
struct HopRec {
    double2 a[8];  // c_det: alpha, beta ; gamma, delta of vertices 0..3
    double4 au;    // {c2, c2c[0..1], c2c[2..3], pad}
};

__device__ __forceinline__ void hop_load(const StepArgs& P, int cell, HopRec& R) {
    const double2* __restrict__ d = reinterpret_cast<const double2*>(P.det + 16 * (int64_t)cell);
#pragma unroll
    for (int q = 0; q < 8; q++) R.a[q] = d[q];
    R.au = *reinterpret_cast<const double4*>(P.aux + cell);
}

__device__ __forceinline__ int hop_calc(const StepArgs& P, const HopRec& R, double x, double y, double z, int& cell,
                                        double& l0, double& l1, double& l2, double& l3) {
    const double2 a0 = R.a[0], a1 = R.a[1], a2 = R.a[2], a3 = R.a[3], a4 = R.a[4], a5 = R.a[5], a6 = R.a[6],
                  a7 = R.a[7];
    const double c2 = R.au.x;
    l0 = c2 * fma(-a1.y, z, fma(a1.x, y, fma(-a0.y, x, a0.x)));
    l1 = c2 * fma(-a3.y, z, fma(a3.x, y, fma(-a2.y, x, a2.x)));
    l2 = c2 * fma(-a5.y, z, fma(a5.x, y, fma(-a4.y, x, a4.x)));
    l3 = c2 * fma(-a7.y, z, fma(a7.x, y, fma(-a6.y, x, a6.x)));
    if ((l0 > -1e-10) && ((l0 - 1.0) < 1e-10) && (l1 > -1e-10) && ((l1 - 1.0) < 1e-10) && (l2 > -1e-10) &&
        ((l2 - 1.0) < 1e-10) && (l3 > -1e-10) && ((l3 - 1.0) < 1e-10))
        return kHopInside;
    const int2 n01 = __builtin_bit_cast(int2, R.au.y), n23 = __builtin_bit_cast(int2, R.au.z);
    int nb = n01.x;
    double ml = l0;
    if (l1 < ml) { ml = l1; nb = n01.y; }
    if (l2 < ml) { ml = l2; nb = n23.x; }
    if (l3 < ml) { ml = l3; nb = n23.y; }
    if (nb < 0) {
        cell = kMaxCell;
        return kHopExit;
    }
    if (nb >= P.n_cells) return kHopBad;
    cell = nb;
    return kHopMove;
}

__device__ __forceinline__ int hop_eval(const StepArgs& P, double x, double y, double z, int& cell, double& l0,
                                        double& l1, double& l2, double& l3) {
    HopRec R;
    hop_load(P, cell, R);
    return hop_calc(P, R, x, y, z, cell, l0, l1, l2, l3);
}

// How an evaluation's result ends a walk of the first pass (h evaluations so far): kWalkOn (still moving),
// kWalkIn (inside its cell), kWalkOut (removed), kWalkBad (fault) or kWalkHandOff (to the worklist).
enum WalkState { kWalkOn = 0, kWalkIn = 1, kWalkOut = 2, kWalkBad = 3, kWalkHandOff = 4 };
__device__ __forceinline__ int walk_state(int r, int h, int inline_hops) {
    if (r == kHopInside) return kWalkIn;
    if (r == kHopExit) return kWalkOut;
    if (r == kHopBad || h >= OPPX_MAX_HOPS) return kWalkBad;
    return h >= inline_hops ? kWalkHandOff : kWalkOn;
}

// Cell records shared by a block's walks. Sorted particles cluster in a few cells, and a wave's
// record loads are bound by the addresses the vector memory unit processes (one per lane per 16-byte
// load), so in the first kCacheRounds evaluations each distinct cell of the block is fetched once,
// 160 bytes by ten lanes, into LDS, and every lane of that cell reads it from there. A lane whose cell
// finds no slot (kCacheProbe probes) loads its own records. One round of 64 slots (10.5 KB, inside the
// deposit table's LDS) measured best: a second round's barriers cost more than its loads save.
constexpr int kCacheBits = 6;
constexpr int kCacheSlots = 1 << kCacheBits;
constexpr int kCacheProbe = 16;
constexpr int kCacheRounds = 1;
struct CellCache {
    int key[kCacheSlots];
    double2 rec[kCacheSlots][10];  // c_det (8 pairs), aux (2 pairs)
};

__device__ __forceinline__ int cache_insert(CellCache& C, int cell) {
    uint32_t h = ((uint32_t)cell * 2654435761u) >> (32 - kCacheBits);
    for (int p = 0; p < kCacheProbe; p++) {
        const int k = C.key[h];
        if (k == cell) return (int)h;
        if (k == -1) {
            const int prev = atomicCAS(&C.key[h], -1, cell);
            if (prev == -1 || prev == cell) return (int)h;
        }
        h = (h + 1) & (kCacheSlots - 1);
    }
    return -1;
}

__device__ __forceinline__ void cache_fill(const StepArgs& P, CellCache& C) {
    for (int c = threadIdx.x; c < kCacheSlots * 10; c += blockDim.x) {
        const int sl = c / 10, q = c - 10 * sl;
        const int key = C.key[sl];
        if (key >= 0) {
            const double2* src = q < 8 ? reinterpret_cast<const double2*>(P.det + 16 * (int64_t)key) + q
                                       : reinterpret_cast<const double2*>(P.aux + key) + (q - 8);
            C.rec[sl][q] = *src;
        }
    }
}

__device__ __forceinline__ void cache_read(const CellCache& C, int sl, HopRec& R) {
#pragma unroll
    for (int q = 0; q < 8; q++) R.a[q] = C.rec[sl][q];
    const double2 u0 = C.rec[sl][8], u1 = C.rec[sl][9];
    R.au = double4{u0.x, u0.y, u1.x, u1.y};
}

// LDS of a 256-thread first-pass block: the record staging, the cell cache and the deposit table in turn
// (side by side, 22.8 KB instead of 12 KB, the pass measured 1-3% slower).
union FirstLds {
    DepositTable dep;
    CellCache cache;
};
static_assert(sizeof(CellCache) <= sizeof(DepositTable), "the cell cache lives inside the deposit table's LDS");

// Worklist of the second pass: one list per XCD, `cap` entries each, {particle, hops so far}.
struct WorkList {
    int2* __restrict__ entry;
    int64_t cap;
};

template <int BS, bool PUSH, bool DH, bool DEPOSIT, bool BATCH, bool GATHER>
__global__ void __launch_bounds__(BS) fempic_first_kernel(StepArgs P, uint32_t nblocks, int inline_hops, WorkList W) {
    int64_t n = P.ctr->size;
    n = n < 0 ? 0 : (n > P.n_max ? P.n_max : n);
    // workgroups past the blocks of live particles (the grid is sized by the host's bound) leave at
    // once; the removal-bitmap words past the live count are never read
    const uint32_t nlive = live_blocks(n, BS, nblocks);
    if (blockIdx.x >= nlive) return;
    const uint32_t b = xcd_block(blockIdx.x, nlive);
    const int64_t i = (int64_t)b * BS + threadIdx.x;
    const bool live = i < n;
    const int64_t S = P.stride;
    const int lane = threadIdx.x & (kWave - 1);
    const double* __restrict__ spos = GATHER ? P.src_pos : P.pos;
    const double* __restrict__ svel = GATHER ? P.src_vel : P.vel;
    const int32_t* __restrict__ sp2c = GATHER ? P.src_p2c : P.p2c;
    // streamed particle loads are non-temporal; loads through a permutation are ordinary cached
    // loads, so the neighbours of a gathered word still find its line in L2
    auto ld = [](const auto* p) { return GATHER ? *p : ld_stream(p); };
    int64_t src = i;
    bool fault = false;
    if (GATHER && live) {
        src = P.perm[i];
        if (!((uint64_t)src < (uint64_t)S)) {  // a corrupt permutation entry: report, touch nothing
            fault = true;
            src = i;
        }
    }

    // the sort may have written the sorted p2c into the target already (src_p2c null): read in place
    const bool p2c_in_place = GATHER && P.src_p2c == nullptr;
    int cell = (live && !fault) ? (p2c_in_place ? ld_stream(P.p2c + i) : ld(sp2c + src)) : kMaxCell;
    // a particle already holding MAX_CELL_INDEX is skipped, as the reference loops do
    // (apps_codegen/app_fempic_cg/seq/move_kernel_loop.hpp:92-94)
    const bool skip = live && !fault && cell == kMaxCell;
    fault = fault || (live && !skip && !((unsigned)cell < (unsigned)P.n_cells));
    bool removed = false, pending = false, pushed = false, p2c_done = false, lc_done = false;
    int h = 0;
    double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
    constexpr bool CACHE = BS == kBlock;
    using Lds = std::conditional_t<CACHE, FirstLds, DepositTable>;
    __shared__ Lds s_lds;
    DepositTable& s_table = *reinterpret_cast<DepositTable*>(&s_lds);

    // push: the particle's new position and velocity (or, for a particle this pass does not push, the
    // ones it carries over)
    const bool go = live && !skip && !fault;
    double x = 0.0, y = 0.0, z = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    DhView D = P.dh;
    if (go) {
        double kk = P.k, dt = P.dt;
        if (BATCH) {  // the instance of the particle's cell supplies k, dt and its overlay
            const int in = P.cell_inst[cell];
            if ((unsigned)in < (unsigned)P.n_insts) {
                const oppx_batch_inst* I = P.insts + in;
                kk = I->k;
                dt = I->dt;
                D.cell = I->dh_cell;
                D.dx = I->dims[0]; D.dy = I->dims[1]; D.dz = I->dims[2];
                D.mx = I->min[0]; D.my = I->min[1]; D.mz = I->min[2];
                D.inv_gs = I->inv_gs;
            } else {
                fault = true;
            }
        }
        if (!fault) {
            if (GATHER && P.src_rec) {  // one 48-byte record instead of six SoA words at permuted addresses
                const double2* r = reinterpret_cast<const double2*>(P.src_rec + 6 * src);
                const double2 r0 = r[0], r1 = r[1], r2 = r[2];
                x = r0.x; y = r0.y; z = r1.x; vx = r1.y; vy = r2.x; vz = r2.y;
            } else {
                x = ld(spos + src); y = ld(spos + S + src); z = ld(spos + 2 * S + src);
                if (PUSH) { vx = ld(svel + src); vy = ld(svel + S + src); vz = ld(svel + 2 * S + src); }
            }
            if (PUSH) {
                // calculate_new_pos_vel_kernel: v = fma(k, E, v); x = fma(v, dt, x)
                const double* e = P.ef + 3 * (int64_t)cell;
                const double ex = e[0], ey = e[1], ez = e[2];
                vx = fma(kk, ex, vx);
                x = fma(vx, dt, x);
                vy = fma(kk, ey, vy);
                y = fma(vy, dt, y);
                vz = fma(kk, ez, vz);
                z = fma(vz, dt, z);
                if (!P.rec_out) {
                    st_stream(P.vel + i, vx); st_stream(P.vel + S + i, vy); st_stream(P.vel + 2 * S + i, vz);
                    st_stream(P.pos + i, x); st_stream(P.pos + S + i, y); st_stream(P.pos + 2 * S + i, z);
                }
                pushed = true;
            }
        }
    }
    if ((GATHER || P.rec_out) && live && !pushed) {
        // every field of the output slot is written: what the pass did not produce is carried over
        double pv[6];
        if (GATHER && P.src_rec) {
            for (int d = 0; d < 6; d++) pv[d] = P.src_rec[6 * src + d];
        } else {
            for (int d = 0; d < 3; d++) {
                pv[d] = spos[d * S + src];
                pv[3 + d] = svel[d * S + src];
            }
        }
        if (P.rec_out) {
            x = pv[0]; y = pv[1]; z = pv[2]; vx = pv[3]; vy = pv[4]; vz = pv[5];
        } else {
            for (int d = 0; d < 3; d++) {
                st_stream(P.pos + d * S + i, pv[d]);
                st_stream(P.vel + d * S + i, pv[3 + d]);
            }
        }
    }
    // records out (block-uniform): every live particle's {x, y, z, vx, vy, vz}. With 256-thread blocks the
    // wave's 64 records pass through LDS (the deposit table's storage, not yet in use) so that each of
    // the three 16-byte stores of a lane lands next to its neighbours' (1 KB contiguous per instruction)
    if (P.rec_out) {
        if constexpr (BS * 48 <= (int)sizeof(Lds)) {
            const int wave = threadIdx.x / kWave;
            f64x2* slice = reinterpret_cast<f64x2*>(&s_lds) + wave * (3 * kWave);
            if (live) {
                slice[3 * lane] = f64x2{x, y};
                slice[3 * lane + 1] = f64x2{z, vx};
                slice[3 * lane + 2] = f64x2{vy, vz};
            }
            __syncthreads();
            const int64_t base = (int64_t)b * BS + wave * kWave;  // first particle of the wave
            const int64_t nw = n - base;                          // live particles of the wave
            f64x2* dst = reinterpret_cast<f64x2*>(P.rec_out + 6 * base);
            for (int k = 0; k < 3; k++) {
                const int q = k * kWave + lane;  // pair q of the wave's 192: record q / 3
                if (q / 3 < nw) __builtin_nontemporal_store(slice[q], dst + q);
            }
            __syncthreads();
        } else if (live) {
            st_record(P.rec_out + 6 * i, x, y, z, vx, vy, vz);
        }
    }
    if constexpr (!CACHE) {
        if (DEPOSIT) table_clear(s_table);
    }

    // direct hop
    bool moving = false;
    int walk = kWalkOn;
    if (go && !fault) {
        if (DH && (!BATCH || D.cell != nullptr)) {
            cell = dh_lookup(D, x, y, z);
            if (cell == kMaxCell) removed = true;
            else if (!((unsigned)cell < (unsigned)P.n_cells)) fault = true;
        }
        moving = !removed && !fault;
    }
    // walk: the block's first evaluations through the cell cache (every thread takes part in its barriers),
    // then each lane on its own
    if constexpr (CACHE) {
        CellCache& C = s_lds.cache;
        for (int round = 0; round < kCacheRounds; round++) {
            __syncthreads();  // the LDS's last users (record staging, the previous round's readers) are done
            for (int q = threadIdx.x; q < kCacheSlots; q += BS) C.key[q] = -1;
            if (!__syncthreads_or(moving)) break;  // a barrier (the keys are clear), the same on every thread
            const int sl = moving ? cache_insert(C, cell) : -1;
            __syncthreads();
            cache_fill(P, C);
            __syncthreads();
            if (moving) {
                HopRec R;
                if (sl >= 0) cache_read(C, sl, R);
                else hop_load(P, cell, R);
                const int r = hop_calc(P, R, x, y, z, cell, l0, l1, l2, l3);
                h++;
                walk = walk_state(r, h, inline_hops);
                moving = walk == kWalkOn;
            }
        }
    }
    while (moving) {
        const int r = hop_eval(P, x, y, z, cell, l0, l1, l2, l3);
        h++;
        walk = walk_state(r, h, inline_hops);
        moving = walk == kWalkOn;
    }
    removed = removed || walk == kWalkOut;
    fault = fault || walk == kWalkBad;
    pending = walk == kWalkHandOff;
    if constexpr (CACHE) {
        if (DEPOSIT) {
            __syncthreads();  // the cache's readers are done with the LDS
            table_clear(s_table);
        }
    }
    if (go) {
        if (!fault) {
            if (!pending && !removed && h > 0) {
                st_stream(P.lc + i, l0); st_stream(P.lc + S + i, l1);
                st_stream(P.lc + 2 * S + i, l2); st_stream(P.lc + 3 * S + i, l3);
                lc_done = true;
            }
            st_stream(P.p2c + i, removed ? kMaxCell : cell);  // a pending particle resumes from `cell`
            p2c_done = true;
        }
        if (P.hops) st_stream(P.hops + i, h);
    }
    if (GATHER && live) {
        // (a pending particle's lc is final only after the worklist walk, which overwrites this)
        if (!p2c_done && !p2c_in_place) st_stream(P.p2c + i, sp2c[src]);
        if (!lc_done)
            for (int d = 0; d < 4; d++) st_stream(P.lc + d * S + i, P.src_lc[d * S + src]);
        if (P.hops && !(live && !skip && !fault)) P.hops[i] = h;
    }

    // hand the still-moving particles to worklist (blockIdx mod kSlots): same XCD, spread counters
    const unsigned long long pmask = __ballot(pending);
    if (pmask) {
        const uint32_t q = blockIdx.x % kSlots;
        int base = 0;
        if (lane == 0) base = atomicAdd(&P.qcount[q], __popcll(pmask));
        base = __shfl(base, 0);
        if (pending) {
            const int64_t slot = base + __popcll(pmask & ((1ull << lane) - 1ull));
            if (slot < W.cap) W.entry[q * W.cap + slot] = make_int2((int)i, h);
        }
    }
    if (DEPOSIT) {
        __syncthreads();  // table cleared
        if (live && !skip && !fault && !removed && !pending) {
            const int4 m = *reinterpret_cast<const int4*>(P.c2n + 4 * (int64_t)cell);
            table_add(s_table, P.charge, m.x, l0);
            table_add(s_table, P.charge, m.y, l1);
            table_add(s_table, P.charge, m.z, l2);
            table_add(s_table, P.charge, m.w, l3);
        }
        __syncthreads();
        table_flush(s_table, P.charge);
    }
    // removal bitmap: one word per wave, written for every word the grid covers
    const unsigned long long rmask = __ballot(live && removed);
    if (lane == 0) P.hole_bits[i >> 6] = rmask;
    const bool done = live && !skip && !pending;
    block_stats(P, done ? h : 0, done ? h : 0, (live && removed) ? 1 : 0, fault ? 1 : 0);
}

// ---------------------------------------------------------------------------------------------
// Quad variant of the first pass: four lanes walk one particle. Lane r of a quad owns vertex r of the
// current tet: it loads that vertex's four plane coefficients (32 B of the tet's 128 B c_det line) and
// one 8 B word of the 32 B aux record, and computes lc_r; the quad combines the inside test and the
// first-minimum neighbour choice through quad DPP moves, so every branch is the same in all four lanes.
// Per hop a lane issues 3 loads instead of the one-lane walk's 10, the four lanes of a quad touch one
// line per load, and a wave's longest walk is the longest of 16 particles instead of 64. The push is
// split the same way (lane d < 3 owns component d), so are the deposit's four node adds and the lc stores.
// Same results as fempic_first_kernel.
constexpr int kQuadBlock = 256;                  // threads per block
constexpr int kQuadParts = kQuadBlock / 4;       // particles per block (one 64-bit removal word)

// Value of lane SRC of the quad (DPP quad_perm broadcast); every lane of the quad must be active.
template <int SRC>
__device__ __forceinline__ int quad_bcast_i32(int v) {
    return __builtin_amdgcn_mov_dpp(v, SRC * 0x55, 0xF, 0xF, false);
}
template <int SRC>
__device__ __forceinline__ double quad_bcast(double v) {
    const int2 h = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(quad_bcast_i32<SRC>(h.x), quad_bcast_i32<SRC>(h.y)));
}
// AND over the quad (quad_perm [1,0,3,2] then [2,3,0,1]).
__device__ __forceinline__ int quad_and(int v) {
    v &= __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);
    v &= __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);
    return v;
}
// Bits 0, 4, 8, ..., 60 of m packed into bits 0..15.
__device__ __forceinline__ uint32_t every_fourth_bit(unsigned long long m) {
    m &= 0x1111111111111111ull;
    m = (m | (m >> 3)) & 0x0303030303030303ull;
    m = (m | (m >> 6)) & 0x000F000F000F000Full;
    m = (m | (m >> 12)) & 0x000000FF000000FFull;
    m = (m | (m >> 24)) & 0xFFFFull;
    return (uint32_t)m;
}

// One evaluation of apps/app_fempic/kernels.h:106-161 by the quad at `cell` (quad-uniform): lane r
// returns lc_r in l; the result code and the new cell are the same in all four lanes.
__device__ __forceinline__ int quad_hop(const StepArgs& P, int r, double x, double y, double z, int& cell, double& l) {
    const double2* __restrict__ d = reinterpret_cast<const double2*>(P.det + 16 * (int64_t)cell + 4 * r);
    const double2 ab = d[0], gd = d[1];  // alpha, beta ; gamma, delta of vertex r
    const double w = reinterpret_cast<const double*>(P.aux + cell)[r];  // c2 | c2c[0..1] | c2c[2..3] | pad
    const double c2 = quad_bcast<0>(w);
    l = c2 * fma(-gd.y, z, fma(gd.x, y, fma(-ab.y, x, ab.x)));
    const int in = (l > -1e-10) && ((l - 1.0) < 1e-10);
    if (quad_and(in)) return kHopInside;
    const double l0 = quad_bcast<0>(l), l1 = quad_bcast<1>(l), l2 = quad_bcast<2>(l), l3 = quad_bcast<3>(l);
    const int2 n01 = __builtin_bit_cast(int2, quad_bcast<1>(w)), n23 = __builtin_bit_cast(int2, quad_bcast<2>(w));
    int nb = n01.x;
    double ml = l0;
    if (l1 < ml) { ml = l1; nb = n01.y; }
    if (l2 < ml) { ml = l2; nb = n23.x; }
    if (l3 < ml) { ml = l3; nb = n23.y; }
    if (nb < 0) {
        cell = kMaxCell;
        return kHopExit;
    }
    if (nb >= P.n_cells) return kHopBad;
    cell = nb;
    return kHopMove;
}

template <bool PUSH, bool DH, bool DEPOSIT, bool BATCH, bool GATHER>
__global__ void __launch_bounds__(kQuadBlock) fempic_quad_kernel(StepArgs P, uint32_t nblocks, int inline_hops, WorkList W) {
    int64_t n = P.ctr->size;
    n = n < 0 ? 0 : (n > P.n_max ? P.n_max : n);
    const uint32_t nlive = live_blocks(n, kQuadParts, nblocks);
    if (blockIdx.x >= nlive) return;  // no block of live particles left for this workgroup
    const uint32_t b = xcd_block(blockIdx.x, nlive);
    const int r = threadIdx.x & 3;
    const int64_t i = (int64_t)b * kQuadParts + (threadIdx.x >> 2);
    const bool live = i < n;                     // quad-uniform
    const int64_t S = P.stride;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const double* __restrict__ spos = GATHER ? P.src_pos : P.pos;
    const double* __restrict__ svel = GATHER ? P.src_vel : P.vel;
    const int32_t* __restrict__ sp2c = GATHER ? P.src_p2c : P.p2c;
    auto ld = [](const auto* p) { return GATHER ? *p : ld_stream(p); };
    int64_t src = i;
    bool fault = false;
    if (GATHER && live) {
        src = P.perm[i];
        if (!((uint64_t)src < (uint64_t)S)) {
            fault = true;
            src = i;
        }
    }
    const bool p2c_in_place = GATHER && P.src_p2c == nullptr;
    int cell = (live && !fault) ? (p2c_in_place ? ld_stream(P.p2c + i) : ld(sp2c + src)) : kMaxCell;
    const bool skip = live && !fault && cell == kMaxCell;
    fault = fault || (live && !skip && !((unsigned)cell < (unsigned)P.n_cells));
    bool removed = false, pending = false, pushed = false, p2c_done = false, lc_done = false;
    int h = 0;
    double l = 0.0;
    __shared__ DepositTable s_table;
    __shared__ uint32_t s_bits[kQuadBlock / kWave];
    if (DEPOSIT) table_clear(s_table);

    if (live && !skip && !fault) {
        double kk = P.k, dt = P.dt;
        DhView D = P.dh;
        if (BATCH) {
            const int in = P.cell_inst[cell];
            if ((unsigned)in < (unsigned)P.n_insts) {
                const oppx_batch_inst* I = P.insts + in;
                kk = I->k;
                dt = I->dt;
                D.cell = I->dh_cell;
                D.dx = I->dims[0]; D.dy = I->dims[1]; D.dz = I->dims[2];
                D.mx = I->min[0]; D.my = I->min[1]; D.mz = I->min[2];
                D.inv_gs = I->inv_gs;
            } else {
                fault = true;
            }
        }
        if (!fault) {
            // lane d < 3 owns component d: x_d (and, pushing, v_d = fma(k, E_d, v_d); x_d = fma(v_d, dt, x_d))
            const int d = r < 3 ? r : 2;
            const bool from_rec = GATHER && P.src_rec;
            double xd = from_rec ? P.src_rec[6 * src + d] : ld(spos + d * S + src);
            if (PUSH) {
                const double e = P.ef[3 * (int64_t)cell + d];
                double vd = from_rec ? P.src_rec[6 * src + 3 + d] : ld(svel + d * S + src);
                vd = fma(kk, e, vd);
                xd = fma(vd, dt, xd);
                if (r < 3) {
                    if (P.rec_out) {
                        st_stream(P.rec_out + 6 * i + d, xd);
                        st_stream(P.rec_out + 6 * i + 3 + d, vd);
                    } else {
                        st_stream(P.vel + d * S + i, vd);
                        st_stream(P.pos + d * S + i, xd);
                    }
                }
                pushed = true;
            }
            const double x = quad_bcast<0>(xd), y = quad_bcast<1>(xd), z = quad_bcast<2>(xd);
            if (DH && (!BATCH || D.cell != nullptr)) {
                cell = dh_lookup(D, x, y, z);
                if (cell == kMaxCell) removed = true;
                else if (!((unsigned)cell < (unsigned)P.n_cells)) fault = true;
            }
            while (!removed && !fault) {
                const int res = quad_hop(P, r, x, y, z, cell, l);
                h++;
                if (res == kHopInside) break;
                if (res == kHopExit) { removed = true; break; }
                if (res == kHopBad || h >= OPPX_MAX_HOPS) { fault = true; break; }
                if (h >= inline_hops) { pending = true; break; }
            }
        }
        if (!fault) {
            if (!pending && !removed && h > 0) {
                st_stream(P.lc + r * S + i, l);
                lc_done = true;
            }
            if (r == 0) st_stream(P.p2c + i, removed ? kMaxCell : cell);
            p2c_done = true;
        }
        if (P.hops && r == 0) st_stream(P.hops + i, h);
    }
    if ((GATHER || P.rec_out) && live && !pushed && r < 3) {  // carry over what the pass did not push
        const bool from_rec = GATHER && P.src_rec;
        const double xr = from_rec ? P.src_rec[6 * src + r] : spos[r * S + src];
        const double vr = from_rec ? P.src_rec[6 * src + 3 + r] : svel[r * S + src];
        if (P.rec_out) {
            st_stream(P.rec_out + 6 * i + r, xr);
            st_stream(P.rec_out + 6 * i + 3 + r, vr);
        } else {
            st_stream(P.pos + r * S + i, xr);
            st_stream(P.vel + r * S + i, vr);
        }
    }
    if (GATHER && live) {
        if (!p2c_done && !p2c_in_place && r == 0) st_stream(P.p2c + i, sp2c[src]);
        if (!lc_done) st_stream(P.lc + r * S + i, P.src_lc[r * S + src]);
        if (P.hops && r == 0 && !(live && !skip && !fault)) P.hops[i] = h;
    }

    // still-moving particles go to the worklist (one entry per quad, from its lane 0)
    const bool lead = r == 0;
    const unsigned long long pmask = __ballot(pending && lead);
    if (pmask) {
        const uint32_t q = blockIdx.x % kSlots;
        int base = 0;
        if (lane == 0) base = atomicAdd(&P.qcount[q], __popcll(pmask));
        base = __shfl(base, 0);
        if (pending && lead) {
            const int64_t slot = base + __popcll(pmask & ((1ull << lane) - 1ull));
            if (slot < W.cap) W.entry[q * W.cap + slot] = make_int2((int)i, h);
        }
    }
    if (DEPOSIT) {
        __syncthreads();  // table cleared
        if (live && !skip && !fault && !removed && !pending) {
            const int node = P.c2n[4 * (int64_t)cell + r];
            table_add(s_table, P.charge, node, l);
        }
    }
    // removal bitmap: the block's 64 particles are one word, 16 bits per wave
    const uint32_t bits16 = every_fourth_bit(__ballot(live && removed && lead));
    if (lane == 0) s_bits[wave] = bits16;
    __syncthreads();
    if (DEPOSIT) table_flush(s_table, P.charge);
    if (threadIdx.x == 0) {
        unsigned long long word = 0ull;
        for (int w = 0; w < kQuadBlock / kWave; w++) word |= (unsigned long long)s_bits[w] << (16 * w);
        P.hole_bits[b] = word;
    }
    const bool done = live && !skip && !pending && lead;
    block_stats(P, done ? h : 0, done ? h : 0, (live && removed && lead) ? 1 : 0, (fault && lead) ? 1 : 0);
}

template <bool DEPOSIT>
__global__ void __launch_bounds__(kBlock) fempic_walk_kernel(StepArgs P, WorkList W) {
    const uint32_t q = blockIdx.x % kSlots;
    const uint32_t nb_q = gridDim.x / kSlots;  // blocks draining this queue (grid is a multiple of kSlots)
    const uint32_t k_q = blockIdx.x / kSlots;
    int64_t m = P.qcount[q];
    m = m < 0 ? 0 : (m > W.cap ? W.cap : m);
    const int64_t S = P.stride;
    __shared__ DepositTable s_table;
    if (DEPOSIT) table_clear(s_table);
    __syncthreads();

    long long hsum = 0;
    int hmax = 0, nrem = 0, nfault = 0;
    // block-uniform rounds of kBlock entries; with the deposit, the LDS table is flushed after every
    // round so that it never fills (a full table sends every further add to a global atomic)
    for (int64_t base = (int64_t)k_q * kBlock; base < m; base += (int64_t)nb_q * kBlock) {
        const int64_t e = base + threadIdx.x;
        if (DEPOSIT && base != (int64_t)k_q * kBlock) {
            __syncthreads();  // the previous round's adds are in the table
            table_flush(s_table, P.charge);
            __syncthreads();
            table_clear(s_table);
            __syncthreads();
        }
        if (e >= m) continue;
        const int2 w = W.entry[q * W.cap + e];
        const int64_t i = w.x;
        int h = w.y;
        int cell = P.p2c[i];
        const bool rec = P.rec_out != nullptr;  // the first pass left the pushed position in a record
        const double x = rec ? P.rec_out[6 * i] : P.pos[i];
        const double y = rec ? P.rec_out[6 * i + 1] : P.pos[S + i];
        const double z = rec ? P.rec_out[6 * i + 2] : P.pos[2 * S + i];
        double l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;
        bool removed = false, fault = !((unsigned)cell < (unsigned)P.n_cells);
        while (!fault) {
            const int r = hop_eval(P, x, y, z, cell, l0, l1, l2, l3);
            h++;
            if (r == kHopInside) break;
            if (r == kHopExit) { removed = true; break; }
            if (r == kHopBad || h >= OPPX_MAX_HOPS) fault = true;
        }
        if (!fault) {
            P.lc[i] = l0; P.lc[S + i] = l1; P.lc[2 * S + i] = l2; P.lc[3 * S + i] = l3;
            P.p2c[i] = removed ? kMaxCell : cell;
            if (removed) {
                atomicOr((unsigned long long*)&P.hole_bits[i >> 6], 1ull << (i & 63));
                nrem++;
            } else if (DEPOSIT) {
                const int4 c = *reinterpret_cast<const int4*>(P.c2n + 4 * (int64_t)cell);
                table_add(s_table, P.charge, c.x, l0);
                table_add(s_table, P.charge, c.y, l1);
                table_add(s_table, P.charge, c.z, l2);
                table_add(s_table, P.charge, c.w, l3);
            }
        } else {
            nfault++;
        }
        if (P.hops) P.hops[i] = h;
        hsum += h;
        hmax = hmax > h ? hmax : h;
    }
    __syncthreads();
    if (DEPOSIT) table_flush(s_table, P.charge);
    block_stats(P, hsum, hmax, nrem, nfault);
}

// Zero the per-XCD worklist / queue counters before a launch that uses them.
__global__ void reset_queue_kernel(oppx_move_counters* ctr) {
    if (threadIdx.x < kXcds) ctr->queue[threadIdx.x] = 0;
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

template <bool PUSH, bool DH, bool DEPOSIT, bool STATS, bool PRE, bool BATCH>
__global__ void __launch_bounds__(kBlock) fempic_tile_kernel(StepArgs P, uint32_t nblocks) {
    const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    int64_t n = P.ctr->size;
    n = n < 0 ? 0 : (n > P.n_max ? P.n_max : n);
    const int tile = P.tile, tile_words = P.tile / kWave;
    const int64_t S = P.stride;

    __shared__ unsigned long long s_holes[kWavesPerBlock][kMaxTileWords];
    __shared__ DepositTable s_table;
    __shared__ uint32_t s_next;

    // Block tiles (kWavesPerBlock consecutive wave tiles) are split into one contiguous range per
    // XCD. A plain launch maps block ids onto those ranges; a persistent launch keeps a resident grid
    // whose blocks take the next block tile of their XCD's range from a queue in the counters.
    // Only the block tiles that hold live particles are split, so every XCD gets 1/8 of them.
    const uint32_t xcd = blockIdx.x % kXcds;
    const uint32_t nlive = live_blocks(n, (int64_t)kWavesPerBlock * tile, nblocks);
    const uint32_t q8 = nlive / kXcds, r8 = nlive % kXcds;
    const uint32_t range0 = xcd * q8 + (xcd < r8 ? xcd : r8), range_n = q8 + (xcd < r8 ? 1u : 0u);

    const unsigned long long below = (1ull << lane) - 1ull;
    long long hsum = 0;
    int hmax = 0, nfault = 0, nrem_total = 0;

    for (uint32_t round = 0;; round++) {
        uint32_t b;
        if (P.persistent) {
            if (threadIdx.x == 0) s_next = (uint32_t)atomicAdd(&P.ctr->queue[xcd], 1);
            __syncthreads();
            const uint32_t t = s_next;
            __syncthreads();
            if (t >= range_n) break;  // uniform over the block: every wave leaves here
            b = range0 + t;
        } else {
            if (round > 0 || blockIdx.x >= nlive) break;
            b = xcd_block(blockIdx.x, nlive);
        }
        // block tiles past the live count hold no particle (and later ones of this range neither)
        if ((int64_t)b * kWavesPerBlock * tile >= n) break;
        const int64_t tile0 = ((int64_t)b * kWavesPerBlock + wave) * tile;
        const int64_t tile_end = (tile0 + tile < n) ? tile0 + tile : n;

        if (lane < tile_words) s_holes[wave][lane] = 0ull;
        if (DEPOSIT) table_clear(s_table);
        __syncthreads();

        int64_t next = tile0 + kWave;  // wave-uniform: first particle no lane has taken yet
        int64_t p = tile0 + lane;
        bool have = false, walking = false, removed = false, fault = false, skip = false;
        int cell = 0, h = 0;
        double x = 0.0, y = 0.0, z = 0.0, l0 = 0.0, l1 = 0.0, l2 = 0.0, l3 = 0.0;

        auto begin = [&](int64_t q) {
            have = true;
            walking = false;
            removed = false;
            fault = false;
            skip = false;
            h = 0;
            cell = P.p2c[q];
            if (cell == kMaxCell) {
                if (PRE) removed = true;  // removed by the relocation pass
                else skip = true;         // already removed: skipped, as the reference loops do
                return;
            }
            if (!((unsigned)cell < (unsigned)P.n_cells)) {
                fault = true;
                return;
            }
            double kk = P.k, dt = P.dt;
            DhView D = P.dh;
            if (BATCH) {  // the instance of the particle's cell supplies k, dt and its overlay
                const int in = P.cell_inst[cell];
                if (!((unsigned)in < (unsigned)P.n_insts)) {
                    fault = true;
                    return;
                }
                const oppx_batch_inst* I = P.insts + in;
                kk = I->k;
                dt = I->dt;
                D.cell = I->dh_cell;
                D.dx = I->dims[0]; D.dy = I->dims[1]; D.dz = I->dims[2];
                D.mx = I->min[0]; D.my = I->min[1]; D.mz = I->min[2];
                D.inv_gs = I->inv_gs;
            }
            x = P.pos[q];
            y = P.pos[S + q];
            z = P.pos[2 * S + q];
            if (PUSH) {
                const double* e = P.ef + 3 * (int64_t)cell;
                const double ex = e[0], ey = e[1], ez = e[2];
                double vx = P.vel[q], vy = P.vel[S + q], vz = P.vel[2 * S + q];
                vx = fma(kk, ex, vx);
                x = fma(vx, dt, x);
                vy = fma(kk, ey, vy);
                y = fma(vy, dt, y);
                vz = fma(kk, ez, vz);
                z = fma(vz, dt, z);
                P.vel[q] = vx; P.vel[S + q] = vy; P.vel[2 * S + q] = vz;
                P.pos[q] = x; P.pos[S + q] = y; P.pos[2 * S + q] = z;
            }
            if (DH && (!BATCH || D.cell != nullptr)) {
                cell = dh_lookup(D, x, y, z);
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
                if (skip) {
                    // nothing to write: the particle keeps MAX_CELL_INDEX
                } else if (!fault) {
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
        if (threadIdx.x < kWavesPerBlock * tile_words) {
            const int w = threadIdx.x / tile_words, k = threadIdx.x % tile_words;
            const unsigned long long bits = s_holes[w][k];
            P.hole_bits[((int64_t)b * kWavesPerBlock + w) * tile_words + k] = bits;
            nrem_total += __popcll(bits);
        }
        if (DEPOSIT) table_flush(s_table, P.charge);
        __syncthreads();  // the next round clears the table and the hole words
    }

    block_stats(P, hsum, hmax, nrem_total, nfault);
}

// Per-move reset of the counters (a kernel, not a memset, so the step can be graph-captured).
__global__ void reset_counters_kernel(oppx_move_counters* ctr) {
    ctr->removed = 0;
    ctr->holes_low = 0;
    ctr->max_hops = 0;
    ctr->total_hops = 0;
    for (int x = 0; x < kXcds; x++) ctr->queue[x] = 0;
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
    if (j < kXcds) ctr->queue[j] = 0;
    if (charge)
        for (int64_t q = j; q < n_nodes; q += (int64_t)gridDim.x * blockDim.x) charge[q] = 0.0;
}

// c2c entries are copied as they are: a negative entry is a boundary face (the walk removes the
// particle, as the reference's `opp_c2c[min_i] >= 0` test does, apps/app_fempic/kernels.h:148-159);
// an entry at or past n_cells is a corrupt map, which the walk reports as a fault.
__global__ void tet_aux_kernel(const double* __restrict__ vol, const int32_t* __restrict__ c2c, int64_t n,
                               oppx_tet_aux* __restrict__ aux) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    oppx_tet_aux a;
    a.c2 = (1.0 / 6.0) / vol[c];  // coefficient2 of apps/app_fempic/kernels.h:110, IEEE division
    for (int q = 0; q < 4; q++) a.c2c[q] = c2c[4 * c + q];
    a.pad[0] = a.pad[1] = 0;
    aux[c] = a;
}

// pos/vel SoA from 48-byte records: slot i takes record perm[i] (perm null: record i)
__global__ void __launch_bounds__(kBlock) gather_records_kernel(int64_t n_max, int64_t stride,
                                                                const oppx_move_counters* __restrict__ ctr,
                                                                const int32_t* __restrict__ perm,
                                                                const double* __restrict__ rec, double* __restrict__ pos,
                                                                double* __restrict__ vel, int32_t* __restrict__ faults) {
    int64_t n = ctr->size;
    n = n < 0 ? 0 : (n > n_max ? n_max : n);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const int64_t src = perm ? (int64_t)perm[i] : i;
        if (!((uint64_t)src < (uint64_t)stride)) {  // a corrupt permutation entry: report, touch nothing
            atomicAdd(faults, 1);
            continue;
        }
        const f64x2* r = reinterpret_cast<const f64x2*>(rec + 6 * src);
        const f64x2 a = r[0], b = r[1], c = r[2];
        st_stream(pos + i, a.x); st_stream(pos + stride + i, a.y); st_stream(pos + 2 * stride + i, b.x);
        st_stream(vel + i, b.y); st_stream(vel + stride + i, c.x); st_stream(vel + 2 * stride + i, c.y);
    }
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
    P.tile = g_tuning.tile;
    P.refill_min = g_tuning.refill_min;
    P.persistent = g_tuning.persistent;
    if (dh) {
        OPPX_ARG_CHECK(dh->inv_gs > 0, "fempic: overlay grid spacing must be positive");
        OPPX_ARG_CHECK(dh->cell && dh->dims[0] > 0 && dh->dims[1] > 0 && dh->dims[2] > 0, "fempic: bad overlay");
        P.dh.cell = dh->cell;
        P.dh.dx = dh->dims[0]; P.dh.dy = dh->dims[1]; P.dh.dz = dh->dims[2];
        P.dh.mx = dh->min[0]; P.dh.my = dh->min[1]; P.dh.mz = dh->min[2];
        P.dh.inv_gs = dh->inv_gs;
    }
    return OPPX_OK;
}

// Grid of a tile-kernel launch: one block per block tile, or (persistent) the resident grid,
// a multiple of the XCD count so every XCD's queue has blocks draining it.
unsigned tile_grid(const void* kernel, unsigned block_tiles, bool persistent) {
    if (!persistent) return block_tiles;
    unsigned g = resident_blocks(kernel, kBlock);
    const unsigned need = ((block_tiles + kXcds - 1) / kXcds) * kXcds;
    if (g > need) g = need;
    g = (g / kXcds) * kXcds;
    return g < (unsigned)kXcds ? (unsigned)kXcds : g;
}

template <bool PUSH, bool DH, bool DEPOSIT, bool STATS, bool PRE, bool BATCH>
void launch_tile(const StepArgs& P, int64_t n_max, hipStream_t s) {
    const int64_t per_block = (int64_t)P.tile * kWavesPerBlock;
    const unsigned nbt = (unsigned)((n_max + per_block - 1) / per_block);
    auto kern = fempic_tile_kernel<PUSH, DH, DEPOSIT, STATS, PRE, BATCH>;
    const unsigned g = tile_grid(reinterpret_cast<const void*>(kern), nbt, P.persistent != 0);
    if (P.persistent) reset_queue_kernel<<<1, kWave, 0, s>>>(P.ctr);
    kern<<<g, kBlock, 0, s>>>(P, nbt);
}

// Work buffer of the move: worklist lengths [kSlots] int32, statistics slots [kSlots], then kSlots
// worklists of work_cap(n_max) {particle, hops} entries.
constexpr size_t kWorkHeader = kSlots * sizeof(int32_t) + kSlots * sizeof(StatSlot);

int64_t work_cap(int64_t n_max) {  // per queue, for either first-pass block size
    const int64_t a = ((n_max + kBlock - 1) / kBlock / kSlots + 1) * kBlock;
    const int64_t b = ((n_max + kMaxBlock - 1) / kMaxBlock / kSlots + 1) * kMaxBlock;
    return a > b ? a : b;
}

size_t work_need(int64_t n_max) { return kWorkHeader + (size_t)kSlots * (size_t)work_cap(n_max) * sizeof(int2); }

// Point P's statistics at the slots of the work buffer (when it holds them) and zero them.
bool attach_slots(StepArgs& P, void* work, size_t work_bytes, hipStream_t s) {
    if (!work || work_bytes < kWorkHeader) return false;
    P.qcount = reinterpret_cast<int32_t*>(work);
    P.slots = reinterpret_cast<StatSlot*>(reinterpret_cast<char*>(work) + kSlots * sizeof(int32_t));
    slots_reset_kernel<<<1, kSlots, 0, s>>>(P.slots, P.qcount);
    return true;
}

void fold_slots(const StepArgs& P, hipStream_t s) {
    if (P.slots) slots_reduce_kernel<<<1, kSlots, 0, s>>>(P.slots, P.ctr);
}

// Two-pass move (first-hop pass + worklist walk) when the work buffer is large enough and the
// tuning asks for it; returns false when the caller should use the cooperative tile walk instead.
template <bool PUSH, bool DH, bool DEPOSIT, bool BATCH, bool GATHER = false>
bool launch_two_pass(const StepArgs& P, int64_t n_max, void* work, size_t work_bytes, hipStream_t s) {
    if (!P.slots || g_tuning.inline_hops < 1 || work_bytes < work_need(n_max)) return false;
    WorkList W{reinterpret_cast<int2*>(reinterpret_cast<char*>(work) + kWorkHeader), work_cap(n_max)};
    if (g_tuning.lanes == 4) {
        const unsigned g = grid_for(n_max, kQuadParts);
        fempic_quad_kernel<PUSH, DH, DEPOSIT, BATCH, GATHER><<<g, kQuadBlock, 0, s>>>(P, g, g_tuning.inline_hops, W);
    } else if (g_tuning.first_block == kMaxBlock) {
        const unsigned g = grid_for(n_max, kMaxBlock);
        fempic_first_kernel<kMaxBlock, PUSH, DH, DEPOSIT, BATCH, GATHER><<<g, kMaxBlock, 0, s>>>(P, g, g_tuning.inline_hops, W);
    } else {
        const unsigned g = grid_for(n_max, kBlock);
        fempic_first_kernel<kBlock, PUSH, DH, DEPOSIT, BATCH, GATHER><<<g, kBlock, 0, s>>>(P, g, g_tuning.inline_hops, W);
    }
    auto walk = fempic_walk_kernel<DEPOSIT>;
    unsigned wg = resident_blocks(reinterpret_cast<const void*>(walk), kBlock);
    wg = wg < (unsigned)kSlots ? (unsigned)kSlots : (wg / kSlots) * kSlots;
    walk<<<wg, kBlock, 0, s>>>(P, W);
    return true;
}

template <bool PUSH, bool MOVE, bool DEPOSIT>
void launch_pass(const StepArgs& P, bool dh, bool stats, int64_t n_max, hipStream_t s) {
    if constexpr (MOVE) {
        if (dh) {
            if (stats) launch_tile<PUSH, true, DEPOSIT, true, false, false>(P, n_max, s);
            else launch_tile<PUSH, true, DEPOSIT, false, false, false>(P, n_max, s);
        } else {
            if (stats) launch_tile<PUSH, false, DEPOSIT, true, false, false>(P, n_max, s);
            else launch_tile<PUSH, false, DEPOSIT, false, false, false>(P, n_max, s);
        }
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

template <bool DH, bool BATCH, bool GATHER>
void launch_rec(const StepArgs& P, bool deposit, int64_t n_max, void* work, size_t work_bytes, hipStream_t s) {
    if (deposit) launch_two_pass<true, DH, true, BATCH, GATHER>(P, n_max, work, work_bytes, s);
    else launch_two_pass<true, DH, false, BATCH, GATHER>(P, n_max, work, work_bytes, s);
}

// The fused step with 48-byte position/velocity records on either side (oppx_fempic_step_fused_rec and its
// batched form): argument checks shared by both, then the two-pass move with the gather mode when a pending
// permutation is given. P carries the mesh, particle arrays, push coefficients and (BATCH) instance tables.
template <bool BATCH>
int step_rec(StepArgs& P, const char* who, int64_t n_max, const int32_t* perm, const double* src_rec,
             const double* src_lc, const oppx_fempic_parts* parts, double* rec_out, bool dh, uint64_t* hole_bits,
             void* work, size_t work_bytes, int32_t* hops, double* node_charge, hipStream_t s) {
    const std::string w(who);
    OPPX_ARG_CHECK(perm ? (src_rec && src_lc) : (!src_rec && !src_lc),
                   w + ": src_rec and src_lc go with perm (all three or none)");
    OPPX_ARG_CHECK(perm || rec_out, w + ": neither records in nor records out");
    OPPX_ARG_CHECK(parts->lc && parts->p2c && (rec_out || (parts->pos && parts->vel)) &&
                       (perm || (parts->pos && parts->vel)),
                   w + ": particle arrays missing");
    OPPX_ARG_CHECK(!rec_out || rec_out != src_rec, w + ": rec_out must not be src_rec");
    OPPX_ARG_CHECK(!perm || src_lc != parts->lc, w + ": src_lc and parts->lc must be separate");
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, w + ": n_max beyond capacity");
    OPPX_ARG_CHECK(hole_bits != nullptr, w + ": hole_bits required");
    OPPX_ARG_CHECK(g_tuning.inline_hops >= 1 && work && work_bytes >= work_need(n_max),
                   w + ": needs the two-pass move (tuning.inline_hops >= 1 and its workspace)");
    if (n_max == 0) return OPPX_OK;
    P.n_max = n_max;
    P.hole_bits = hole_bits;
    P.hops = hops;
    P.charge = node_charge;
    P.rec_out = rec_out;
    attach_slots(P, work, work_bytes, s);
    if (perm) {
        P.perm = perm;
        P.src_rec = src_rec;
        P.src_lc = src_lc;
        P.src_pos = nullptr;  // never read: records replace them
        P.src_vel = nullptr;
        P.src_p2c = nullptr;  // the sort wrote the sorted p2c into parts->p2c: read in place
    }
    // batched: each instance brings its own overlay (the instance table), so the DH form is the one built
    if (BATCH || dh) {
        if (perm) launch_rec<true, BATCH, true>(P, node_charge != nullptr, n_max, work, work_bytes, s);
        else launch_rec<true, BATCH, false>(P, node_charge != nullptr, n_max, work, work_bytes, s);
    } else if constexpr (!BATCH) {
        if (perm) launch_rec<false, false, true>(P, node_charge != nullptr, n_max, work, work_bytes, s);
        else launch_rec<false, false, false>(P, node_charge != nullptr, n_max, work, work_bytes, s);
    }
    fold_slots(P, s);
    OPPX_LAUNCH_CHECK(who);
    return OPPX_OK;
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
                     const oppx_dh_grid* dh, oppx_move_counters* counters, uint64_t* hole_bits, void* work,
                     size_t work_bytes, int32_t* hops, void* stream) {
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
    attach_slots(P, work, work_bytes, s);
    const bool two = dh ? launch_two_pass<false, true, false, false>(P, n_max, work, work_bytes, s)
                        : launch_two_pass<false, false, false, false>(P, n_max, work, work_bytes, s);
    if (!two) launch_pass<false, true, false>(P, dh != nullptr, true, n_max, s);
    fold_slots(P, s);
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
                           uint64_t* hole_bits, void* work, size_t work_bytes, int32_t* hops, double* node_charge,
                           void* stream) {
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
    attach_slots(P, work, work_bytes, s);
    bool two;
    if (dh) two = node_charge ? launch_two_pass<true, true, true, false>(P, n_max, work, work_bytes, s)
                              : launch_two_pass<true, true, false, false>(P, n_max, work, work_bytes, s);
    else two = node_charge ? launch_two_pass<true, false, true, false>(P, n_max, work, work_bytes, s)
                           : launch_two_pass<true, false, false, false>(P, n_max, work, work_bytes, s);
    if (two) {
    } else if (g_tuning.split == 1 && dh != nullptr) {
        // two passes: coalesced push + overlay relocation, then the cooperative walk (+ deposit)
        fempic_pass_kernel<true, false, true, false, false><<<grid_for(n_max, kBlock), kBlock, 0, s>>>(
            P, grid_for(n_max, kBlock));
        OPPX_LAUNCH_CHECK("fempic_push_relocate");
        if (node_charge) launch_tile<false, false, true, true, true, false>(P, n_max, s);
        else launch_tile<false, false, false, true, true, false>(P, n_max, s);
    } else if (node_charge) {
        launch_pass<true, true, true>(P, dh != nullptr, true, n_max, s);
    } else {
        launch_pass<true, true, false>(P, dh != nullptr, true, n_max, s);
    }
    fold_slots(P, s);
    OPPX_LAUNCH_CHECK("fempic_step_fused");
    return OPPX_OK;
}

int oppx_fempic_step_fused_gather(int64_t n_max, const int32_t* perm, const oppx_fempic_parts* src,
                                  const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh, const oppx_dh_grid* dh,
                                  double k, double dt, oppx_move_counters* counters, uint64_t* hole_bits, void* work,
                                  size_t work_bytes, int32_t* hops, double* node_charge, void* stream) {
    StepArgs P;
    int rc = fill_args(P, parts, mesh, dh, counters);
    if (rc) return rc;
    OPPX_ARG_CHECK(perm && src && src->pos && src->vel && src->lc,
                   "oppx_fempic_step_fused_gather: permutation and source state required");
    OPPX_ARG_CHECK(src->stride == parts->stride, "oppx_fempic_step_fused_gather: source and target strides differ");
    OPPX_ARG_CHECK(src->pos != parts->pos && src->vel != parts->vel && src->lc != parts->lc && src->p2c != parts->p2c,
                   "oppx_fempic_step_fused_gather: source and target must be separate buffers");
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, "oppx_fempic_step_fused_gather: n_max beyond capacity");
    OPPX_ARG_CHECK(hole_bits != nullptr, "oppx_fempic_step_fused_gather: hole_bits required");
    OPPX_ARG_CHECK(g_tuning.inline_hops >= 1 && work && work_bytes >= work_need(n_max),
                   "oppx_fempic_step_fused_gather: needs the two-pass move (tuning.inline_hops >= 1 and its workspace)");
    hipStream_t s = as_stream(stream);
    if (n_max == 0) return OPPX_OK;
    P.n_max = n_max;
    P.k = k;
    P.dt = dt;
    P.hole_bits = hole_bits;
    P.hops = hops;
    P.charge = node_charge;
    P.perm = perm;
    P.src_pos = src->pos;
    P.src_vel = src->vel;
    P.src_lc = src->lc;
    P.src_p2c = src->p2c;
    attach_slots(P, work, work_bytes, s);
    if (dh) {
        if (node_charge) launch_two_pass<true, true, true, false, true>(P, n_max, work, work_bytes, s);
        else launch_two_pass<true, true, false, false, true>(P, n_max, work, work_bytes, s);
    } else {
        if (node_charge) launch_two_pass<true, false, true, false, true>(P, n_max, work, work_bytes, s);
        else launch_two_pass<true, false, false, false, true>(P, n_max, work, work_bytes, s);
    }
    fold_slots(P, s);
    OPPX_LAUNCH_CHECK("fempic_step_fused_gather");
    return OPPX_OK;
}

int oppx_fempic_step_fused_rec(int64_t n_max, const int32_t* perm, const double* src_rec, const double* src_lc,
                               const oppx_fempic_parts* parts, double* rec_out, const oppx_tet_mesh* mesh,
                               const oppx_dh_grid* dh, double k, double dt, oppx_move_counters* counters,
                               uint64_t* hole_bits, void* work, size_t work_bytes, int32_t* hops, double* node_charge,
                               void* stream) {
    StepArgs P;
    int rc = fill_args(P, parts, mesh, dh, counters);
    if (rc) return rc;
    P.k = k;
    P.dt = dt;
    return step_rec<false>(P, "oppx_fempic_step_fused_rec", n_max, perm, src_rec, src_lc, parts, rec_out, dh != nullptr,
                           hole_bits, work, work_bytes, hops, node_charge, as_stream(stream));
}

int oppx_fempic_batch_step_fused_rec(int64_t n_max, const int32_t* perm, const double* src_rec, const double* src_lc,
                                     const oppx_fempic_parts* parts, double* rec_out, const oppx_tet_mesh* mesh,
                                     const int32_t* cell_inst, const oppx_batch_inst* insts, int32_t n_insts,
                                     oppx_move_counters* counters, uint64_t* hole_bits, void* work, size_t work_bytes,
                                     int32_t* hops, double* node_charge, void* stream) {
    StepArgs P;
    int rc = fill_args(P, parts, mesh, nullptr, counters);
    if (rc) return rc;
    OPPX_ARG_CHECK(cell_inst && insts && n_insts > 0, "oppx_fempic_batch_step_fused_rec: instance tables required");
    P.cell_inst = cell_inst;
    P.insts = insts;
    P.n_insts = n_insts;
    return step_rec<true>(P, "oppx_fempic_batch_step_fused_rec", n_max, perm, src_rec, src_lc, parts, rec_out, true,
                          hole_bits, work, work_bytes, hops, node_charge, as_stream(stream));
}

int oppx_gather_records(int64_t n_max, int64_t stride, oppx_move_counters* counters, const int32_t* perm,
                        const double* rec, double* pos, double* vel, void* stream) {
    OPPX_ARG_CHECK(counters && rec && pos && vel, "oppx_gather_records: null array");
    OPPX_ARG_CHECK(n_max >= 0 && stride >= n_max, "oppx_gather_records: need 0 <= n_max <= stride");
    if (n_max == 0) return OPPX_OK;
    unsigned g = grid_for(n_max, kBlock);
    if (g > 8192) g = 8192;
    gather_records_kernel<<<g, kBlock, 0, as_stream(stream)>>>(n_max, stride, counters, perm, rec, pos, vel,
                                                              &counters->faults);
    OPPX_LAUNCH_CHECK("gather_records");
    return OPPX_OK;
}

int oppx_fempic_batch_step_fused(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                                 const int32_t* cell_inst, const oppx_batch_inst* insts, int32_t n_insts,
                                 oppx_move_counters* counters, uint64_t* hole_bits, void* work, size_t work_bytes,
                                 int32_t* hops, double* node_charge, void* stream) {
    StepArgs P;
    int rc = fill_args(P, parts, mesh, nullptr, counters);
    if (rc) return rc;
    OPPX_ARG_CHECK(cell_inst && insts && n_insts > 0, "oppx_fempic_batch_step_fused: instance tables required");
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, "oppx_fempic_batch_step_fused: n_max beyond capacity");
    OPPX_ARG_CHECK(hole_bits != nullptr, "oppx_fempic_batch_step_fused: hole_bits required");
    if (n_max == 0) return OPPX_OK;
    P.n_max = n_max;
    P.hole_bits = hole_bits;
    P.hops = hops;
    P.charge = node_charge;
    P.cell_inst = cell_inst;
    P.insts = insts;
    P.n_insts = n_insts;
    hipStream_t s = as_stream(stream);
    attach_slots(P, work, work_bytes, s);
    const bool two = node_charge ? launch_two_pass<true, true, true, true>(P, n_max, work, work_bytes, s)
                                 : launch_two_pass<true, true, false, true>(P, n_max, work, work_bytes, s);
    if (!two) {
        if (node_charge) launch_tile<true, true, true, true, false, true>(P, n_max, s);
        else launch_tile<true, true, false, true, false, true>(P, n_max, s);
    }
    fold_slots(P, s);
    OPPX_LAUNCH_CHECK("fempic_batch_step_fused");
    return OPPX_OK;
}

size_t oppx_fempic_work_bytes(int64_t n_max) { return n_max < 0 ? 0 : work_need(n_max); }

int oppx_set_tuning(const oppx_tuning* t) {
    OPPX_ARG_CHECK(t != nullptr, "oppx_set_tuning: null");
    OPPX_ARG_CHECK(t->tile >= kWave && t->tile <= kMaxTile && t->tile % kWave == 0,
                   "oppx_set_tuning: tile must be a multiple of 64 in [64, 512]");
    OPPX_ARG_CHECK(t->refill_min >= 1 && t->refill_min <= kWave, "oppx_set_tuning: refill_min in [1, 64]");
    OPPX_ARG_CHECK((t->split == 0 || t->split == 1) && (t->persistent == 0 || t->persistent == 1),
                   "oppx_set_tuning: split and persistent are 0 or 1");
    OPPX_ARG_CHECK(t->sort_items == 4 || t->sort_items == 8 || t->sort_items == 16 || t->sort_items == 32,
                   "oppx_set_tuning: sort_items must be 4, 8, 16 or 32");
    OPPX_ARG_CHECK(t->inline_hops >= 0 && t->inline_hops <= 64, "oppx_set_tuning: inline_hops in [0, 64]");
    OPPX_ARG_CHECK(t->first_block == kBlock || t->first_block == kMaxBlock, "oppx_set_tuning: first_block 256 or 1024");
    OPPX_ARG_CHECK(t->lanes == 1 || t->lanes == 4, "oppx_set_tuning: lanes (per particle in the first pass) 1 or 4");
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
