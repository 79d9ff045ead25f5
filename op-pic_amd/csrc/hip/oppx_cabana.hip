// CabanaPIC particle move with per-hop current deposit on gfx950 (apps/app_cabanapic/kernels.h:
// 109-353; generated loop apps_codegen/app_cabanapic_cg/hip/move_deposit_kernel_loop.hpp): on the
// first evaluation a Boris push and the streak midpoint test; then streak by streak, each streak's
// current weighted into the 12 accumulator entries of the cell it crosses and the particle handed to
// the face neighbour, until its track ends inside a cell.
//
// One lane per particle on a resident grid; each block walks one contiguous run of 256-particle
// tiles. Deposits go into a workgroup LDS table keyed by cell (ds_add_f64), flushed with one
// device-scope atomic per (cell, entry) every few tiles: particles of one cell share a block, so this
// removes most of the global atomics; a wave whose first evaluations all deposit into one cell sums
// them with shuffles first. The arithmetic is the reference GCC build's, FMA for FMA (the file is built with
// -ffp-contract=off): the interpolated fields and the Boris cross products are FMA chains, the
// half-advance products stay unfused (their second use lies past the errno branch GCC places around
// sqrt), v5 = q*ux*uy*uz/3 is fused into all twelve accumulator terms and v1 = v4*d into both of
// its uses in CALC_J.

#include "oppx_internal.h"

namespace oppx {
namespace {

constexpr int kBlock = 256;
constexpr int kCabTable = 64;  // cells per workgroup table (a 256-particle tile spans a few)
constexpr int kCabProbe = 8;
constexpr int kCabFlushTiles = 4;  // tiles of one block between table flushes

struct CabArgs {
    int64_t n_cells;
    const int32_t* __restrict__ c2c6;
    const double* __restrict__ interp;
    double* __restrict__ acc;
    int64_t stride;
    double* __restrict__ pos;
    double* __restrict__ vel;
    double* __restrict__ streak;
    const double* __restrict__ weight;
    int32_t* __restrict__ p2c;
    double qsp, qdt_2mc, cdx, cdy, cdz;
    int64_t n_max;
    oppx_move_counters* ctr;
    int32_t* __restrict__ hops;
};

template <typename T>
__device__ __forceinline__ T ld_nt(const T* p) { return __builtin_nontemporal_load(p); }
template <typename T>
__device__ __forceinline__ void st_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }

struct CabTable {
    int key[kCabTable];
    double val[kCabTable][12];
};

__device__ __forceinline__ void cab_table_clear(CabTable& T) {
    for (int q = threadIdx.x; q < kCabTable; q += blockDim.x) {
        T.key[q] = -1;
#pragma unroll
        for (int j = 0; j < 12; j++) T.val[q][j] = 0.0;
    }
}

// The 12 values of one streak go to the cell's table entry, or straight to global memory when no
// entry is free within kCabProbe probes.
__device__ __forceinline__ void cab_table_add(CabTable& T, double* __restrict__ acc, int cell, const double* v) {
    uint32_t h = ((uint32_t)cell * 2654435761u) >> 26;  // 6 bits
    for (int p = 0; p < kCabProbe; p++) {
        int k = T.key[h];
        if (k == -1) {
            const int prev = atomicCAS(&T.key[h], -1, cell);
            k = prev == -1 ? cell : prev;
        }
        if (k == cell) {
#pragma unroll
            for (int j = 0; j < 12; j++) atomicAdd(&T.val[h][j], v[j]);
            return;
        }
        h = (h + 1) & (kCabTable - 1);
    }
#pragma unroll
    for (int j = 0; j < 12; j++) atomicAdd(acc + 12 * (int64_t)cell + j, v[j]);
}

// Flushes the used entries and leaves the table cleared for the next tile (the caller separates the
// tile's adds from the flush, and the flush from the next tile's adds, with barriers).
__device__ __forceinline__ void cab_table_flush(CabTable& T, double* __restrict__ acc) {
    for (int e = threadIdx.x; e < kCabTable * 12; e += blockDim.x) {
        const int q = e / 12, j = e - 12 * q;
        const int k = T.key[q];
        if (k >= 0) {
            atomicAdd(acc + 12 * (int64_t)k + j, T.val[q][j]);
            T.val[q][j] = 0.0;
        }
    }
    __syncthreads();  // every entry's values are read before its key is released
    for (int q = threadIdx.x; q < kCabTable; q += blockDim.x) T.key[q] = -1;
}

// CALC_J(u, dA, dB) of apps/app_cabanapic/kernels.h:11-25, with v5 = w / 3 fused into its four uses.
__device__ __forceinline__ void cab_calc_j(double q, double w, double u, double dA, double dB, double* out) {
    const double third = 1.0 / 3.0;
    const double v4 = q * u;
    double v0 = fma(-v4, dA, v4);
    double v1 = fma(v4, dA, v4);
    const double p = 1.0 + dB;
    const double v2 = v0 * p;
    const double v3 = v1 * p;
    const double m = 1.0 - dB;
    v0 = v0 * m;
    v1 = v1 * m;
    out[0] = fma(w, third, v0);
    out[1] = fma(-w, third, v1);
    out[2] = fma(-w, third, v2);
    out[3] = fma(w, third, v3);
}

// weight_current_to_accumulator_kernel (kernels.h:109-136) into v[12]: jfx 0..3, jfy 4..7, jfz 8..11
__device__ __forceinline__ void cab_weight(double q, double ux, double uy, double uz, double dx, double dy, double dz,
                                           double* v) {
    const double w = q * ux * uy * uz;
    cab_calc_j(q, w, ux, dy, dz, v);
    cab_calc_j(q, w, uy, dz, dx, v + 4);
    cab_calc_j(q, w, uz, dx, dy, v + 8);
}

enum CabStreak { kCabOn = 0, kCabEnd = 1, kCabFault = 2 };

// One streak of the track (kernels.h:262-353): the part of the remaining streak s up to the first
// face it reaches (or all of it), its current weighted into out[12] for `cell`, the particle advanced
// to the face, and -- when the track leaves the cell -- handed to the face neighbour (entering it
// exactly on the opposite face). kCabEnd: the track ends inside `cell`; kCabFault: the neighbour index
// is out of range or the walk reached OPPX_MAX_HOPS (`cell` keeps the cell the streak was in).
__device__ __forceinline__ int cab_streak(const CabArgs& P, int& cell, double& px, double& py, double& pz, double& sx,
                                          double& sy, double& sz, double q, int h, double* out) {
    const double dirx = (sx > 0) ? 1.0 : -1.0, diry = (sy > 0) ? 1.0 : -1.0, dirz = (sz > 0) ? 1.0 : -1.0;
    const double f0 = (sx == 0) ? 3.4e38 : (dirx - px) / sx;
    const double f1 = (sy == 0) ? 3.4e38 : (diry - py) / sy;
    const double f2 = (sz == 0) ? 3.4e38 : (dirz - pz) / sz;
    double f = 2;
    int axis = 3;
    if (f0 < f) f = f0, axis = 0;
    if (f1 < f) f = f1, axis = 1;
    if (f2 < f) f = f2, axis = 2;
    f *= 0.5;
    const double dsx = sx * f, dsy = sy * f, dsz = sz * f;
    cab_weight(q, dsx, dsy, dsz, px + dsx, py + dsy, pz + dsz, out);
    sx -= dsx;
    sy -= dsy;
    sz -= dsz;
    px += (dsx + dsx);
    py += (dsy + dsy);
    pz += (dsz + dsz);
    if (axis == 3) return kCabEnd;
    const double d = axis == 0 ? dirx : (axis == 1 ? diry : dirz);
    const int face = axis + (d > 0 ? 3 : 0);
    const int nb = P.c2c6[6 * (int64_t)cell + face];
    // value selects, not stores through a chosen reference (which would put px/py/pz in scratch)
    px = axis == 0 ? -d : px;
    py = axis == 1 ? -d : py;
    pz = axis == 2 ? -d : pz;
    if (!((unsigned)nb < (unsigned)P.n_cells) || h + 1 >= OPPX_MAX_HOPS) return kCabFault;
    cell = nb;
    return kCabOn;
}

__global__ void __launch_bounds__(kBlock) cabana_move_deposit_kernel(CabArgs P) {
    __shared__ CabTable T;
    int64_t n = P.ctr->size;
    n = n < 0 ? 0 : (n > P.n_max ? P.n_max : n);
    const int64_t S = P.stride;
    long long hsum = 0;
    int hmax = 0, nfault = 0;
    cab_table_clear(T);
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    // each block walks one contiguous run of tiles, so consecutive tiles share cells and the table is
    // flushed only every kCabFlushTiles tiles (and after the last)
    const int64_t tiles = (n + kBlock - 1) / kBlock;
    const int64_t per = (tiles + gridDim.x - 1) / gridDim.x;
    const int64_t t_begin = (int64_t)blockIdx.x * per;
    const int64_t t_end = t_begin + per < tiles ? t_begin + per : tiles;
    int since_flush = 0;
    for (int64_t t = t_begin; t < t_end; t++) {
        const int64_t i = t * kBlock + threadIdx.x;
        int h = 0;
        // The first evaluation of every lane (push, then the move inside the cell or the first
        // streak) ends with its deposit in v1st; the wave adds those together before any lane walks
        // on, so v1st is dead during the later streaks.
        double v1st[12];
        int c1st = -1;
        bool active = false, fault = false;
        int cell = kMaxCell;
        double px = 0, py = 0, pz = 0, sx = 0, sy = 0, sz = 0, q = 0;
        if (i < n) {
            cell = ld_nt(P.p2c + i);
            if (cell != kMaxCell) {
                active = true;
                px = ld_nt(P.pos + i); py = ld_nt(P.pos + S + i); pz = ld_nt(P.pos + 2 * S + i);
                sx = ld_nt(P.streak + i); sy = ld_nt(P.streak + S + i); sz = ld_nt(P.streak + 2 * S + i);
                const double ux0 = ld_nt(P.vel + i), uy0 = ld_nt(P.vel + S + i), uz0 = ld_nt(P.vel + 2 * S + i);
                q = ld_nt(P.weight + i) * P.qsp;
                fault = !((unsigned)cell < (unsigned)P.n_cells);
                // first evaluation (OPP_DO_ONCE): Boris push, then either a move inside the cell
                // or the first streak
                if (!fault) {
                    const double* ip = P.interp + 18 * (int64_t)cell;
                    const double ex = ip[0], dexdy = ip[1], dexdz = ip[2], d2exdydz = ip[3];
                    const double ey = ip[4], deydz = ip[5], deydx = ip[6], d2eydzdx = ip[7];
                    const double ez = ip[8], dezdx = ip[9], dezdy = ip[10], d2ezdxdy = ip[11];
                    const double dx = px, dy = py, dz = pz;
                    const double tx = fma(dz, fma(dy, d2exdydz, dexdz), fma(dy, dexdy, ex));
                    const double ty = fma(dx, fma(dz, d2eydzdx, deydx), fma(dz, deydz, ey));
                    const double tz = fma(dy, fma(dx, d2ezdxdy, dezdy), fma(dx, dezdx, ez));
                    const double cbx = fma(dx, ip[13], ip[12]);
                    const double cby = fma(dy, ip[15], ip[14]);
                    const double cbz = fma(dz, ip[17], ip[16]);
                    const double hax = P.qdt_2mc * tx, hay = P.qdt_2mc * ty, haz = P.qdt_2mc * tz;
                    double ux = ux0 + hax, uy = uy0 + hay, uz = uz0 + haz;
                    double v0 = P.qdt_2mc / sqrt(1.0 + fma(ux, ux, fma(uy, uy, uz * uz)));
                    double v1 = fma(cbx, cbx, fma(cby, cby, cbz * cbz));
                    double v2 = (v0 * v0) * v1;
                    double v3 = v0 * fma(v2, fma(v2, 2.0 / 15.0, 1.0 / 3.0), 1.0);
                    double v4 = v3 / fma(v1, v3 * v3, 1.0);
                    v4 = v4 + v4;
                    v0 = fma(v3, fma(uy, cbz, -(uz * cby)), ux);
                    v1 = fma(v3, fma(uz, cbx, -(ux * cbz)), uy);
                    v2 = fma(v3, fma(ux, cby, -(uy * cbx)), uz);
                    ux = fma(v4, fma(v1, cbz, -(v2 * cby)), ux);
                    uy = fma(v4, fma(v2, cbx, -(v0 * cbz)), uy);
                    uz = fma(v4, fma(v0, cby, -(v1 * cbx)), uz);
                    ux = ux + hax;
                    uy = uy + hay;
                    uz = uz + haz;
                    st_nt(P.vel + i, ux);
                    st_nt(P.vel + S + i, uy);
                    st_nt(P.vel + 2 * S + i, uz);
                    v0 = 1.0 / sqrt(1.0 + fma(ux, ux, fma(uy, uy, uz * uz)));
                    ux = (ux * P.cdx) * v0;
                    uy = (uy * P.cdy) * v0;
                    uz = (uz * P.cdz) * v0;
                    v0 = dx + ux;
                    v1 = dy + uy;
                    v2 = dz + uz;
                    v3 = v0 + ux;
                    v4 = v1 + uy;
                    const double v5 = v2 + uz;
                    h = 1;
                    c1st = cell;
                    if (v3 <= 1.0 && v4 <= 1.0 && v5 <= 1.0 && -v3 <= 1.0 && -v4 <= 1.0 && -v5 <= 1.0) {
                        px = v3;
                        py = v4;
                        pz = v5;
                        cab_weight(q, ux, uy, uz, v0, v1, v2, v1st);
                        active = false;  // the track ends inside the cell
                    } else {
                        sx = ux;
                        sy = uy;
                        sz = uz;
                        const int st = cab_streak(P, cell, px, py, pz, sx, sy, sz, q, h, v1st);
                        active = st == kCabOn;
                        fault = st == kCabFault;
                    }
                } else {
                    active = false;
                }
            }
        }
        // First-evaluation deposits: in cell order nearly every wave's particles share one cell, and
        // 64 lanes adding to the same 12 LDS words would serialise; such a wave sums its 12 values
        // with a butterfly and one lane adds them. A wave spanning several cells adds lane by lane.
        const unsigned long long has = __ballot(c1st >= 0);
        if (has) {
            const int lead = __shfl(c1st, __ffsll((long long)has) - 1);
            if ((has & ~__ballot(c1st == lead)) == 0ull) {
                double sum[12];
#pragma unroll
                for (int j = 0; j < 12; j++) {
                    double x = c1st >= 0 ? v1st[j] : 0.0;
#pragma unroll
                    for (int o = 1; o < kWave; o <<= 1) x += __shfl_xor(x, o);
                    sum[j] = x;
                }
                if (lane == 0) cab_table_add(T, P.acc, lead, sum);
            } else if (c1st >= 0) {
                cab_table_add(T, P.acc, c1st, v1st);
            }
        }
        // later streaks: each evaluation deposits one streak and hops across the face it ends on
        while (active) {
            h++;
            double v[12];
            const int c = cell;
            const int st = cab_streak(P, cell, px, py, pz, sx, sy, sz, q, h, v);
            cab_table_add(T, P.acc, c, v);
            active = st == kCabOn;
            fault = st == kCabFault;
        }
        if (c1st >= 0) {  // a particle that was moved
            if (fault) {
                nfault++;
            } else {
                st_nt(P.pos + i, px);
                st_nt(P.pos + S + i, py);
                st_nt(P.pos + 2 * S + i, pz);
                st_nt(P.streak + i, sx);
                st_nt(P.streak + S + i, sy);
                st_nt(P.streak + 2 * S + i, sz);
                st_nt(P.p2c + i, cell);
            }
        } else if (fault) {
            nfault++;  // a cell index out of range on entry
        }
        if (i < n && P.hops) P.hops[i] = h;
        hsum += h;
        hmax = hmax > h ? hmax : h;
        if (++since_flush == kCabFlushTiles || t + 1 == t_end) {  // uniform over the block
            __syncthreads();
            cab_table_flush(T, P.acc);
            __syncthreads();
            since_flush = 0;
        }
    }
    __shared__ long long s_sum[kBlock / kWave];
    __shared__ int s_max[kBlock / kWave], s_flt[kBlock / kWave];
    const int w = threadIdx.x / kWave;
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

__global__ void cabana_reset_kernel(oppx_move_counters* ctr) {
    ctr->removed = 0;
    ctr->holes_low = 0;
    ctr->max_hops = 0;
    ctr->total_hops = 0;
    ctr->faults = 0;
}

}  // namespace
}  // namespace oppx

using namespace oppx;

extern "C" {

int oppx_cabana_move_deposit(int64_t n_max, const oppx_cabana_parts* parts, const oppx_cabana_mesh* mesh, double* acc,
                             double qsp, double qdt_2mc, const double cdt_d[3], oppx_move_counters* counters,
                             int32_t* hops, void* stream) {
    OPPX_ARG_CHECK(parts && mesh && acc && cdt_d && counters, "oppx_cabana_move_deposit: null pointer");
    OPPX_ARG_CHECK(parts->pos && parts->vel && parts->streak && parts->weight && parts->p2c,
                   "oppx_cabana_move_deposit: incomplete particle dats");
    OPPX_ARG_CHECK(mesh->c2c6 && mesh->interp && mesh->n_cells > 0 && mesh->n_cells < kMaxCell,
                   "oppx_cabana_move_deposit: bad mesh");
    OPPX_ARG_CHECK(n_max >= 0 && n_max <= parts->stride, "oppx_cabana_move_deposit: n_max beyond capacity");
    hipStream_t s = as_stream(stream);
    cabana_reset_kernel<<<1, 1, 0, s>>>(counters);
    OPPX_LAUNCH_CHECK("cabana_reset");
    if (n_max == 0) return OPPX_OK;
    CabArgs P{};
    P.n_cells = mesh->n_cells;
    P.c2c6 = mesh->c2c6;
    P.interp = mesh->interp;
    P.acc = acc;
    P.stride = parts->stride;
    P.pos = parts->pos;
    P.vel = parts->vel;
    P.streak = parts->streak;
    P.weight = parts->weight;
    P.p2c = parts->p2c;
    P.qsp = qsp;
    P.qdt_2mc = qdt_2mc;
    P.cdx = cdt_d[0];
    P.cdy = cdt_d[1];
    P.cdz = cdt_d[2];
    P.n_max = n_max;
    P.ctr = counters;
    P.hops = hops;
    unsigned g = resident_blocks(reinterpret_cast<const void*>(cabana_move_deposit_kernel), kBlock);
    const unsigned need = grid_for(n_max, kBlock);
    if (g > need) g = need;
    cabana_move_deposit_kernel<<<g, kBlock, 0, s>>>(P);
    OPPX_LAUNCH_CHECK("cabana_move_deposit");
    return OPPX_OK;
}

}  // extern "C"
