// Oracle: CPU restatement of OP-PIC's particle hot path. TEST INFRASTRUCTURE ONLY: it is the
// checker the GPU path is compared with (tests/, the smoke check, and the CPU-baseline leg of
// bench.py); no product code links or calls it.
//
// Arithmetic follows the reference's GCC -O3 build, which contracts a*b+c into FMAs; the
// contractions are written out here with std::fma and the file is built with
// -ffp-contract=off so nothing else is fused. Each function cites the reference lines it
// restates (paths relative to the OP-PIC source tree).

#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

constexpr int kMaxCell = INT_MAX;  // MAX_CELL_INDEX, opp_lib/include/opp_defs.h:150

enum Status { kDone = 0, kNeedMove = 1, kNeedRemove = 2 };

// Elemental move kernel, apps/app_fempic/kernels.h:106-161 (generated copy:
// apps_codegen/app_fempic_cg/seq/move_kernel_loop.hpp:6-63). GCC contracts
// c2*(d0 - d1*x + d2*y - d3*z) into c2*fnma(d3,z,fma(d2,y,fnma(d1,x,d0))).
inline int fempic_move_eval(const double* pos, double* lc, const double* det, double vol, const int* c2c,
                            int* p2c) {
    const double c2 = (1.0 / 6.0) / vol;
    for (int i = 0; i < 4; i++) {
        const double* d = det + 4 * i;
        lc[i] = c2 * std::fma(-d[3], pos[2], std::fma(d[2], pos[1], std::fma(-d[1], pos[0], d[0])));
    }
    if ((lc[0] > -1e-10) && ((lc[0] - 1.0) < 1e-10) && (lc[1] > -1e-10) && ((lc[1] - 1.0) < 1e-10) &&
        (lc[2] > -1e-10) && ((lc[2] - 1.0) < 1e-10) && (lc[3] > -1e-10) && ((lc[3] - 1.0) < 1e-10))
        return kDone;
    int min_i = 0;
    double min_lc = lc[0];
    for (int i = 1; i < 4; i++)
        if (lc[i] < min_lc) {
            min_lc = lc[i];
            min_i = i;
        }
    if (c2c[min_i] >= 0) {
        *p2c = c2c[min_i];
        return kNeedMove;
    }
    *p2c = kMaxCell;
    return kNeedRemove;
}

// static_cast<size_t>(double) of opp_direct_hop_core.h:413-415 as x86-64 evaluates it:
// values in (-1, 2^64) truncate toward zero; anything else (<= -1, NaN, huge) yields an
// index no grid reaches. Returns false for the latter.
inline bool dh_axis_index(double t, uint64_t* out) {
    if (!(t > -1.0) || !(t < 18446744073709551616.0)) return false;
    *out = (uint64_t)std::trunc(t < 0 ? 0.0 : t);
    return true;
}

// Overlay lookup + omp removal rules: findStructuredCellIndex3D (opp_direct_hop_core.h:410-426),
// opp_part_checkForGlobalMove_util (opp_lib/include/opp_omp.h:203-292, non-MPI branch).
// Returns the new cell or kMaxCell (= remove).
inline int dh_lookup(const double* p, const int* overlay, const int64_t dims[3], const double bmin[3],
                     double inv_gs) {
    uint64_t ix, iy, iz;
    if (!dh_axis_index((p[0] - bmin[0]) * inv_gs, &ix) || !dh_axis_index((p[1] - bmin[1]) * inv_gs, &iy) ||
        !dh_axis_index((p[2] - bmin[2]) * inv_gs, &iz))
        return kMaxCell;
    if (ix >= (uint64_t)dims[0] || iy >= (uint64_t)dims[1] || iz >= (uint64_t)dims[2]) return kMaxCell;
    const uint64_t idx = ix + iy * (uint64_t)dims[0] + iz * (uint64_t)(dims[0] * dims[1]);
    return overlay[idx];
}

int nthreads_of(int threads) { return threads > 0 ? threads : omp_get_max_threads(); }

}  // namespace

extern "C" {

// apps_codegen/app_fempic_cg/seq/calculate_new_pos_vel_kernel_loop.hpp:5-52; elemental
// apps/app_fempic/kernels.h:93-103. coefficient1 = charge/mass*dt; v = fma(k,E,v); x = fma(v,dt,x).
void orc_fempic_push(int64_t n, const double* c_ef, const int* p2c, double* pos, double* vel, double charge,
                     double mass, double dt, int threads) {
    const double k = charge / mass * dt;
#pragma omp parallel for schedule(static) num_threads(nthreads_of(threads))
    for (int64_t i = 0; i < n; i++) {
        const double* e = c_ef + 3 * (int64_t)p2c[i];
        for (int d = 0; d < 3; d++) {
            vel[3 * i + d] = std::fma(k, e[d], vel[3 * i + d]);
            pos[3 * i + d] = std::fma(vel[3 * i + d], dt, pos[3 * i + d]);
        }
    }
}

// apps_codegen/app_fempic_cg/omp/move_kernel_loop.hpp:67-274 (global-move pass :146-170 then the
// multi-hop pass :181-199), status rules opp_lib/include/opp_omp.h:176-199.
int64_t orc_fempic_move(int64_t n, const int* c2c, const double* c_det, const double* c_vol, const int* overlay,
                        const int64_t dims[3], const double bmin[3], double inv_gs, const double* pos, double* lc,
                        int* p2c, int* hops, int threads) {
    int64_t removed = 0;
#pragma omp parallel for schedule(static) reduction(+ : removed) num_threads(nthreads_of(threads))
    for (int64_t i = 0; i < n; i++) {
        int h = 0;
        if (overlay) {
            p2c[i] = dh_lookup(pos + 3 * i, overlay, dims, bmin, inv_gs);
            if (p2c[i] == kMaxCell) removed++;
        }
        if (p2c[i] != kMaxCell) {
            int st;
            do {
                const int c = p2c[i];
                st = fempic_move_eval(pos + 3 * i, lc + 4 * i, c_det + 16 * (int64_t)c, c_vol[c], c2c + 4 * (int64_t)c,
                                      p2c + i);
                h++;
            } while (st == kNeedMove);
            if (st == kNeedRemove) removed++;
        }
        if (hops) hops[i] = h;
    }
    return removed;
}

// seq: apps_codegen/app_fempic_cg/seq/deposit_charge_on_nodes_kernel_loop.hpp:21-60 (particle order);
// omp: apps_codegen/app_fempic_cg/omp/deposit_charge_on_nodes_kernel_loop.hpp:21-90 with the
// thread arrays of opp_lib/include/opp_omp.h:59-138 (thread 0 adds into the dat itself, threads
// 1.. into zeroed copies that are then added in thread order).
void orc_fempic_deposit(int64_t n, int64_t n_nodes, const int* c2n, const int* p2c, const double* lc,
                        double* node_charge, int threads) {
    const int T = nthreads_of(threads);
    if (T <= 1) {
        for (int64_t i = 0; i < n; i++) {
            const int* m = c2n + 4 * (int64_t)p2c[i];
            for (int k = 0; k < 4; k++) node_charge[m[k]] += lc[4 * i + k];
        }
        return;
    }
    std::vector<std::vector<double>> priv(T);
#pragma omp parallel num_threads(T)
    {
        const int t = omp_get_thread_num();
        double* dst = node_charge;
        if (t > 0) {
            priv[t].assign(n_nodes, 0.0);
            dst = priv[t].data();
        }
        const int64_t s = (n * t) / T, e = (n * (t + 1)) / T;
        for (int64_t i = s; i < e; i++) {
            const int* m = c2n + 4 * (int64_t)p2c[i];
            for (int k = 0; k < 4; k++) dst[m[k]] += lc[4 * i + k];
        }
    }
#pragma omp parallel for schedule(static) num_threads(T)
    for (int64_t j = 0; j < n_nodes; j++)
        for (int t = 1; t < T; t++) node_charge[j] += priv[t][j];
}

// particle_hole_fill_core, opp_lib/src/core/opp_particle_organize_core.cpp:124-218: walk the holes
// upward, take the last live record from the back, stop when the two meet.
int64_t orc_holefill(int64_t n, int64_t removed, const int* p2c_marks, void** dats, const int* dims,
                     const int* elem_bytes, int n_dats) {
    if (removed <= 0) return n;
    std::vector<int> marks(p2c_marks, p2c_marks + n);  // the p2c dat is itself moved below
    std::vector<std::pair<int64_t, int64_t>> pairs;
    int64_t filled = 0, back = n - 1;
    for (int64_t j = 0; j < n; j++) {
        if (marks[j] != kMaxCell) continue;
        while (back >= 0 && marks[back] == kMaxCell) back--;
        if (back <= j || filled >= removed) break;
        pairs.emplace_back(j, back);
        back--;
        filled++;
    }
    for (int d = 0; d < n_dats; d++) {
        const size_t rec = (size_t)dims[d] * (size_t)elem_bytes[d];
        char* base = (char*)dats[d];
        for (const auto& pr : pairs) std::memcpy(base + pr.first * rec, base + pr.second * rec, rec);
    }
    return n - removed;
}

// Sort contract of this build: stable by (p2c, index) -- opp_lib/src/core/opp_particle_organize_core.cpp:98-121
// sorts with std::sort (unstable), so only the sorted key array and per-cell multisets are
// reference-defined.
void orc_sort_by_cell(int64_t n, const int* p2c, void** dats, const int* dims, const int* elem_bytes, int n_dats,
                      int64_t* perm) {
    std::vector<int64_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::vector<int> key(p2c, p2c + n);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
    for (int d = 0; d < n_dats; d++) {
        const size_t rec = (size_t)dims[d] * (size_t)elem_bytes[d];
        std::vector<char> tmp((size_t)n * rec);
        const char* src = (const char*)dats[d];
        for (int64_t j = 0; j < n; j++) std::memcpy(tmp.data() + j * rec, src + idx[j] * rec, rec);
        std::memcpy(dats[d], tmp.data(), (size_t)n * rec);
    }
    if (perm)
        for (int64_t j = 0; j < n; j++) perm[j] = idx[j];
}

// One step of the FEM-PIC main loop restricted to the hot path
// (apps_codegen/app_fempic_cg/fempic_opp.cpp:172-193): reset, push, move + hole-fill, deposit.
int64_t orc_fempic_step(int64_t n, int64_t n_nodes, const int* c2n, const int* c2c, const double* c_det,
                        const double* c_vol, const double* c_ef, const int* overlay, const int64_t dims[3],
                        const double bmin[3], double inv_gs, double* pos, double* vel, double* lc, int* p2c,
                        double* node_charge, double charge, double mass, double dt, int threads, int64_t* stats) {
    std::fill(node_charge, node_charge + n_nodes, 0.0);
    orc_fempic_push(n, c_ef, p2c, pos, vel, charge, mass, dt, threads);
    std::vector<int> hops(stats ? n : 0);
    const int64_t removed = orc_fempic_move(n, c2c, c_det, c_vol, overlay, dims, bmin, inv_gs, pos, lc, p2c,
                                            stats ? hops.data() : nullptr, threads);
    if (stats) {
        int64_t tot = 0;
        int mx = 0;
        for (int64_t i = 0; i < n; i++) {
            tot += hops[i];
            mx = std::max(mx, hops[i]);
        }
        stats[0] = removed;
        stats[1] = tot;
        stats[2] = mx;
        stats[3] = 0;
    }
    void* dats[4] = {pos, vel, lc, p2c};
    const int dims_[4] = {3, 3, 4, 1};
    const int eb[4] = {8, 8, 8, 4};
    const int64_t n_new = orc_holefill(n, removed, p2c, dats, dims_, eb, 4);
    orc_fempic_deposit(n_new, n_nodes, c2n, p2c, lc, node_charge, threads);
    return n_new;
}

// CellMapper::generateStructuredMesh, opp_lib/src/seq/opp_direct_hop_seq.cpp:134-265 with the
// all_cell_checker of apps_codegen/app_fempic_cg/seq/move_kernel_loop.hpp:193-217 (first cell in
// index order whose elemental kernel reports MOVE_DONE), grid of opp_direct_hop_core.h:317-336,
// centroid rule of opp_lib/include/opp_defs.h:121-127, corner clamp GET_VERT (opp_defs.h:115).
int orc_dh_overlay_bruteforce(int64_t n_cells, const int* c2c, const double* c_det, const double* c_vol,
                              const double bmin[3], const double bmax[3], double gs, int* overlay, int threads) {
    int64_t dims[3];
    for (int d = 0; d < 3; d++) {
        int64_t a = 0;
        for (double x = bmin[d]; x < bmax[d]; x += gs) a++;
        dims[d] = a + 1;
    }
    const int64_t X = dims[0], XY = dims[0] * dims[1], G = dims[0] * dims[1] * dims[2];
    auto checker = [&](const double p[3]) -> int {
        double lc[4];
        for (int64_t c = 0; c < n_cells; c++) {
            int tmp = (int)c;
            if (fempic_move_eval(p, lc, c_det + 16 * c, c_vol[c], c2c + 4 * c, &tmp) == kDone) return (int)c;
        }
        return kMaxCell;
    };
    std::vector<unsigned char> unresolved(G, 0);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads_of(threads))
    for (int64_t idx = 0; idx < G; idx++) {
        const int64_t dz = idx / XY, dy = (idx % XY) / X, dx = idx % X;
        const double o[3] = {std::fma((double)dx, gs, bmin[0]), std::fma((double)dy, gs, bmin[1]),
                             std::fma((double)dz, gs, bmin[2])};
        double c[3];
        for (int d = 0; d < 3; d++) c[d] = (o[d] + gs <= bmax[d]) ? o[d] + gs * 0.5 : (o[d] + bmax[d]) * 0.5;
        int cid = checker(c);
        if (cid == kMaxCell) { c[0] += gs / 100.0; cid = checker(c); }
        if (cid == kMaxCell) { c[1] += gs / 100.0; cid = checker(c); }
        if (cid == kMaxCell) { c[2] += gs / 100.0; cid = checker(c); }
        overlay[idx] = cid;
        unresolved[idx] = (cid == kMaxCell);
    }
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads_of(threads))
    for (int64_t idx = 0; idx < G; idx++) {
        if (!unresolved[idx]) continue;
        const int64_t dz = idx / XY, dy = (idx % XY) / X, dx = idx % X;
        const double o[3] = {std::fma((double)dx, gs, bmin[0]), std::fma((double)dy, gs, bmin[1]),
                             std::fma((double)dz, gs, bmin[2])};
        auto gv = [&](int d, double v) { return (v > bmax[d]) ? bmax[d] : v; };
        // vertex order of opp_direct_hop_seq.cpp:232-243 (the minimum is order independent)
        int best = kMaxCell;
        for (int kz = 0; kz < 2; kz++)
            for (int kx = 0; kx < 2; kx++)
                for (int ky = 0; ky < 2; ky++) {
                    const double p[3] = {gv(0, kx ? o[0] + gs : o[0]), gv(1, ky ? o[1] + gs : o[1]),
                                         gv(2, kz ? o[2] + gs : o[2])};
                    const int cid = checker(p);
                    if (cid != kMaxCell && cid < best) best = cid;
                }
        if (best != kMaxCell) overlay[idx] = best;
    }
    return 0;
}

// NESO-advection step: update_pos_kernel (apps/app_neso_advection/kernels.h:41-54; generated
// apps_codegen/app_neso_advection_mdir_cg/seq/update_pos_kernel_loop.hpp:5-19) then the move loop
// (kernels.h:56-113; seq/move_kernel_loop.hpp), OPP_DO_ONCE true on the first hop only.
// GCC contracts pos += vel*dt into fma(vel,dt,pos), temp = pos + n*ext into fma(n,ext,pos) and
// FMOD's a - b*trunc(a/b) into fma(-b,trunc(a/b),a).
int64_t orc_neso_step(int64_t n, int nx, int ny, double cell_width, double dt, double ext_x, double ext_y,
                      const int* c2c, const double* c_pos_ll, double* pos, const double* vel, int* mdir, int* p2c,
                      int* hops, int threads) {
    (void)nx;
    (void)ny;
    const double ext[2] = {ext_x, ext_y};
    int64_t total = 0;
#pragma omp parallel for schedule(static) reduction(+ : total) num_threads(nthreads_of(threads))
    for (int64_t i = 0; i < n; i++) {
        double* p = pos + 2 * i;
        for (int d = 0; d < 2; d++) {
            p[d] = std::fma(vel[2 * i + d], dt, p[d]);
            const int off = (int)(std::abs(p[d]) + 2.0);
            const double t = std::fma((double)off, ext[d], p[d]);
            p[d] = std::fma(-ext[d], std::trunc(t / ext[d]), t);
        }
        int h = 0;
        if (p2c[i] != kMaxCell) {
            bool first = true;
            int* md = mdir + 2 * i;
            for (;;) {
                const double* ll = c_pos_ll + 2 * (int64_t)p2c[i];
                const int* cc = c2c + 4 * (int64_t)p2c[i];
                const double dx = p[0] - ll[0];
                const double dy = p[1] - ll[1];
                h++;
                if (first) {
                    const bool xp = ((dx > cell_width) && (ext[0] > 2 * dx - cell_width)) ||
                                    ((dx < 0) && (ext[0] < cell_width - 2 * dx));
                    const bool yp = ((dy > cell_width) && (ext[1] > 2 * dy - cell_width)) ||
                                    ((dy < 0) && (ext[1] < cell_width - 2 * dy));
                    md[0] = xp ? 1 : -1;
                    md[1] = yp ? 1 : -1;
                    first = false;
                }
                if ((dx >= 0.0) && (dx <= cell_width)) {
                    md[0] = 0;
                } else if (md[0] > 0) {
                    p2c[i] = cc[1];
                    continue;
                } else if (md[0] < 0) {
                    p2c[i] = cc[0];
                    continue;
                }
                if ((dy >= 0.0) && (dy <= cell_width)) {
                    md[1] = 0;
                } else if (md[1] > 0) {
                    p2c[i] = cc[3];
                    continue;
                } else if (md[1] < 0) {
                    p2c[i] = cc[2];
                    continue;
                }
                break;
            }
        }
        if (hops) hops[i] = h;
        total += h;
    }
    return total;
}

// ---------------------------------------------------------------------------------------
// Injection and the loops either side of the field solve.

// opp_inc_part_count_with_distribution's relation fill, seq core
// (opp_lib/src/core/opp_increase_part_count_core.cpp:122-146): new particle i of [0, count) gets
// face j, where j advances by one whenever i reaches dist[j] (dist = cumulative counts).
void orc_inc_part_distribution(int64_t start, int64_t count, const int* dist, int n_faces, int* rel) {
    int j = 0;
    for (int64_t i = 0; i < count; i++) {
        if (j < n_faces && i >= dist[j]) j++;
        rel[start + i] = j;
    }
}

// inject_ions_kernel (apps/app_fempic/kernels.h:63-90) in the loop of
// apps_codegen/app_fempic_cg/seq/inject_ions_kernel_loop.hpp: particles [first, first + count),
// p2c = inlet face on entry, random pair k for the k-th of them. GCC contracts
// a*u + b*v + n into fma(a, u, b*v) + n and n*iv - ((q/m)*E)*(dt/2) into fma(-((q/m)*E), dt/2, n*iv).
void orc_fempic_inject(int64_t first, int64_t count, const int* if2c, const double* c_ef, const double* if_u,
                       const double* if_v, const double* if_norm, const double* if_npos, const double* rnd,
                       double* pos, double* vel, int* p2c, double ion_velocity, double charge, double mass,
                       double dt) {
    for (int64_t k = 0; k < count; k++) {
        const int64_t i = first + k;
        const int f = p2c[i];
        const int c = if2c[f];
        double a = rnd[2 * k], b = rnd[2 * k + 1];
        if ((a + b) > 1) {
            a = (1 - a);
            b = (1 - b);
        }
        for (int d = 0; d < 3; d++) {
            pos[3 * i + d] = std::fma(a, if_u[3 * f + d], b * if_v[3 * f + d]) + if_npos[9 * f + d];
            const double v0 = if_norm[3 * f + d] * ion_velocity;
            vel[3 * i + d] = std::fma(-(charge / mass * c_ef[3 * c + d]), 0.5 * dt, v0);
        }
        p2c[i] = c;
    }
}

// compute_node_charge_density_kernel (kernels.h:178-183)
void orc_node_charge_density(int64_t n, double* q, const double* vol, double spwt) {
    for (int64_t i = 0; i < n; i++) q[i] *= (spwt / vol[i]);
}

// compute_electric_field_kernel (kernels.h:186-202): GCC contracts (c1 + c2 + c3 + c4) into
// fma(sd3, p3, fma(sd2, p2, fma(sd0, p0, sd1*p1))).
void orc_electric_field(int64_t n_cells, double* ef, const double* sd, const int* c2n, const double* pot) {
    for (int64_t c = 0; c < n_cells; c++) {
        const double p0 = pot[c2n[4 * c]], p1 = pot[c2n[4 * c + 1]], p2 = pot[c2n[4 * c + 2]],
                     p3 = pot[c2n[4 * c + 3]];
        for (int d = 0; d < 3; d++) {
            double t = std::fma(sd[12 * c + d], p0, sd[12 * c + 3 + d] * p1);
            t = std::fma(sd[12 * c + 6 + d], p2, t);
            t = std::fma(sd[12 * c + 9 + d], p3, t);
            ef[3 * c + d] -= t;
        }
    }
}

// init_boundary_pot_kernel (kernels.h:48-59)
void orc_boundary_potential(int64_t n, const int* type, double wall, double* bnd) {
    for (int64_t i = 0; i < n; i++) bnd[i] = type[i] == 3 ? -1 * wall : 0;
}

// An OPP_MAX global as the seq loops fold it: m = MAX(v, m) in index order
// (get_max_cef_kernel / get_final_max_values_kernel, kernels.h:205-224; MAX of opp_util.h:45).
double orc_max_fold(int64_t n, const double* x, int abs_values, double init) {
    double m = init;
    for (int64_t i = 0; i < n; i++) {
        const double v = abs_values ? std::fabs(x[i]) : x[i];
        m = (v > m) ? v : m;
    }
    return m;
}

// ---------------------------------------------------------------------------------------
// CabanaPIC move with per-hop current deposit.

namespace {

// weight_current_to_accumulator_kernel (apps/app_cabanapic/kernels.h:109-136) with CALC_J (:11-25).
// The reference GCC build fuses v5 = q*ux*uy*uz/3 into each of the twelve v0 += v5 / v1 -= v5 /
// v2 -= v5 / v3 += v5 (it is formed first, and all its uses are additions), and v1 = v4*d into both
// v0 = v4 - v1 and v1 += v4.
void cabana_weight(double* acc, double q, double ux, double uy, double uz, double dx, double dy, double dz) {
    const double w = q * ux * uy * uz;  // v5 = w * (1/3), fused below
    const double third = 1.0 / 3.0;
    auto calc_j = [&](double u, double dA, double dB, double* out) {
        const double v4 = q * u;
        double v0 = std::fma(-v4, dA, v4);
        double v1 = std::fma(v4, dA, v4);
        const double p = 1.0 + dB;
        const double v2 = v0 * p;
        const double v3 = v1 * p;
        const double m = 1.0 - dB;
        v0 = v0 * m;
        v1 = v1 * m;
        out[0] += std::fma(w, third, v0);
        out[1] += std::fma(-w, third, v1);
        out[2] += std::fma(-w, third, v2);
        out[3] += std::fma(w, third, v3);
    };
    calc_j(ux, dy, dz, acc + 0);
    calc_j(uy, dz, dx, acc + 4);
    calc_j(uz, dx, dy, acc + 8);
}

enum { kCabDone = 0, kCabMove = 1 };

// move_deposit_kernel (apps/app_cabanapic/kernels.h:139-353); `first` is OPP_DO_ONCE. Contractions
// of the reference build: the interpolated E and B as fma chains, the Boris scalars and cross
// products as written below; products whose uses straddle a sqrt stay unfused.
int cabana_eval(bool first, double* vel, double* pos, double* streak, double weight, const double* ip, double* acc,
                const int* c2c6, int* cell, double qsp, double qdt_2mc, const double cdt_d[3]) {
    if (first) {
        const double ex = ip[0], dexdy = ip[1], dexdz = ip[2], d2exdydz = ip[3];
        const double ey = ip[4], deydz = ip[5], deydx = ip[6], d2eydzdx = ip[7];
        const double ez = ip[8], dezdx = ip[9], dezdy = ip[10], d2ezdxdy = ip[11];
        double cbx = ip[12];
        const double dcbxdx = ip[13];
        double cby = ip[14];
        const double dcbydy = ip[15];
        double cbz = ip[16];
        const double dcbzdz = ip[17];
        const double dx = pos[0], dy = pos[1], dz = pos[2];
        // hax = qdt_2mc * t stays a product: its second use sits after the errno branch GCC puts
        // around sqrt (-fmath-errno), in another basic block, so neither use is fused
        const double tx = std::fma(dz, std::fma(dy, d2exdydz, dexdz), std::fma(dy, dexdy, ex));
        const double ty = std::fma(dx, std::fma(dz, d2eydzdx, deydx), std::fma(dz, deydz, ey));
        const double tz = std::fma(dy, std::fma(dx, d2ezdxdy, dezdy), std::fma(dx, dezdx, ez));
        cbx = std::fma(dx, dcbxdx, cbx);
        cby = std::fma(dy, dcbydy, cby);
        cbz = std::fma(dz, dcbzdz, cbz);
        const double hax = qdt_2mc * tx, hay = qdt_2mc * ty, haz = qdt_2mc * tz;
        double ux = vel[0] + hax;
        double uy = vel[1] + hay;
        double uz = vel[2] + haz;
        double v0 = qdt_2mc / std::sqrt(1.0 + std::fma(ux, ux, std::fma(uy, uy, uz * uz)));
        double v1 = std::fma(cbx, cbx, std::fma(cby, cby, cbz * cbz));
        double v2 = (v0 * v0) * v1;
        double v3 = v0 * std::fma(v2, std::fma(v2, 2.0 / 15.0, 1.0 / 3.0), 1.0);
        double v4 = v3 / std::fma(v1, v3 * v3, 1.0);
        v4 = v4 + v4;
        v0 = std::fma(v3, std::fma(uy, cbz, -(uz * cby)), ux);
        v1 = std::fma(v3, std::fma(uz, cbx, -(ux * cbz)), uy);
        v2 = std::fma(v3, std::fma(ux, cby, -(uy * cbx)), uz);
        ux = std::fma(v4, std::fma(v1, cbz, -(v2 * cby)), ux);
        uy = std::fma(v4, std::fma(v2, cbx, -(v0 * cbz)), uy);
        uz = std::fma(v4, std::fma(v0, cby, -(v1 * cbx)), uz);
        ux = ux + hax;
        uy = uy + hay;
        uz = uz + haz;
        vel[0] = ux;
        vel[1] = uy;
        vel[2] = uz;
        v0 = 1.0 / std::sqrt(1.0 + std::fma(ux, ux, std::fma(uy, uy, uz * uz)));
        ux = ux * cdt_d[0];
        uy = uy * cdt_d[1];
        uz = uz * cdt_d[2];
        ux = ux * v0;
        uy = uy * v0;
        uz = uz * v0;
        v0 = dx + ux;
        v1 = dy + uy;
        v2 = dz + uz;
        v3 = v0 + ux;
        v4 = v1 + uy;
        const double v5 = v2 + uz;
        if (v3 <= 1.0 && v4 <= 1.0 && v5 <= 1.0 && -v3 <= 1.0 && -v4 <= 1.0 && -v5 <= 1.0) {
            pos[0] = v3;
            pos[1] = v4;
            pos[2] = v5;
            cabana_weight(acc, weight * qsp, ux, uy, uz, v0, v1, v2);
            return kCabDone;
        }
        streak[0] = ux;
        streak[1] = uy;
        streak[2] = uz;
    }
    double s_midx = pos[0], s_midy = pos[1], s_midz = pos[2];
    double s_dispx = streak[0], s_dispy = streak[1], s_dispz = streak[2];
    const double s_dir[3] = {(s_dispx > 0) ? 1.0 : -1.0, (s_dispy > 0) ? 1.0 : -1.0, (s_dispz > 0) ? 1.0 : -1.0};
    const double v0 = (s_dispx == 0) ? 3.4e38 : (s_dir[0] - s_midx) / s_dispx;
    const double v1 = (s_dispy == 0) ? 3.4e38 : (s_dir[1] - s_midy) / s_dispy;
    const double v2 = (s_dispz == 0) ? 3.4e38 : (s_dir[2] - s_midz) / s_dispz;
    double v3 = 2;
    int axis = 3;
    if (v0 < v3) v3 = v0, axis = 0;
    if (v1 < v3) v3 = v1, axis = 1;
    if (v2 < v3) v3 = v2, axis = 2;
    v3 *= 0.5;
    s_dispx *= v3;
    s_dispy *= v3;
    s_dispz *= v3;
    s_midx += s_dispx;
    s_midy += s_dispy;
    s_midz += s_dispz;
    cabana_weight(acc, weight * qsp, s_dispx, s_dispy, s_dispz, s_midx, s_midy, s_midz);
    streak[0] -= s_dispx;
    streak[1] -= s_dispy;
    streak[2] -= s_dispz;
    pos[0] += (s_dispx + s_dispx);
    pos[1] += (s_dispy + s_dispy);
    pos[2] += (s_dispz + s_dispz);
    if (axis != 3) {
        const double d = s_dir[axis];
        pos[axis] = d;
        int face = axis;
        if (d > 0) face += 3;
        *cell = c2c6[face];
        pos[axis] = -d;
        return kCabMove;
    }
    return kCabDone;
}

}  // namespace

// The move_deposit loop of apps_codegen/app_cabanapic_cg/seq/move_deposit_kernel_loop.hpp over
// particles [0, n) in order (pos/vel/streak AoS, 3 per particle; interp 18 and acc 12 per cell;
// c2c6 the six face neighbours {x-, y-, z-, x+, y+, z+}). acc is accumulated, not cleared. Returns
// the total evaluations; hops (nullable) receives them per particle.
int64_t orc_cabana_move_deposit(int64_t n, const int* c2c6, const double* interp, double* acc, double* pos,
                                double* vel, double* streak, const double* weight, int* p2c, double qsp,
                                double qdt_2mc, const double* cdt_d, int* hops) {
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) {
        if (p2c[i] == kMaxCell) {
            if (hops) hops[i] = 0;
            continue;
        }
        bool first = true;
        int h = 0;
        for (;;) {
            const int c = p2c[i];
            const int r = cabana_eval(first, vel + 3 * i, pos + 3 * i, streak + 3 * i, weight[i], interp + 18 * (int64_t)c,
                                      acc + 12 * (int64_t)c, c2c6 + 6 * (int64_t)c, p2c + i, qsp, qdt_2mc, cdt_d);
            first = false;
            h++;
            if (r != kCabMove) break;
        }
        if (hops) hops[i] = h;
        total += h;
    }
    return total;
}

}  // extern "C"
