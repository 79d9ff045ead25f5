// Host-side builders for the FEM-PIC particle hot path:
//   * a box tetrahedral mesh (n_x x n_y x n_z cubes, each cut into the 6 tets that share the
//     cube's main diagonal) with the FEM geometry the mover consumes,
//   * the per-cell electric field of a fixed potential,
//   * seeded particle populations,
//   * the direct-hop overlay grid (structured box -> owning tet).
//
// This file is compiled with -ffp-contract=off: every operation below is one IEEE
// operation unless it is an explicit std::fma, so the arrays are identical on every host
// that runs the same glibc.
//
// Geometry follows the reference mesh loader apps/app_fempic/minifempic_funcs.h:284-443
// (volume sign flip :300-310, alpha/beta/gamma/delta :316-380, face neighbours :388-443)
// and the c_det packing of apps/app_fempic/fempic_misc_mesh_loader.h:119-127.
// The overlay follows opp_lib/include/opp_direct_hop_core.h:317-380 (grid dims) and
// opp_lib/src/seq/opp_direct_hop_seq.cpp:134-265 (box ownership rules).

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "oppx.h"
#include "tet_geometry.h"

using oppx_host::det3;
using oppx_host::tet_contains;

namespace {

// The 6 tets of a unit cube that share the diagonal (0,0,0)-(1,1,1). Tet t walks the
// cube corners along axis order kKuhnPerm[t]: v0=(0,0,0), v1=e_a, v2=e_a+e_b, v3=(1,1,1).
// A point with fractional coordinates f lies in tet t iff f[a] >= f[b] >= f[c].
const int kKuhnPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

inline int64_t node_id(int i, int j, int k, int nx, int ny) {
    return (int64_t)i + (int64_t)(nx + 1) * ((int64_t)j + (int64_t)(ny + 1) * (int64_t)k);
}

// splitmix64 finaliser over a (seed, counter) pair: a counter-based stream, so any
// particle's numbers can be generated independently of the others.
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline uint64_t rng_u64(uint64_t seed, uint64_t ctr) {
    return mix64(mix64(seed + 0x9E3779B97F4A7C15ULL) ^ (ctr * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
}
// uniform in [0,1)
inline double rng_u01(uint64_t seed, uint64_t ctr) {
    return (double)(rng_u64(seed, ctr) >> 11) * (1.0 / 9007199254740992.0);
}
// standard normal by Box-Muller (cosine branch), u1 in (0,1]
inline double rng_normal(uint64_t seed, uint64_t ctr) {
    const double u1 = 1.0 - rng_u01(seed, 2 * ctr);
    const double u2 = rng_u01(seed, 2 * ctr + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// Numbering of a structured n0 x n1 x n2 grid: rank[i + n0*(j + n1*k)]. Order 0 is that linear
// index itself (x fastest); order 1 is the Morton (bit-interleaved) order of (i, j, k), which keeps
// spatial neighbours close in index -- and in memory -- in all three directions.
std::vector<int64_t> grid_rank(int64_t n0, int64_t n1, int64_t n2, int order) {
    const int64_t n = n0 * n1 * n2;
    std::vector<int64_t> rank(n);
    if (order == 0) {
        for (int64_t q = 0; q < n; q++) rank[q] = q;
        return rank;
    }
    auto spread = [](uint64_t v) {  // 21 bits -> every third bit
        uint64_t x = v & 0x1FFFFFull;
        x = (x | x << 32) & 0x1F00000000FFFFull;
        x = (x | x << 16) & 0x1F0000FF0000FFull;
        x = (x | x << 8) & 0x100F00F00F00F00Full;
        x = (x | x << 4) & 0x10C30C30C30C30C3ull;
        x = (x | x << 2) & 0x1249249249249249ull;
        return x;
    };
    std::vector<std::pair<uint64_t, int64_t>> key(n);
    for (int64_t k = 0; k < n2; k++)
        for (int64_t j = 0; j < n1; j++)
            for (int64_t i = 0; i < n0; i++) {
                const int64_t q = i + n0 * (j + n1 * k);
                key[q] = {spread((uint64_t)i) | spread((uint64_t)j) << 1 | spread((uint64_t)k) << 2, q};
            }
    std::sort(key.begin(), key.end());
    for (int64_t r = 0; r < n; r++) rank[key[r].second] = r;
    return rank;
}

}  // namespace

extern "C" {

int oppx_box_mesh_sizes(int nx, int ny, int nz, int64_t* n_nodes, int64_t* n_cells) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !n_nodes || !n_cells) return OPPX_ERR_ARG;
    *n_nodes = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    *n_cells = (int64_t)6 * nx * ny * nz;
    if (*n_cells > INT_MAX - 1) return OPPX_ERR_ARG;
    return OPPX_OK;
}

int oppx_box_mesh_build(int nx, int ny, int nz, double lx, double ly, double lz, int order,
                        double* node_pos, int* c2n, int* c2c, double* c_det, double* c_vol) {
    int64_t nn = 0, nc = 0;
    int rc = oppx_box_mesh_sizes(nx, ny, nz, &nn, &nc);
    if (rc != OPPX_OK) return rc;
    if (!node_pos || !c2n || !c2c || !c_det || !c_vol || !(lx > 0) || !(ly > 0) || !(lz > 0) ||
        (order != 0 && order != 1))
        return OPPX_ERR_ARG;
    const std::vector<int64_t> nrank = grid_rank(nx + 1, ny + 1, nz + 1, order);
    const std::vector<int64_t> crank = grid_rank(nx, ny, nz, order);
    auto nid = [&](int i, int j, int k) { return nrank[node_id(i, j, k, nx, ny)]; };

    // nodes: x = lx * i / nx (one rounding of the exact quotient, then one product)
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= nz; k++)
        for (int j = 0; j <= ny; j++)
            for (int i = 0; i <= nx; i++) {
                const int64_t n = nid(i, j, k);
                node_pos[3 * n + 0] = lx * ((double)i / (double)nx);
                node_pos[3 * n + 1] = ly * ((double)j / (double)ny);
                node_pos[3 * n + 2] = lz * ((double)k / (double)nz);
            }

    // cells: cube (ci,cj,ck) of rank r owns tets 6r .. 6r+5
#pragma omp parallel for schedule(static)
    for (int ck = 0; ck < nz; ck++)
        for (int cj = 0; cj < ny; cj++)
            for (int ci = 0; ci < nx; ci++) {
                const int64_t cube = crank[ci + (int64_t)nx * (cj + (int64_t)ny * ck)];
                for (int t = 0; t < 6; t++) {
                    const int64_t c = 6 * cube + t;
                    int off[3] = {0, 0, 0};
                    int con[4];
                    con[0] = (int)nid(ci, cj, ck);
                    for (int s = 0; s < 3; s++) {
                        off[kKuhnPerm[t][s]] = 1;
                        con[s + 1] = (int)nid(ci + off[0], cj + off[1], ck + off[2]);
                    }
                    oppx_host::tet_geometry(node_pos, con, c_det + 16 * c, c_vol + c);
                    for (int v = 0; v < 4; v++) c2n[4 * c + v] = con[v];
                }
            }

    oppx_host::face_neighbours(nn, nc, c2n, c2c);
    return OPPX_OK;
}

int oppx_node_cell_csr(int64_t n_cells, int64_t n_nodes, const int* c2n, int* node_start, int* node_ref) {
    // For every node, the entries c*4+k with c2n[4c+k] == node, in ascending (c, k) order: the
    // fixed summation order of the sorted FP32 deposit.
    if (n_cells <= 0 || n_nodes <= 0 || !c2n || !node_start || !node_ref) return OPPX_ERR_ARG;
    if (4 * n_cells >= (int64_t)INT_MAX) return OPPX_ERR_ARG;
    std::vector<int64_t> cnt(n_nodes + 1, 0);
    for (int64_t e = 0; e < 4 * n_cells; e++) {
        const int v = c2n[e];
        if (v < 0 || v >= n_nodes) return OPPX_ERR_ARG;
        cnt[v + 1]++;
    }
    for (int64_t v = 0; v < n_nodes; v++) cnt[v + 1] += cnt[v];
    for (int64_t v = 0; v <= n_nodes; v++) node_start[v] = (int)cnt[v];
    for (int64_t e = 0; e < 4 * n_cells; e++) node_ref[cnt[c2n[e]]++] = (int)e;
    return OPPX_OK;
}

int oppx_box_field_sines(int64_t n_nodes, int64_t n_cells, const double* node_pos, const int* c2n,
                         const double* c_det, const double* c_vol, double* c_ef) {
    // E = -grad(phi_h), phi_h the P1 interpolant of phi = sin(pi x) sin(pi y) sin(pi z).
    // With lc_i = (a_i - b_i x + g_i y - d_i z)/(6V): grad lc_i = (-b_i, g_i, -d_i)/(6V).
    if (!node_pos || !c2n || !c_det || !c_vol || !c_ef || n_nodes <= 0 || n_cells <= 0) return OPPX_ERR_ARG;
    const double pi = 3.141592653589793;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; c++) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int v = 0; v < 4; v++) {
            const int64_t n = c2n[4 * c + v];
            const double* p = node_pos + 3 * n;
            const double phi = std::sin(pi * p[0]) * std::sin(pi * p[1]) * std::sin(pi * p[2]);
            const double* d = c_det + (4 * c + v) * 4;
            gx = gx + phi * (-d[1]);
            gy = gy + phi * d[2];
            gz = gz + phi * (-d[3]);
        }
        const double s = 1.0 / (6.0 * c_vol[c]);
        c_ef[3 * c + 0] = -(gx * s);
        c_ef[3 * c + 1] = -(gy * s);
        c_ef[3 * c + 2] = -(gz * s);
    }
    return OPPX_OK;
}

int oppx_box_particles_uniform(int64_t n_part, uint64_t seed, double sigma, int nx, int ny, int nz,
                               double lx, double ly, double lz, int order, double* pos, double* vel, int* p2c) {
    // Uniform positions in the box, N(0, sigma^2) velocity components, and the containing
    // tet from the closed form of the cube split (ordering of the fractional coordinates).
    if (n_part < 0 || !pos || !vel || !p2c || nx <= 0 || ny <= 0 || nz <= 0 || (order != 0 && order != 1))
        return OPPX_ERR_ARG;
    const std::vector<int64_t> crank = grid_rank(nx, ny, nz, order);
    const int n[3] = {nx, ny, nz};
    const double L[3] = {lx, ly, lz};
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n_part; p++) {
        double f[3];
        int ci[3];
        for (int d = 0; d < 3; d++) {
            const double u = rng_u01(seed, (uint64_t)p * 8 + d);
            pos[3 * p + d] = L[d] * u;
            const double s = u * (double)n[d];
            int c = (int)s;
            if (c >= n[d]) c = n[d] - 1;
            ci[d] = c;
            f[d] = s - (double)c;
            vel[3 * p + d] = sigma * rng_normal(seed ^ 0x5851F42D4C957F2DULL, (uint64_t)p * 4 + d);
        }
        int t = 0;
        for (int q = 0; q < 6; q++) {
            const int a = kKuhnPerm[q][0], b = kKuhnPerm[q][1], cc = kKuhnPerm[q][2];
            if (f[a] >= f[b] && f[b] >= f[cc]) { t = q; break; }
        }
        const int64_t cube = crank[ci[0] + (int64_t)nx * (ci[1] + (int64_t)ny * ci[2])];
        p2c[p] = (int)(6 * cube + t);
    }
    return OPPX_OK;
}

int oppx_box12_particles(int64_t n_part, uint64_t seed, double sigma, int nx, int ny, int nz, double lx, double ly,
                         double lz, const double* c_det, const double* c_vol, double* pos, double* vel, int* p2c) {
    // Uniform positions in the box of a 12-tet box mesh (tets 12q .. 12q+11 fill cube q, cubes x fastest),
    // N(0, sigma^2) velocity components (the streams of oppx_box_particles_uniform), and the containing
    // tet: the first of the cube's 12 that the mover's inside test accepts.
    if (n_part < 0 || !pos || !vel || !p2c || !c_det || !c_vol || nx <= 0 || ny <= 0 || nz <= 0) return OPPX_ERR_ARG;
    const int n[3] = {nx, ny, nz};
    const double L[3] = {lx, ly, lz};
    int64_t lost = 0;
#pragma omp parallel for schedule(static) reduction(+ : lost)
    for (int64_t p = 0; p < n_part; p++) {
        int ci[3];
        for (int d = 0; d < 3; d++) {
            const double u = rng_u01(seed, (uint64_t)p * 8 + d);
            pos[3 * p + d] = L[d] * u;
            int c = (int)(u * (double)n[d]);
            ci[d] = c >= n[d] ? n[d] - 1 : c;
            vel[3 * p + d] = sigma * rng_normal(seed ^ 0x5851F42D4C957F2DULL, (uint64_t)p * 4 + d);
        }
        const int64_t cube = ci[0] + (int64_t)nx * (ci[1] + (int64_t)ny * ci[2]);
        int t = -1;
        for (int k = 0; k < 12 && t < 0; k++) {
            const int64_t c = 12 * cube + k;
            if (tet_contains(c_det + 16 * c, c_vol[c], pos[3 * p], pos[3 * p + 1], pos[3 * p + 2])) t = (int)c;
        }
        if (t < 0) lost++;
        p2c[p] = t < 0 ? (int)(12 * cube) : t;
    }
    return lost ? OPPX_ERR_STATE : OPPX_OK;
}

int oppx_dh_grid_dims(const double bmin[3], const double bmax[3], double gs, int64_t dims[3]) {
    // CellMapper::init_host (opp_direct_hop_core.h:317-336): dims = (#steps of a floating
    // accumulation that stays below max) + 1 per axis.
    if (!bmin || !bmax || !dims || !(gs > 0)) return OPPX_ERR_ARG;
    for (int d = 0; d < 3; d++) {
        int64_t a = 0;
        for (double x = bmin[d]; x < bmax[d]; x += gs) {
            a++;
            if (a > (1 << 20)) return OPPX_ERR_ARG;
        }
        dims[d] = a + 1;
    }
    return OPPX_OK;
}

}  // extern "C"

namespace {

// Overlay builder over per-cell bounding boxes box[6c..6c+5] = {min xyz, max xyz}.
int build_overlay(int64_t n_cells, const double* c_det, const double* c_vol, const int* c_gbl_id,
                  std::vector<double>& box, const double bmin[3], const double bmax[3], double gs, int* overlay) {
    int64_t dims[3];
    int rc = oppx_dh_grid_dims(bmin, bmax, gs, dims);
    if (rc != OPPX_OK) return rc;
    const int64_t X = dims[0], XY = dims[0] * dims[1];
    const int64_t gsize = dims[0] * dims[1] * dims[2];
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t c = 0; c < n_cells; c++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], box[6 * c + d]);
            hi[d] = std::max(hi[d], box[6 * c + 3 + d]);
        }
    double ext = 0.0;
    for (int d = 0; d < 3; d++) ext = std::max(ext, hi[d] - lo[d]);
    const double margin = 1e-7 * ext;
    for (int d = 0; d < 3; d++) { lo[d] -= 2 * margin; hi[d] += 2 * margin; }
    int nb = (int)std::cbrt((double)n_cells / 4.0);
    nb = std::max(1, std::min(nb, 512));
    const int NB[3] = {nb, nb, nb};
    double inv[3];
    for (int d = 0; d < 3; d++) inv[d] = (double)NB[d] / (hi[d] - lo[d]);
    auto bin_of = [&](int d, double x) {
        long b = (long)std::floor((x - lo[d]) * inv[d]);
        return (int)std::max(0L, std::min(b, (long)NB[d] - 1));
    };
    const int64_t nbins = (int64_t)NB[0] * NB[1] * NB[2];
    std::vector<int64_t> bstart(nbins + 1, 0);
    auto for_bins = [&](int64_t c, auto&& fn) {
        const double* b = &box[6 * c];
        int b0[3], b1[3];
        for (int d = 0; d < 3; d++) { b0[d] = bin_of(d, b[d] - margin); b1[d] = bin_of(d, b[3 + d] + margin); }
        for (int k = b0[2]; k <= b1[2]; k++)
            for (int j = b0[1]; j <= b1[1]; j++)
                for (int i = b0[0]; i <= b1[0]; i++) fn(i + (int64_t)NB[0] * (j + (int64_t)NB[1] * k));
    };
    for (int64_t c = 0; c < n_cells; c++) for_bins(c, [&](int64_t bi) { bstart[bi + 1]++; });
    for (int64_t b = 0; b < nbins; b++) bstart[b + 1] += bstart[b];
    std::vector<int> blist(bstart[nbins]);
    {
        std::vector<int64_t> cursor(bstart.begin(), bstart.end() - 1);
        for (int64_t c = 0; c < n_cells; c++) for_bins(c, [&](int64_t bi) { blist[cursor[bi]++] = (int)c; });
    }
    // first tet (ascending id) containing the point, or INT_MAX
    auto find_cell = [&](double x, double y, double z) -> int {
        const double q[3] = {x, y, z};
        for (int d = 0; d < 3; d++)
            if (q[d] < lo[d] || q[d] > hi[d]) return INT_MAX;
        const int64_t bi = bin_of(0, x) + (int64_t)NB[0] * (bin_of(1, y) + (int64_t)NB[1] * bin_of(2, z));
        for (int64_t p = bstart[bi]; p < bstart[bi + 1]; p++) {
            const int c = blist[p];
            const double* b = &box[6 * (int64_t)c];
            if (x < b[0] - margin || x > b[3] + margin || y < b[1] - margin || y > b[4] + margin ||
                z < b[2] - margin || z > b[5] + margin)
                continue;
            if (tet_contains(c_det + 16 * (int64_t)c, c_vol[c], x, y, z)) return c;
        }
        return INT_MAX;
    };
    auto gid = [&](int c) { return c_gbl_id ? c_gbl_id[c] : c; };

    // ---- step 1: box centroids with nudges
    std::vector<unsigned char> unresolved(gsize, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t idx = 0; idx < gsize; idx++) {
        const int64_t dz = idx / XY, dy = (idx % XY) / X, dx = idx % X;
        const double org[3] = {std::fma((double)dx, gs, bmin[0]), std::fma((double)dy, gs, bmin[1]),
                               std::fma((double)dz, gs, bmin[2])};
        double cen[3];
        for (int d = 0; d < 3; d++) {
            if (org[d] + gs <= bmax[d]) cen[d] = org[d] + gs * 0.5;
            else cen[d] = (org[d] + bmax[d]) * 0.5;
        }
        int cid = find_cell(cen[0], cen[1], cen[2]);
        if (cid == INT_MAX) { cen[0] += gs / 100.0; cid = find_cell(cen[0], cen[1], cen[2]); }
        if (cid == INT_MAX) { cen[1] += gs / 100.0; cid = find_cell(cen[0], cen[1], cen[2]); }
        if (cid == INT_MAX) { cen[2] += gs / 100.0; cid = find_cell(cen[0], cen[1], cen[2]); }
        if (cid == INT_MAX) {
            overlay[idx] = INT_MAX;
            unresolved[idx] = 1;
        } else {
            overlay[idx] = gid(cid);
        }
    }
    // ---- step 3: corner sweep for boxes whose centroid hit no tet
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t idx = 0; idx < gsize; idx++) {
        if (!unresolved[idx]) continue;
        const int64_t dz = idx / XY, dy = (idx % XY) / X, dx = idx % X;
        const double x = std::fma((double)dx, gs, bmin[0]);
        const double y = std::fma((double)dy, gs, bmin[1]);
        const double z = std::fma((double)dz, gs, bmin[2]);
        auto clampv = [&](int d, double v) { return (v > bmax[d]) ? bmax[d] : v; };
        const double xs[2] = {clampv(0, x), clampv(0, x + gs)};
        const double ys[2] = {clampv(1, y), clampv(1, y + gs)};
        const double zs[2] = {clampv(2, z), clampv(2, z + gs)};
        int best = INT_MAX;
        for (int kz = 0; kz < 2; kz++)
            for (int kx = 0; kx < 2; kx++)
                for (int ky = 0; ky < 2; ky++) {
                    const int cid = find_cell(xs[kx], ys[ky], zs[kz]);
                    if (cid != INT_MAX && gid(cid) < best) best = gid(cid);
                }
        if (best != INT_MAX) overlay[idx] = best;
    }
    return OPPX_OK;
}

}  // namespace

extern "C" {

int oppx_dh_build_overlay_nodes(int64_t n_cells, const double* c_det, const double* c_vol, const int* c_gbl_id,
                                const int* c2n, const double* node_pos, const double bmin[3],
                                const double bmax[3], double gs, int* overlay) {
    if (n_cells <= 0 || !c_det || !c_vol || !c2n || !node_pos || !bmin || !bmax || !overlay)
        return OPPX_ERR_ARG;
    std::vector<double> box(6 * (size_t)n_cells);
    for (int64_t c = 0; c < n_cells; c++) {
        double* b = &box[6 * c];
        for (int d = 0; d < 3; d++) { b[d] = 1e300; b[3 + d] = -1e300; }
        for (int v = 0; v < 4; v++) {
            const double* p = node_pos + 3 * (int64_t)c2n[4 * c + v];
            for (int d = 0; d < 3; d++) {
                b[d] = std::min(b[d], p[d]);
                b[3 + d] = std::max(b[3 + d], p[d]);
            }
        }
    }
    return build_overlay(n_cells, c_det, c_vol, c_gbl_id, box, bmin, bmax, gs, overlay);
}

int oppx_dh_build_overlay_geom(int64_t n_cells, const double* c_det, const double* c_vol, const int* c_gbl_id,
                               const double bmin[3], const double bmax[3], double gs, int* overlay) {
    // Vertex j of a tet is where the three barycentric functions of the other vertices vanish:
    // -b_i x + g_i y - d_i z = -a_i for i != j (Cramer's rule); the boxes only need to be conservative.
    if (n_cells <= 0 || !c_det || !c_vol || !bmin || !bmax || !overlay) return OPPX_ERR_ARG;
    std::vector<double> box(6 * (size_t)n_cells);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n_cells; c++) {
        const double* t = c_det + 16 * c;
        double* b = &box[6 * c];
        for (int d = 0; d < 3; d++) { b[d] = 1e300; b[3 + d] = -1e300; }
        for (int j = 0; j < 4; j++) {
            double A[3][3], r[3];
            int row = 0;
            for (int i = 0; i < 4; i++) {
                if (i == j) continue;
                A[row][0] = -t[4 * i + 1];
                A[row][1] = t[4 * i + 2];
                A[row][2] = -t[4 * i + 3];
                r[row] = -t[4 * i + 0];
                row++;
            }
            const double D = det3(A);
            if (D == 0.0 || !std::isfinite(D)) {  // degenerate tet: contains no point, never a candidate
                for (int d = 0; d < 3; d++) { b[d] = 1e300; b[3 + d] = -1e300; }
                break;
            }
            double v[3];
            for (int k = 0; k < 3; k++) {
                double M[3][3];
                for (int q = 0; q < 3; q++)
                    for (int w = 0; w < 3; w++) M[q][w] = (w == k) ? r[q] : A[q][w];
                v[k] = det3(M) / D;
            }
            for (int d = 0; d < 3; d++) {
                b[d] = std::min(b[d], v[d]);
                b[3 + d] = std::max(b[3 + d], v[d]);
            }
        }
    }
    return build_overlay(n_cells, c_det, c_vol, c_gbl_id, box, bmin, bmax, gs, overlay);
}


}  // extern "C"

namespace oppx_host {

// Face neighbours: c2c[4c+v] = the tet sharing the face opposite local vertex v of tet c, -1 on the
// boundary (apps/app_fempic/minifempic_funcs.h:388-443; in a conforming mesh an interior face belongs
// to exactly two tets, so the answer does not depend on the search order). Faces are bucketed by
// their smallest node.
void face_neighbours(int64_t nn, int64_t nc, const int* c2n, int* c2c) {
    std::vector<int64_t> start(nn + 1, 0);
    for (int64_t c = 0; c < nc; c++)
        for (int v = 0; v < 4; v++) {
            int a = INT_MAX;
            for (int w = 0; w < 4; w++)
                if (w != v) a = std::min(a, c2n[4 * c + w]);
            start[a + 1]++;
        }
    for (int64_t n = 0; n < nn; n++) start[n + 1] += start[n];
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    std::vector<int64_t> face_of(4 * nc);  // entries (c*4+v), grouped by smallest node
    for (int64_t c = 0; c < nc; c++)
        for (int v = 0; v < 4; v++) {
            int a = INT_MAX;
            for (int w = 0; w < 4; w++)
                if (w != v) a = std::min(a, c2n[4 * c + w]);
            face_of[fill[a]++] = 4 * c + v;
        }
    auto face_key = [&](int64_t fv, int out[3]) {
        const int64_t c = fv / 4;
        const int v = (int)(fv % 4);
        int m = 0;
        for (int w = 0; w < 4; w++)
            if (w != v) out[m++] = c2n[4 * c + w];
        std::sort(out, out + 3);
    };
#pragma omp parallel for schedule(dynamic, 4096)
    for (int64_t n = 0; n < nn; n++) {
        for (int64_t p = start[n]; p < start[n + 1]; p++) {
            const int64_t f = face_of[p];
            int kf[3];
            face_key(f, kf);
            int nb = -1;
            for (int64_t q = start[n]; q < start[n + 1]; q++) {
                if (q == p) continue;
                int kg[3];
                face_key(face_of[q], kg);
                if (kg[0] == kf[0] && kg[1] == kf[1] && kg[2] == kf[2]) {
                    nb = (int)(face_of[q] / 4);
                    break;
                }
            }
            c2c[f] = nb;
        }
    }
}

}  // namespace oppx_host
