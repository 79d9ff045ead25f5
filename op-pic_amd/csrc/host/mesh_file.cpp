// FEM-PIC text mesh files: mesh.dat (nodes in mm + elements), inlet.dat (inlet nodes + triangles) and
// wall.dat (fixed-potential nodes), the input format of the reference app's loaders
// (apps/app_fempic/minifempic_funcs.h:230-478 LoadVolumeMesh, :481-588 LoadSurfaceMesh; packed into
// the app's dats by apps_codegen/app_fempic_cg/fempic_misc_mesh_loader.h:66-177), and a writer of the
// 12-tet box meshes the reference's generator produces (scripts/batch/fempic/box_mesh_gen.py:19-178:
// a centre node per cube, each of the cube's 6 faces split into 2 triangles, each triangle closed by
// the centre into a tet).
//
// Built with -ffp-contract=off: every operation is one IEEE operation unless it is an explicit
// std::fma; where the reference's GCC build contracts an expression into an FMA, the FMA is written
// out here (node volumes, inlet normals), so the arrays equal the reference loader's bit for bit.

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "oppx.h"
#include "tet_geometry.h"

namespace {

// Whole-file reader with strtod/strtol parsing (the reference reads with std::ifstream >>, which
// gives the same correctly rounded doubles).
struct TextIn {
    std::string buf;
    const char* p = nullptr;
    bool ok = false;
    explicit TextIn(const char* path) {
        FILE* f = path ? fopen(path, "rb") : nullptr;
        if (!f) return;
        char chunk[1 << 16];
        size_t n;
        while ((n = fread(chunk, 1, sizeof(chunk), f)) > 0) buf.append(chunk, n);
        fclose(f);
        p = buf.c_str();
        ok = true;
    }
    bool num(long& v) {
        char* e = nullptr;
        v = strtol(p, &e, 10);
        if (e == p) return false;
        p = e;
        return true;
    }
    bool num(double& v) {
        char* e = nullptr;
        v = strtod(p, &e);
        if (e == p) return false;
        p = e;
        return true;
    }
    void skip_line() {
        while (*p && *p != '\n') p++;
        if (*p) p++;
    }
};

int64_t box12_node(int x, int y, int z, int nx, int ny) {
    return (int64_t)x + (int64_t)y * (nx + 1) + (int64_t)z * (int64_t)(ny + 1) * (nx + 1);
}

// The 12 tets of cube (x, y, z), as 1-based node numbers (corner a, corner b, corner c, centre) in
// the generator's order.
void box12_cube_tets(int x, int y, int z, int nx, int ny, int nz, int64_t out[12][4]) {
    const int64_t p1 = box12_node(x, y, z, nx, ny) + 1, p2 = box12_node(x, y, z + 1, nx, ny) + 1;
    const int64_t p3 = box12_node(x, y + 1, z, nx, ny) + 1, p4 = box12_node(x, y + 1, z + 1, nx, ny) + 1;
    const int64_t p5 = box12_node(x + 1, y, z, nx, ny) + 1, p6 = box12_node(x + 1, y, z + 1, nx, ny) + 1;
    const int64_t p7 = box12_node(x + 1, y + 1, z, nx, ny) + 1, p8 = box12_node(x + 1, y + 1, z + 1, nx, ny) + 1;
    const int64_t ctr = (int64_t)(nx + 1) * (ny + 1) * (nz + 1) + (int64_t)z * nx * ny + (int64_t)y * nx + x + 1;
    const int64_t t[12][3] = {{p1, p3, p5}, {p3, p5, p7}, {p4, p2, p6}, {p4, p6, p8}, {p5, p6, p7}, {p6, p7, p8},
                              {p1, p3, p2}, {p3, p4, p2}, {p3, p4, p7}, {p4, p7, p8}, {p1, p5, p2}, {p2, p5, p6}};
    for (int k = 0; k < 12; k++) {
        out[k][0] = t[k][0];
        out[k][1] = t[k][1];
        out[k][2] = t[k][2];
        out[k][3] = ctr;
    }
}

}  // namespace

extern "C" {

int oppx_box12_write(int nx, int ny, int nz, double cube_length, const char* dir) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !(cube_length > 0) || !dir) return OPPX_ERR_ARG;
    const int64_t n_corner = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
    const int64_t n_nodes = n_corner + (int64_t)nx * ny * nz, n_tets = (int64_t)12 * nx * ny * nz;
    if (n_tets > INT_MAX - 1 || n_nodes > INT_MAX - 1) return OPPX_ERR_ARG;
    // node coordinates in mm: corners (i * L), centres ((i + 0.5) * L)
    std::vector<std::array<double, 3>> pos((size_t)n_nodes);
    for (int z = 0; z <= nz; z++)
        for (int y = 0; y <= ny; y++)
            for (int x = 0; x <= nx; x++)
                pos[(size_t)box12_node(x, y, z, nx, ny)] = {x * cube_length, y * cube_length, z * cube_length};
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                pos[(size_t)(n_corner + (int64_t)z * nx * ny + (int64_t)y * nx + x)] = {
                    (x + 0.5) * cube_length, (y + 0.5) * cube_length, (z + 0.5) * cube_length};
    const std::string d(dir);
    FILE* fm = fopen((d + "/mesh.dat").c_str(), "w");
    FILE* fi = fopen((d + "/inlet.dat").c_str(), "w");
    FILE* fw = fopen((d + "/wall.dat").c_str(), "w");
    if (!fm || !fi || !fw) {
        if (fm) fclose(fm);
        if (fi) fclose(fi);
        if (fw) fclose(fw);
        return OPPX_ERR_ARG;
    }
    auto node_line = [&](FILE* f, int64_t n) {
        fprintf(f, "%lld %.14e %.14e %.14e\n", (long long)(n + 1), pos[(size_t)n][0], pos[(size_t)n][1],
                pos[(size_t)n][2]);
    };
    fprintf(fm, "%lld %lld\n", (long long)n_nodes, (long long)n_tets);
    for (int64_t n = 0; n < n_nodes; n++) node_line(fm, n);
    int64_t idx = 0;
    std::vector<std::array<int64_t, 3>> inlet_faces;
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                int64_t t[12][4];
                box12_cube_tets(x, y, z, nx, ny, nz, t);
                for (int k = 0; k < 12; k++)
                    fprintf(fm, "%lld 304 %lld %lld %lld %lld\n", (long long)++idx, (long long)t[k][0],
                            (long long)t[k][1], (long long)t[k][2], (long long)t[k][3]);
                // a tet whose three corner nodes all lie on z = 0 has its first face on the inlet
                // (the first two tets of a bottom cube)
                if (z == 0) {
                    inlet_faces.push_back({t[0][0], t[0][1], t[0][2]});
                    inlet_faces.push_back({t[1][0], t[1][1], t[1][2]});
                }
            }
    // the generator lists the distinct face rows in lexicographic order, each prefixed by "4 203"
    std::sort(inlet_faces.begin(), inlet_faces.end());
    inlet_faces.erase(std::unique(inlet_faces.begin(), inlet_faces.end()), inlet_faces.end());
    const double X = nx * cube_length, Y = ny * cube_length, Z = nz * cube_length;
    std::vector<int64_t> inlet_nodes, wall_nodes;
    for (int64_t n = 0; n < n_nodes; n++) {
        const auto& p = pos[(size_t)n];
        if (!(p[2] == 0 || p[2] == Z) && (p[0] == 0 || p[1] == 0 || p[0] == X || p[1] == Y)) wall_nodes.push_back(n);
        if (p[2] == 0) inlet_nodes.push_back(n);
    }
    fprintf(fi, "%zu %zu\n", inlet_nodes.size(), inlet_faces.size());
    for (int64_t n : inlet_nodes) node_line(fi, n);
    for (const auto& f : inlet_faces)
        fprintf(fi, "4 203 %lld %lld %lld\n", (long long)f[0], (long long)f[1], (long long)f[2]);
    fprintf(fw, "%zu 0\n", wall_nodes.size());
    for (int64_t n : wall_nodes) node_line(fw, n);
    const bool ok = !ferror(fm) && !ferror(fi) && !ferror(fw);
    fclose(fm);
    fclose(fi);
    fclose(fw);
    return ok ? OPPX_OK : OPPX_ERR_ARG;
}

int oppx_mesh_file_sizes(const char* mesh_path, const char* inlet_path, int64_t* n_nodes, int64_t* n_cells,
                         int64_t* n_ifaces) {
    if (!n_nodes || !n_cells || !n_ifaces) return OPPX_ERR_ARG;
    TextIn m(mesh_path);
    long nn = 0, ne = 0;
    if (!m.ok || !m.num(nn) || !m.num(ne) || nn <= 0 || ne <= 0) return OPPX_ERR_ARG;
    m.skip_line();
    for (long n = 0; n < nn; n++) m.skip_line();
    long cells = 0;
    for (long e = 0; e < ne; e++) {
        long index = 0, type = 0;
        if (!m.num(index) || !m.num(type)) return OPPX_ERR_ARG;
        if (type == 304) cells++;
        m.skip_line();
    }
    long faces = 0;
    if (inlet_path) {
        TextIn f(inlet_path);
        long fn = 0, fe = 0;
        if (!f.ok || !f.num(fn) || !f.num(fe)) return OPPX_ERR_ARG;
        f.skip_line();
        for (long n = 0; n < fn; n++) f.skip_line();
        for (long e = 0; e < fe; e++) {
            long index = 0, type = 0;
            if (!f.num(index) || !f.num(type)) return OPPX_ERR_ARG;
            if (type == 203) faces++;
            f.skip_line();
        }
    }
    if (cells > INT_MAX - 1) return OPPX_ERR_ARG;
    *n_nodes = nn;
    *n_cells = cells;
    *n_ifaces = faces;
    return OPPX_OK;
}

int oppx_mesh_file_load(const char* mesh_path, const char* inlet_path, const char* wall_path, double* node_pos,
                        int32_t* node_type, double* node_vol, int32_t* c2n, int32_t* c2c, double* c_det,
                        double* c_vol, int32_t* if2c, int32_t* if2n, double* if_u, double* if_v, double* if_normal,
                        double* if_area) {
    int64_t nn = 0, nc = 0, nf = 0;
    int rc = oppx_mesh_file_sizes(mesh_path, inlet_path, &nn, &nc, &nf);
    if (rc) return rc;
    if (!node_pos || !node_type || !node_vol || !c2n || !c2c || !c_det || !c_vol) return OPPX_ERR_ARG;
    if (nf > 0 && (!if2c || !if2n || !if_u || !if_v || !if_normal || !if_area)) return OPPX_ERR_ARG;
    // ---- volume mesh (LoadVolumeMesh): nodes in mm -> m, elements of type 304
    TextIn m(mesh_path);
    long hn = 0, he = 0;
    m.num(hn);
    m.num(he);
    for (int64_t n = 0; n < nn; n++) {
        long index = 0;
        double x, y, z;
        if (!m.num(index) || !m.num(x) || !m.num(y) || !m.num(z) || index != n + 1) return OPPX_ERR_ARG;
        node_pos[3 * n + 0] = x / 1000.;
        node_pos[3 * n + 1] = y / 1000.;
        node_pos[3 * n + 2] = z / 1000.;
        node_type[n] = 0;  // NORMAL
        node_vol[n] = 0.0;
    }
    int64_t c = 0;
    for (long e = 0; e < he; e++) {
        long index = 0, type = 0;
        m.num(index);
        m.num(type);
        if (type != 304) {
            m.skip_line();
            continue;
        }
        long a[4];
        for (int v = 0; v < 4; v++)
            if (!m.num(a[v]) || a[v] < 1 || a[v] > nn) return OPPX_ERR_ARG;
        int con[4] = {(int)a[0] - 1, (int)a[1] - 1, (int)a[2] - 1, (int)a[3] - 1};
        oppx_host::tet_geometry<true>(node_pos, con, c_det + 16 * c, c_vol + c);
        for (int v = 0; v < 4; v++) c2n[4 * c + v] = con[v];
        c++;
    }
    oppx_host::face_neighbours(nn, nc, c2n, c2c);
    // node volumes: every tet scatters lc_v(centroid) * volume to its nodes, in tet order, with the
    // weights of the loader's XtoLtet ((1/6)(alpha - x beta + y gamma - z delta)/V, the reference
    // build's contractions); nodes of boundary faces become OPEN (1)
    for (int64_t t = 0; t < nc; t++) {
        double p[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; d++)
            for (int v = 0; v < 4; v++) p[d] = std::fma(0.25, node_pos[3 * (int64_t)c2n[4 * t + v] + d], p[d]);
        const double* det = c_det + 16 * t;
        for (int v = 0; v < 4; v++) {
            const double* q = det + 4 * v;
            const double s = std::fma(-p[2], q[3], std::fma(p[1], q[2], std::fma(-p[0], q[1], q[0])));
            const double lc = (1.0 / 6.0) * s / c_vol[t];
            double& nv = node_vol[c2n[4 * t + v]];
            nv = std::fma(lc, c_vol[t], nv);
        }
        for (int v = 0; v < 4; v++)
            if (c2c[4 * t + v] < 0)
                for (int w = 0; w < 4; w++)
                    if (w != v) node_type[c2n[4 * t + w]] = 1;  // OPEN
    }
    // ---- surface meshes (LoadSurfaceMesh): inlet nodes INLET (2) and faces, wall nodes FIXED (3)
    auto surface = [&](const char* path, int type, bool faces) -> int {
        TextIn f(path);
        long fn = 0, fe = 0;
        if (!f.ok || !f.num(fn) || !f.num(fe)) return OPPX_ERR_ARG;
        for (long n = 0; n < fn; n++) {
            long index = 0;
            double x, y, z;
            if (!f.num(index) || !f.num(x) || !f.num(y) || !f.num(z)) return OPPX_ERR_ARG;
            if (index >= 1 && index <= nn) node_type[index - 1] = type;
        }
        if (!faces) return OPPX_OK;
        // the tet a face belongs to: the first (in index order) holding all three of its nodes
        int64_t k = 0;
        for (long e = 0; e < fe; e++) {
            long index = 0, ft = 0;
            f.num(index);
            f.num(ft);
            if (ft != 203) {
                f.skip_line();
                continue;
            }
            long a[3];
            for (int v = 0; v < 3; v++)
                if (!f.num(a[v]) || a[v] < 1 || a[v] > nn) return OPPX_ERR_ARG;
            const int n1 = (int)a[0] - 1, n2 = (int)a[1] - 1, n3 = (int)a[2] - 1;
            int owner = -1;
            for (int64_t t = 0; t < nc && owner < 0; t++) {
                int match = 0;
                for (int v = 0; v < 4; v++) {
                    const int q = c2n[4 * t + v];
                    if (q == n1 || q == n2 || q == n3) match++;
                }
                if (match == 3) owner = (int)t;
            }
            if (owner < 0) return OPPX_ERR_ARG;  // the reference loader exits here
            if2c[k] = owner;
            if2n[3 * k + 0] = n1;
            if2n[3 * k + 1] = n2;
            if2n[3 * k + 2] = n3;
            double u[3], w[3];
            for (int d = 0; d < 3; d++) {
                u[d] = node_pos[3 * (int64_t)n2 + d] - node_pos[3 * (int64_t)n1 + d];
                w[d] = node_pos[3 * (int64_t)n3 + d] - node_pos[3 * (int64_t)n1 + d];
                if_u[3 * k + d] = u[d];
                if_v[3 * k + d] = w[d];
            }
            const double nrm[3] = {std::fma(u[2], w[1], -(u[1] * w[2])), std::fma(u[0], w[2], -(u[2] * w[0])),
                                   std::fma(u[1], w[0], -(u[0] * w[1]))};
            const double len = std::sqrt(std::fma(nrm[2], nrm[2], std::fma(nrm[1], nrm[1], nrm[0] * nrm[0])));
            for (int d = 0; d < 3; d++) if_normal[3 * k + d] = nrm[d] / len;
            if_area[k] = len / 2;
            k++;
        }
        return OPPX_OK;
    };
    if (inlet_path && (rc = surface(inlet_path, 2, true)) != OPPX_OK) return rc;
    if (wall_path && (rc = surface(wall_path, 3, false)) != OPPX_OK) return rc;
    return OPPX_OK;
}

}  // extern "C"
