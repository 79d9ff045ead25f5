// Driver that runs the reference FEM-PIC mesh loaders (apps/app_fempic/minifempic_funcs.h:
// LoadVolumeMesh and LoadSurfaceMesh, compiled from the reference sources by oracle/Makefile) on
// mesh.dat / inlet.dat / wall.dat and dumps, as flat binary arrays, what the app's init_mesh packs
// (apps_codegen/app_fempic_cg/fempic_misc_mesh_loader.h:66-177): node positions, types and volumes,
// c2n, c2c, c_det, c_vol and the inlet faces. Test infrastructure only.
//
// usage: mesh_ref <mesh.dat> <inlet.dat> <wall.dat> <out_dir>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "minifempic_funcs.h"

namespace {

template <typename T>
void write_vec(const std::string& path, const std::vector<T>& v) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
    if (!v.empty() && fwrite(v.data(), sizeof(T), v.size(), f) != v.size()) { fprintf(stderr, "short write\n"); exit(2); }
    fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) { fprintf(stderr, "usage: %s mesh inlet wall out_dir\n", argv[0]); return 2; }
    Volume vol;
    if (!LoadVolumeMesh(argv[1], vol) || !LoadSurfaceMesh(argv[2], vol, INLET, false) ||
        !LoadSurfaceMesh(argv[3], vol, FIXED, false))
        return 3;
    const std::string out = argv[4];
    const size_t nn = vol.nodes.size(), nc = vol.elements.size(), nf = vol.inlet_faces.size();
    std::vector<double> node_pos(3 * nn), node_vol(nn), c_det(16 * nc), c_vol(nc);
    std::vector<int> node_type(nn), c2n(4 * nc), c2c(4 * nc), if2c(nf), if2n(3 * nf);
    std::vector<double> if_u(3 * nf), if_v(3 * nf), if_normal(3 * nf), if_area(nf);
    for (size_t n = 0; n < nn; n++) {
        for (int d = 0; d < 3; d++) node_pos[3 * n + d] = vol.nodes[n].pos[d];
        node_vol[n] = vol.nodes[n].volume;
        node_type[n] = (int)vol.nodes[n].type;
    }
    for (size_t c = 0; c < nc; c++) {
        const Tetra& t = vol.elements[c];
        for (int v = 0; v < 4; v++) {
            c2n[4 * c + v] = t.con[v];
            c2c[4 * c + v] = t.cell_con[v];
            c_det[16 * c + 4 * v + 0] = t.alpha[v];
            c_det[16 * c + 4 * v + 1] = t.beta[v];
            c_det[16 * c + 4 * v + 2] = t.gamma[v];
            c_det[16 * c + 4 * v + 3] = t.delta[v];
        }
        c_vol[c] = t.volume;
    }
    for (size_t f = 0; f < nf; f++) {
        const Face& face = vol.inlet_faces[f];
        if2c[f] = face.cell_con;
        if_area[f] = face.area;
        for (int d = 0; d < 3; d++) {
            if2n[3 * f + d] = face.con[d];
            if_u[3 * f + d] = face.u[d];
            if_v[3 * f + d] = face.v[d];
            if_normal[3 * f + d] = face.normal[d];
        }
    }
    write_vec(out + "/node_pos.bin", node_pos);
    write_vec(out + "/node_vol.bin", node_vol);
    write_vec(out + "/node_type.bin", node_type);
    write_vec(out + "/c2n.bin", c2n);
    write_vec(out + "/c2c.bin", c2c);
    write_vec(out + "/c_det.bin", c_det);
    write_vec(out + "/c_vol.bin", c_vol);
    write_vec(out + "/if2c.bin", if2c);
    write_vec(out + "/if2n.bin", if2n);
    write_vec(out + "/if_u.bin", if_u);
    write_vec(out + "/if_v.bin", if_v);
    write_vec(out + "/if_normal.bin", if_normal);
    write_vec(out + "/if_area.bin", if_area);
    printf("mesh_ref: %zu nodes %zu cells %zu inlet faces\n", nn, nc, nf);
    return 0;
}
