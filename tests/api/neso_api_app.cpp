// NESO-advection (mdir) hot-path application written only against the opp_lib API, as OP-PIC's
// generated main is (apps_codegen/app_neso_advection_mdir_cg/advec_opp.cpp:50-129, minus the verify
// loop). Linked with libopp_mi355x.so: the drop-in check of the API boundary for the 2-D mover.
//
// usage: neso_api_app <params> <in_dir> <out_dir>   (same files as the reference driver in oracle/ref)

#include <cstdio>
#include <string>
#include <vector>

#include "opp_lib.h"

template <typename T>
static std::vector<T> read_arr(const std::string& path, size_t n) {
    std::vector<T> v(n);
    FILE* f = fopen(path.c_str(), "rb");
    if (!f || fread(v.data(), sizeof(T), n, f) != n) opp_abort("cannot read " + path);
    fclose(f);
    return v;
}

template <typename T>
static void write_arr(const std::string& path, const T* p, size_t n) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f || (n && fwrite(p, sizeof(T), n, f) != n)) opp_abort("cannot write " + path);
    fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s params in_dir out_dir\n", argv[0]);
        return 2;
    }
    opp_init(argc, argv);
    const std::string in = argv[2], out = argv[3];
    const int nx = opp_params->get<OPP_INT>("nx"), ny = opp_params->get<OPP_INT>("ny");
    const int n_part = opp_params->get<OPP_INT>("nparts");
    const int steps = opp_params->get<OPP_INT>("num_steps");
    OPP_REAL dt = opp_params->get<OPP_REAL>("dt");
    OPP_REAL cell_width = opp_params->get<OPP_REAL>("cell_width");
    OPP_REAL extents[2] = {nx * cell_width, ny * cell_width};
    OPP_INT ndimcells[2] = {nx, ny};
    const int n_cells = nx * ny;

    auto c_pos_ll = read_arr<double>(in + "/c_pos_ll.bin", 2 * (size_t)n_cells);
    auto c2c = read_arr<int>(in + "/c2c.bin", 4 * (size_t)n_cells);
    auto p_pos = read_arr<double>(in + "/p_pos.bin", 2 * (size_t)n_part);
    auto p_vel = read_arr<double>(in + "/p_vel.bin", 2 * (size_t)n_part);
    auto p_cell = read_arr<int>(in + "/p2c.bin", (size_t)n_part);

    opp_set cell_set = opp_decl_set(n_cells, "mesh_cells");
    opp_map c2c_map = opp_decl_map(cell_set, cell_set, 4, c2c.data(), "c_c_map");
    opp_dat c_pos_ll_d = opp_decl_dat(cell_set, 2, DT_REAL, c_pos_ll.data(), "c_pos_ll");
    opp_set part_set = opp_decl_particle_set("particles", cell_set);
    opp_map p2c_map = opp_decl_map(part_set, cell_set, 1, nullptr, "p_mesh_rel");
    opp_dat p_pos_d = opp_decl_dat(part_set, 2, DT_REAL, nullptr, "p_pos");
    opp_dat p_vel_d = opp_decl_dat(part_set, 2, DT_REAL, nullptr, "p_vel");
    opp_dat p_mdir = opp_decl_dat(part_set, 2, DT_INT, nullptr, "p_move_dir");

    opp_decl_const<OPP_REAL>(2, extents, "CONST_extents");
    opp_decl_const<OPP_REAL>(1, &dt, "CONST_dt");
    opp_decl_const<OPP_REAL>(1, &cell_width, "CONST_cell_width");
    opp_decl_const<OPP_INT>(2, ndimcells, "CONST_ndimcells");

    opp_increase_particle_count(part_set, n_part);
    for (size_t i = 0; i < 2 * (size_t)n_part; i++) {
        opp_get_data<OPP_REAL>(p_pos_d)[i] = p_pos[i];
        opp_get_data<OPP_REAL>(p_vel_d)[i] = p_vel[i];
        opp_get_data<OPP_INT>(p_mdir)[i] = 1;
    }
    for (int i = 0; i < n_part; i++) opp_get_data(p2c_map)[i] = p_cell[i];
    opp_upload_particle_set(part_set);

    FILE* stats = fopen((out + "/steps.txt").c_str(), "w");
    for (OPP_main_loop_iter = 0; OPP_main_loop_iter < steps; OPP_main_loop_iter++) {
        opp_par_loop_all__update_pos_kernel(part_set, opp_arg_dat(p_vel_d, OPP_READ), opp_arg_dat(p_pos_d, OPP_RW));
        opp_particle_move__move_kernel(part_set, c2c_map, p2c_map, opp_arg_dat(p_pos_d, OPP_READ),
                                       opp_arg_dat(p_mdir, OPP_RW), opp_arg_dat(c_pos_ll_d, p2c_map, OPP_READ));
        fprintf(stats, "%d %d %d\n", OPP_main_loop_iter, part_set->size, OPP_move_max_hops);
    }
    fclose(stats);
    opp_download_particle_set(part_set);
    write_arr<double>(out + "/final_pos.bin", opp_get_data<OPP_REAL>(p_pos_d), 2 * (size_t)part_set->size);
    write_arr<int>(out + "/final_p2c.bin", opp_get_data(p2c_map), (size_t)part_set->size);
    write_arr<int>(out + "/final_mdir.bin", opp_get_data<OPP_INT>(p_mdir), 2 * (size_t)part_set->size);
    opp_exit();
    return 0;
}
