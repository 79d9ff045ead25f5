// FEM-PIC hot-path application written only against the opp_lib API, as OP-PIC's generated main is
// (apps_codegen/app_fempic_cg/fempic_opp.cpp:52-231, minus injection and the field solve). Built against
// op-pic_amd/include/opp_lib.h and linked with libopp_mi355x.so: the drop-in check of the API boundary.
//
// usage: fempic_api_app <params> <in_dir> <out_dir>   (same files as the reference driver in oracle/ref)

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
    const int n_nodes = opp_params->get<OPP_INT>("nnodes");
    const int n_cells = opp_params->get<OPP_INT>("ncells");
    const int n_part = opp_params->get<OPP_INT>("nparts");
    const int steps = opp_params->get<OPP_INT>("num_steps");
    OPP_REAL dt = opp_params->get<OPP_REAL>("dt");
    OPP_REAL mass = opp_params->get<OPP_REAL>("mass");
    OPP_REAL charge = opp_params->get<OPP_REAL>("charge");
    const OPP_REAL gs = opp_params->get<OPP_REAL>("grid_spacing");
    const opp_point bmin(opp_params->get<OPP_REAL>("bminx"), opp_params->get<OPP_REAL>("bminy"),
                         opp_params->get<OPP_REAL>("bminz"));
    const opp_point bmax(opp_params->get<OPP_REAL>("bmaxx"), opp_params->get<OPP_REAL>("bmaxy"),
                         opp_params->get<OPP_REAL>("bmaxz"));

    auto c2n = read_arr<int>(in + "/c2n.bin", 4 * (size_t)n_cells);
    auto c2c = read_arr<int>(in + "/c2c.bin", 4 * (size_t)n_cells);
    auto c_det = read_arr<double>(in + "/c_det.bin", 16 * (size_t)n_cells);
    auto c_vol = read_arr<double>(in + "/c_vol.bin", (size_t)n_cells);
    auto c_ef = read_arr<double>(in + "/c_ef.bin", 3 * (size_t)n_cells);
    auto p_pos = read_arr<double>(in + "/p_pos.bin", 3 * (size_t)n_part);
    auto p_vel = read_arr<double>(in + "/p_vel.bin", 3 * (size_t)n_part);
    auto p_cell = read_arr<int>(in + "/p2c.bin", (size_t)n_part);
    std::vector<int> c_id(n_cells);
    for (int i = 0; i < n_cells; i++) c_id[i] = i;
    std::vector<double> zeros(n_nodes, 0.0);

    opp_set node_set = opp_decl_set(n_nodes, "mesh_nodes");
    opp_set cell_set = opp_decl_set(n_cells, "mesh_cells");
    opp_set particle_set = opp_decl_particle_set("particles", cell_set);
    opp_map c2n_map = opp_decl_map(cell_set, node_set, 4, c2n.data(), "c_v_n_map");
    opp_map c2c_map = opp_decl_map(cell_set, cell_set, 4, c2c.data(), "c_v_c_map");
    opp_dat c_det_d = opp_decl_dat(cell_set, 16, DT_REAL, c_det.data(), "c_det");
    opp_dat c_volume = opp_decl_dat(cell_set, 1, DT_REAL, c_vol.data(), "c_volume");
    opp_dat c_ef_d = opp_decl_dat(cell_set, 3, DT_REAL, c_ef.data(), "c_ef");
    opp_dat c_gbl_id = opp_decl_dat(cell_set, 1, DT_INT, c_id.data(), "c_gbl_id");
    opp_dat n_charge_den = opp_decl_dat(node_set, 1, DT_REAL, zeros.data(), "n_charge_den");
    opp_dat p_pos_d = opp_decl_dat(particle_set, 3, DT_REAL, nullptr, "p_position");
    opp_dat p_vel_d = opp_decl_dat(particle_set, 3, DT_REAL, nullptr, "p_velocity");
    opp_dat p_lc = opp_decl_dat(particle_set, 4, DT_REAL, nullptr, "p_lc");
    opp_map p2c_map = opp_decl_map(particle_set, cell_set, 1, nullptr, "p2c_map");

    opp_decl_const<OPP_REAL>(1, &dt, "CONST_dt");
    opp_decl_const<OPP_REAL>(1, &mass, "CONST_mass");
    opp_decl_const<OPP_REAL>(1, &charge, "CONST_charge");

    opp_increase_particle_count(particle_set, n_part);
    for (size_t i = 0; i < 3 * (size_t)n_part; i++) {
        opp_get_data<OPP_REAL>(p_pos_d)[i] = p_pos[i];
        opp_get_data<OPP_REAL>(p_vel_d)[i] = p_vel[i];
    }
    for (int i = 0; i < n_part; i++) opp_get_data(p2c_map)[i] = p_cell[i];
    opp_upload_particle_set(particle_set);

    opp::BoundingBox bounding_box(3, bmin, bmax);
    opp_init_direct_hop_cg(gs, c_gbl_id, bounding_box, c2c_map, p2c_map, opp_arg_dat(p_pos_d, OPP_READ),
                           opp_arg_dat(p_lc, OPP_WRITE), opp_arg_dat(c_volume, p2c_map, OPP_READ),
                           opp_arg_dat(c_det_d, p2c_map, OPP_READ));
    if (useGlobalMove) write_arr<int>(out + "/overlay.bin", cellMapper->structMeshToCellMapping, cellMapper->globalGridSize);

    FILE* stats = fopen((out + "/steps.txt").c_str(), "w");
    for (OPP_main_loop_iter = 0; OPP_main_loop_iter < steps; OPP_main_loop_iter++) {
        const int before = particle_set->size;
        opp_reset_dat(n_charge_den, opp_zero_double16);
        opp_par_loop_all__calculate_new_pos_vel_kernel(particle_set, opp_arg_dat(c_ef_d, p2c_map, OPP_READ),
                                                       opp_arg_dat(p_pos_d, OPP_WRITE), opp_arg_dat(p_vel_d, OPP_WRITE));
        opp_particle_move__move_kernel(particle_set, c2c_map, p2c_map, opp_arg_dat(p_pos_d, OPP_READ),
                                       opp_arg_dat(p_lc, OPP_WRITE), opp_arg_dat(c_volume, p2c_map, OPP_READ),
                                       opp_arg_dat(c_det_d, p2c_map, OPP_READ));
        opp_par_loop_all__deposit_charge_on_nodes_kernel(
            particle_set, opp_arg_dat(p_lc, OPP_READ), opp_arg_dat(n_charge_den, 0, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 1, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 2, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 3, c2n_map, p2c_map, OPP_INC));
        fprintf(stats, "%d %d %d %d\n", OPP_main_loop_iter, before, particle_set->size, OPP_move_max_hops);
        if (OPP_main_loop_iter == 0) {
            opp_download_particle_set(particle_set);
            opp_download_dat(n_charge_den);
            const int sz = particle_set->size;
            write_arr<double>(out + "/s1_pos.bin", opp_get_data<OPP_REAL>(p_pos_d), 3 * (size_t)sz);
            write_arr<double>(out + "/s1_lc.bin", opp_get_data<OPP_REAL>(p_lc), 4 * (size_t)sz);
            write_arr<int>(out + "/s1_p2c.bin", opp_get_data(p2c_map), (size_t)sz);
            write_arr<double>(out + "/s1_charge.bin", opp_get_data<OPP_REAL>(n_charge_den), (size_t)n_nodes);
        }
    }
    fclose(stats);
    opp_download_particle_set(particle_set);
    opp_download_dat(n_charge_den);
    const int sz = particle_set->size;
    write_arr<double>(out + "/final_pos.bin", opp_get_data<OPP_REAL>(p_pos_d), 3 * (size_t)sz);
    write_arr<double>(out + "/final_vel.bin", opp_get_data<OPP_REAL>(p_vel_d), 3 * (size_t)sz);
    write_arr<double>(out + "/final_lc.bin", opp_get_data<OPP_REAL>(p_lc), 4 * (size_t)sz);
    write_arr<int>(out + "/final_p2c.bin", opp_get_data(p2c_map), (size_t)sz);
    write_arr<double>(out + "/final_charge.bin", opp_get_data<OPP_REAL>(n_charge_den), (size_t)n_nodes);
    opp_exit();
    return 0;
}
