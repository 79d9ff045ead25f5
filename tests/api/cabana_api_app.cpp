// CabanaPIC particle move with per-hop current deposit, written only against the opp_lib API as
// OP-PIC's generated main is (apps_codegen/app_cabanapic_cg/cabana_opp.cpp: opp_reset_dat(c_acc) and
// opp_particle_move(move_deposit_kernel) per step, the field loops around them replaced by fixed
// interpolated fields). Built against op-pic_amd/include/opp_lib.h and linked with libopp_mi355x.so:
// the drop-in check of the move_deposit loop symbol.
//
// usage: cabana_api_app <params> <in_dir> <out_dir>   (same files as oracle/ref/cabana_ref_main.cpp)

#include <cstdio>
#include <string>
#include <vector>

#include "opp_lib.h"

template <typename T>
static std::vector<T> read_arr(const std::string& path, size_t n) {
    std::vector<T> v(n);
    FILE* f = fopen(path.c_str(), "rb");
    if (!f || (n && fread(v.data(), sizeof(T), n, f) != n)) opp_abort("cannot read " + path);
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
    const int n_cells = opp_params->get<OPP_INT>("ncells");
    const int n_part = opp_params->get<OPP_INT>("nparts");
    const int steps = opp_params->get<OPP_INT>("num_steps");
    OPP_REAL qsp = opp_params->get<OPP_REAL>("qsp");
    OPP_REAL qdt_2mc = opp_params->get<OPP_REAL>("qdt_half_mc");
    OPP_REAL cdt_d[3] = {opp_params->get<OPP_REAL>("cdt_dx"), opp_params->get<OPP_REAL>("cdt_dy"),
                         opp_params->get<OPP_REAL>("cdt_dz")};

    auto c2ngc = read_arr<int>(in + "/c2ngc.bin", 6 * (size_t)n_cells);
    auto interp = read_arr<double>(in + "/c_interp.bin", 18 * (size_t)n_cells);
    auto p_pos = read_arr<double>(in + "/p_pos.bin", 3 * (size_t)n_part);
    auto p_vel = read_arr<double>(in + "/p_vel.bin", 3 * (size_t)n_part);
    auto p_weight = read_arr<double>(in + "/p_weight.bin", (size_t)n_part);
    auto p_cell = read_arr<int>(in + "/p2c.bin", (size_t)n_part);
    std::vector<double> zeros(12 * (size_t)n_cells, 0.0);

    opp_set c_set = opp_decl_set(n_cells, "mesh_cells");
    opp_set p_set = opp_decl_particle_set("particles", c_set);
    opp_map c2ngc_map = opp_decl_map(c_set, c_set, 6, c2ngc.data(), "c2c_non_ghost_map");
    opp_dat c_acc = opp_decl_dat(c_set, 12, DT_REAL, zeros.data(), "c_acc");
    opp_dat c_interp = opp_decl_dat(c_set, 18, DT_REAL, interp.data(), "c_interp");
    opp_dat pos = opp_decl_dat(p_set, 3, DT_REAL, nullptr, "p_pos");
    opp_dat vel = opp_decl_dat(p_set, 3, DT_REAL, nullptr, "p_vel");
    opp_dat streak = opp_decl_dat(p_set, 3, DT_REAL, nullptr, "p_streak_mid");
    opp_dat weight = opp_decl_dat(p_set, 1, DT_REAL, nullptr, "p_weight");
    opp_map p2c_map = opp_decl_map(p_set, c_set, 1, nullptr, "p_mesh_rel");

    opp_decl_const<OPP_REAL>(1, &qsp, "CONST_qsp");
    opp_decl_const<OPP_REAL>(3, cdt_d, "CONST_cdt_d");
    opp_decl_const<OPP_REAL>(1, &qdt_2mc, "CONST_qdt_2mc");

    opp_increase_particle_count(p_set, n_part);
    {
        double* pp = opp_get_data<OPP_REAL>(pos);
        double* pv = opp_get_data<OPP_REAL>(vel);
        double* ps = opp_get_data<OPP_REAL>(streak);
        double* pw = opp_get_data<OPP_REAL>(weight);
        int* pc = opp_get_data(p2c_map);
        for (size_t i = 0; i < 3 * (size_t)n_part; i++) { pp[i] = p_pos[i]; pv[i] = p_vel[i]; ps[i] = 0.0; }
        for (int i = 0; i < n_part; i++) { pw[i] = p_weight[i]; pc[i] = p_cell[i]; }
    }
    opp_upload_particle_set(p_set);

    FILE* stats = fopen((out + "/steps.txt").c_str(), "w");
    if (!stats) opp_abort("cannot write steps.txt");
    for (OPP_main_loop_iter = 0; OPP_main_loop_iter < steps; OPP_main_loop_iter++) {
        opp_reset_dat(c_acc, opp_zero_double16);
        opp_particle_move__move_deposit_kernel(p_set, c2ngc_map, p2c_map, opp_arg_dat(vel, OPP_RW),
                                               opp_arg_dat(pos, OPP_RW), opp_arg_dat(streak, OPP_RW),
                                               opp_arg_dat(weight, OPP_READ), opp_arg_dat(c_interp, p2c_map, OPP_READ),
                                               opp_arg_dat(c_acc, p2c_map, OPP_INC));
        fprintf(stats, "%d %d %d\n", OPP_main_loop_iter, p_set->size, OPP_move_max_hops);
        if (OPP_main_loop_iter == 0 || OPP_main_loop_iter == steps - 1) {
            const std::string t = OPP_main_loop_iter == 0 ? "/s1_" : "/final_";
            const size_t sz = (size_t)p_set->size;
            opp_download_particle_set(p_set);
            opp_download_dat(c_acc);
            write_arr<double>(out + t + "pos.bin", (const double*)pos->data, 3 * sz);
            write_arr<double>(out + t + "vel.bin", (const double*)vel->data, 3 * sz);
            write_arr<double>(out + t + "streak.bin", (const double*)streak->data, 3 * sz);
            write_arr<int>(out + t + "p2c.bin", (const int*)p2c_map->p2c_dat->data, sz);
            write_arr<double>(out + t + "acc.bin", (const double*)c_acc->data, 12 * (size_t)n_cells);
        }
    }
    fclose(stats);
    opp_exit();
    return 0;
}
