// Driver that runs the reference OP-PIC CabanaPIC particle move with per-hop current deposit
// (seq backend and the generated seq move_deposit loop, built from the reference sources by
// oracle/Makefile) on arrays written by the fixture script, and dumps what the oracle and the GPU
// path are checked against. Test infrastructure only.
//
// Per step, as apps/app_cabanapic/cabana.cpp:153-165 does between the field loops:
// opp_reset_dat(c_acc), then opp_particle_move(move_deposit_kernel) over the 6-face
// non-ghost neighbour map. The interpolated fields (c_interp) are an input, held fixed.
// Per-particle hop counts come from wrapping opp_check_part_move_status.
//
// usage: cabana_ref <params> <in_dir> <out_dir>

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "opp_lib.h"

static std::vector<int> oppx_ref_hops;

#if defined(OPPX_REF_OMP)
#include "opp_omp.h"
#define opp_check_part_move_status(f, i1, m, n, t) (oppx_ref_hops[(n)]++, opp_check_part_move_status(f, i1, m, n, t))
#include OPPX_REF_KERNELS_OMP
#else
#include "opp_seq.h"
#define opp_check_part_move_status(m, n, rc) (oppx_ref_hops[(n)]++, opp_check_part_move_status(m, n, rc))
#include OPPX_REF_KERNELS_SEQ
#endif

namespace {

template <typename T>
std::vector<T> read_arr(const std::string& path, size_t n) {
    std::vector<T> v(n);
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path.c_str()); exit(2); }
    if (n && fread(v.data(), sizeof(T), n, f) != n) { fprintf(stderr, "short read %s\n", path.c_str()); exit(2); }
    fclose(f);
    return v;
}

template <typename T>
void write_arr(const std::string& path, const T* p, size_t n) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path.c_str()); exit(2); }
    if (n && fwrite(p, sizeof(T), n, f) != n) { fprintf(stderr, "short write %s\n", path.c_str()); exit(2); }
    fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 4) { fprintf(stderr, "usage: %s params in_dir out_dir\n", argv[0]); return 2; }
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

    FILE* stats = fopen((out + "/steps.txt").c_str(), "w");
    double loop_s = 0.0;  // the step's loops only (accumulator reset + move-deposit), for CPU timing
    for (OPP_main_loop_iter = 0; OPP_main_loop_iter < steps; OPP_main_loop_iter++) {
        oppx_ref_hops.assign((size_t)p_set->set_capacity + 1, 0);
        const auto t_step = std::chrono::steady_clock::now();
        opp_reset_dat(c_acc, opp_zero_double16);
        opp_particle_move__move_deposit_kernel(p_set, c2ngc_map, p2c_map, opp_arg_dat(vel, OPP_RW),
                                               opp_arg_dat(pos, OPP_RW), opp_arg_dat(streak, OPP_RW),
                                               opp_arg_dat(weight, OPP_READ), opp_arg_dat(c_interp, p2c_map, OPP_READ),
                                               opp_arg_dat(c_acc, p2c_map, OPP_INC));
        loop_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_step).count();
        long long tot = 0;
        int mx = 0;
        for (int i = 0; i < n_part; i++) { tot += oppx_ref_hops[i]; mx = std::max(mx, oppx_ref_hops[i]); }
        fprintf(stats, "%d %d %lld %d\n", OPP_main_loop_iter, p_set->size, tot, mx);
        if (OPP_main_loop_iter == 0 || OPP_main_loop_iter == steps - 1) {
            const std::string t = OPP_main_loop_iter == 0 ? "/s1_" : "/final_";
            const size_t sz = (size_t)p_set->size;
            write_arr<double>(out + t + "pos.bin", (const double*)pos->data, 3 * sz);
            write_arr<double>(out + t + "vel.bin", (const double*)vel->data, 3 * sz);
            write_arr<double>(out + t + "streak.bin", (const double*)streak->data, 3 * sz);
            write_arr<int>(out + t + "p2c.bin", (const int*)p2c_map->p2c_dat->data, sz);
            write_arr<double>(out + t + "acc.bin", (const double*)c_acc->data, 12 * (size_t)n_cells);
            write_arr<int>(out + t + "hops.bin", oppx_ref_hops.data(), sz);
        }
    }
    fclose(stats);
    if (FILE* tf = fopen((out + "/loop_seconds.txt").c_str(), "w")) {
        fprintf(tf, "%.9f\n", loop_s);
        fclose(tf);
    }
    opp_exit();
    return 0;
}
