// Driver that runs the reference OP-PIC NESO-advection (mdir) hot path -- the seq backend and
// the generated seq loops compiled from the reference sources by oracle/Makefile -- on arrays
// written by the oppx input builders. Test infrastructure only.
//
// The step mirrors apps_codegen/app_neso_advection_mdir_cg/advec_opp.cpp:105-129 without the
// verify loop: update_pos then the particle move. Hop counts come from wrapping
// opp_check_part_move_status, called once per elemental-kernel evaluation.
//
// usage: neso_ref <params> <in_dir> <out_dir>

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
    if (fread(v.data(), sizeof(T), n, f) != n) { fprintf(stderr, "short read %s\n", path.c_str()); exit(2); }
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
    std::vector<int> c_id(n_cells);
    for (int i = 0; i < n_cells; i++) c_id[i] = i;

    opp_set cell_set = opp_decl_set(n_cells, "mesh_cells");
    opp_map c2c_map = opp_decl_map(cell_set, cell_set, 4, c2c.data(), "c_c_map");
    opp_dat c_idx = opp_decl_dat(cell_set, 1, DT_INT, c_id.data(), "c_index");
    opp_dat c_pos_ll_d = opp_decl_dat(cell_set, 2, DT_REAL, c_pos_ll.data(), "c_pos_ll");
    opp_set part_set = opp_decl_particle_set("particles", cell_set);
    opp_map p2c_map = opp_decl_map(part_set, cell_set, 1, nullptr, "p_mesh_rel");
    opp_dat p_pos_d = opp_decl_dat(part_set, 2, DT_REAL, nullptr, "p_pos");
    opp_dat p_vel_d = opp_decl_dat(part_set, 2, DT_REAL, nullptr, "p_vel");
    opp_dat p_mdir = opp_decl_dat(part_set, 2, DT_INT, nullptr, "p_move_dir");
    (void)c_idx;

    opp_decl_const<OPP_REAL>(2, extents, "CONST_extents");
    opp_decl_const<OPP_REAL>(1, &dt, "CONST_dt");
    opp_decl_const<OPP_REAL>(1, &cell_width, "CONST_cell_width");
    opp_decl_const<OPP_INT>(2, ndimcells, "CONST_ndimcells");

    opp_increase_particle_count(part_set, n_part);
    {
        double* pp = opp_get_data<OPP_REAL>(p_pos_d);
        double* pv = opp_get_data<OPP_REAL>(p_vel_d);
        int* md = opp_get_data<OPP_INT>(p_mdir);
        int* pc = opp_get_data(p2c_map);
        for (size_t i = 0; i < 2 * (size_t)n_part; i++) { pp[i] = p_pos[i]; pv[i] = p_vel[i]; md[i] = 1; }
        for (int i = 0; i < n_part; i++) pc[i] = p_cell[i];
    }

    FILE* stats = fopen((out + "/steps.txt").c_str(), "w");
    double loop_s = 0.0;  // the step's loops only (update_pos + move), for CPU timing
    for (OPP_main_loop_iter = 0; OPP_main_loop_iter < steps; OPP_main_loop_iter++) {
        oppx_ref_hops.assign((size_t)part_set->set_capacity + 1, 0);
        const auto t_step = std::chrono::steady_clock::now();
        opp_par_loop_all__update_pos_kernel(part_set, opp_arg_dat(p_vel_d, OPP_READ), opp_arg_dat(p_pos_d, OPP_RW));
        opp_particle_move__move_kernel(part_set, c2c_map, p2c_map, opp_arg_dat(p_pos_d, OPP_READ),
                                       opp_arg_dat(p_mdir, OPP_RW), opp_arg_dat(c_pos_ll_d, p2c_map, OPP_READ));
        loop_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_step).count();
        long long tot = 0;
        int mx = 0;
        for (int i = 0; i < part_set->size; i++) { tot += oppx_ref_hops[i]; mx = std::max(mx, oppx_ref_hops[i]); }
        fprintf(stats, "%d %d %lld %d\n", OPP_main_loop_iter, part_set->size, tot, mx);
        if (OPP_main_loop_iter == 0) {
            write_arr<double>(out + "/s1_pos.bin", (const double*)p_pos_d->data, 2 * (size_t)part_set->size);
            write_arr<int>(out + "/s1_p2c.bin", (const int*)p2c_map->p2c_dat->data, (size_t)part_set->size);
            write_arr<int>(out + "/s1_mdir.bin", (const int*)p_mdir->data, 2 * (size_t)part_set->size);
            write_arr<int>(out + "/s1_hops.bin", oppx_ref_hops.data(), (size_t)part_set->size);
        }
    }
    fclose(stats);
    write_arr<double>(out + "/final_pos.bin", (const double*)p_pos_d->data, 2 * (size_t)part_set->size);
    write_arr<int>(out + "/final_p2c.bin", (const int*)p2c_map->p2c_dat->data, (size_t)part_set->size);
    write_arr<int>(out + "/final_mdir.bin", (const int*)p_mdir->data, 2 * (size_t)part_set->size);
    if (FILE* tf = fopen((out + "/loop_seconds.txt").c_str(), "w")) {
        fprintf(tf, "%.9f\n", loop_s);
        fclose(tf);
    }
    opp_exit();
    return 0;
}
