// Driver that runs the reference OP-PIC FEM-PIC hot path (seq or omp backend, built from the
// reference sources by oracle/Makefile) on arrays written by the oppx input builders, and
// dumps what the oracle and the GPU path are checked against. Test infrastructure only.
//
// The step mirrors the MainLoop of apps_codegen/app_fempic_cg/fempic_opp.cpp:152-231 minus
// injection and the field solve: opp_reset_dat(n_charge_den), calculate_new_pos_vel,
// particle move (direct hop when opp_global_move is set), deposit_charge_on_nodes.
// Per-particle hop counts come from wrapping opp_check_part_move_status, which the
// generated move loops call once per elemental-kernel evaluation.
//
// usage: fempic_ref <params> <in_dir> <out_dir>

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

    const int n_nodes = opp_params->get<OPP_INT>("nnodes");
    const int n_cells = opp_params->get<OPP_INT>("ncells");
    const int n_part = opp_params->get<OPP_INT>("nparts");
    const int steps = opp_params->get<OPP_INT>("num_steps");
    OPP_REAL dt = opp_params->get<OPP_REAL>("dt");
    OPP_REAL mass = opp_params->get<OPP_REAL>("mass");
    OPP_REAL charge = opp_params->get<OPP_REAL>("charge");
    const OPP_REAL gs = opp_params->get<OPP_REAL>("grid_spacing");
    const OPP_REAL bx0 = opp_params->get<OPP_REAL>("bminx"), by0 = opp_params->get<OPP_REAL>("bminy"),
                   bz0 = opp_params->get<OPP_REAL>("bminz");
    const OPP_REAL bx1 = opp_params->get<OPP_REAL>("bmaxx"), by1 = opp_params->get<OPP_REAL>("bmaxy"),
                   bz1 = opp_params->get<OPP_REAL>("bmaxz");

    auto node_pos = read_arr<double>(in + "/node_pos.bin", 3 * (size_t)n_nodes);
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
    std::vector<double> zeros_n(n_nodes, 0.0);

    opp_set node_set = opp_decl_set(n_nodes, "mesh_nodes");
    opp_set cell_set = opp_decl_set(n_cells, "mesh_cells");
    opp_set particle_set = opp_decl_particle_set("particles", cell_set);
    opp_map c2n_map = opp_decl_map(cell_set, node_set, 4, c2n.data(), "c_v_n_map");
    opp_map c2c_map = opp_decl_map(cell_set, cell_set, 4, c2c.data(), "c_v_c_map");
    opp_dat c_det_d = opp_decl_dat(cell_set, 16, DT_REAL, c_det.data(), "c_det");
    opp_dat c_vol_d = opp_decl_dat(cell_set, 1, DT_REAL, c_vol.data(), "c_volume");
    opp_dat c_ef_d = opp_decl_dat(cell_set, 3, DT_REAL, c_ef.data(), "c_ef");
    opp_dat c_gbl_id = opp_decl_dat(cell_set, 1, DT_INT, c_id.data(), "c_gbl_id");
    opp_dat n_charge_den = opp_decl_dat(node_set, 1, DT_REAL, zeros_n.data(), "n_charge_den");
    opp_dat n_pos = opp_decl_dat(node_set, 3, DT_REAL, node_pos.data(), "n_pos");
    opp_dat p_pos_d = opp_decl_dat(particle_set, 3, DT_REAL, nullptr, "p_position");
    opp_dat p_vel_d = opp_decl_dat(particle_set, 3, DT_REAL, nullptr, "p_velocity");
    opp_dat p_lc_d = opp_decl_dat(particle_set, 4, DT_REAL, nullptr, "p_lc");
    opp_map p2c_map = opp_decl_map(particle_set, cell_set, 1, nullptr, "p2c_map");
    (void)n_pos;

    opp_decl_const<OPP_REAL>(1, &dt, "CONST_dt");
    opp_decl_const<OPP_REAL>(1, &mass, "CONST_mass");
    opp_decl_const<OPP_REAL>(1, &charge, "CONST_charge");

    opp_increase_particle_count(particle_set, n_part);
    {
        double* pp = opp_get_data<OPP_REAL>(p_pos_d);
        double* pv = opp_get_data<OPP_REAL>(p_vel_d);
        int* pc = opp_get_data(p2c_map);
        for (size_t i = 0; i < 3 * (size_t)n_part; i++) { pp[i] = p_pos[i]; pv[i] = p_vel[i]; }
        for (int i = 0; i < n_part; i++) pc[i] = p_cell[i];
    }

    {
        opp::BoundingBox bbox(3, opp_point(bx0, by0, bz0), opp_point(bx1, by1, bz1));
        opp_init_direct_hop_cg(gs, c_gbl_id, bbox, c2c_map, p2c_map,
                               opp_arg_dat(p_pos_d, OPP_READ), opp_arg_dat(p_lc_d, OPP_WRITE),
                               opp_arg_dat(c_vol_d, p2c_map, OPP_READ), opp_arg_dat(c_det_d, p2c_map, OPP_READ));
        if (useGlobalMove) {
            write_arr<int>(out + "/overlay.bin", cellMapper->structMeshToCellMapping, cellMapper->globalGridSize);
            const long long dims[3] = {(long long)cellMapper->globalGridDimsX, (long long)cellMapper->globalGridDimsY,
                                       (long long)cellMapper->globalGridDimsZ};
            write_arr<long long>(out + "/overlay_dims.bin", dims, 3);
        }
    }

    FILE* stats = fopen((out + "/steps.txt").c_str(), "w");
    std::vector<int> hops_first, hops_last;
    double loop_s = 0.0;  // wall time of the steps' loops alone (no set-up, no dumps)
    for (OPP_main_loop_iter = 0; OPP_main_loop_iter < steps; OPP_main_loop_iter++) {
        oppx_ref_hops.assign((size_t)particle_set->set_capacity + 1, 0);
        const int before = particle_set->size;
        const auto t_step = std::chrono::steady_clock::now();

        opp_reset_dat(n_charge_den, opp_zero_double16);
        opp_par_loop_all__calculate_new_pos_vel_kernel(particle_set, opp_arg_dat(c_ef_d, p2c_map, OPP_READ),
                                                       opp_arg_dat(p_pos_d, OPP_WRITE),
                                                       opp_arg_dat(p_vel_d, OPP_WRITE));
        opp_particle_move__move_kernel(particle_set, c2c_map, p2c_map, opp_arg_dat(p_pos_d, OPP_READ),
                                       opp_arg_dat(p_lc_d, OPP_WRITE), opp_arg_dat(c_vol_d, p2c_map, OPP_READ),
                                       opp_arg_dat(c_det_d, p2c_map, OPP_READ));
        opp_par_loop_all__deposit_charge_on_nodes_kernel(
            particle_set, opp_arg_dat(p_lc_d, OPP_READ), opp_arg_dat(n_charge_den, 0, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 1, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 2, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 3, c2n_map, p2c_map, OPP_INC));
        loop_s += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_step).count();

        long long tot = 0;
        int mx = 0;
        for (int i = 0; i < before; i++) { tot += oppx_ref_hops[i]; mx = std::max(mx, oppx_ref_hops[i]); }
        double qsum = 0.0;
        const double* q = (const double*)n_charge_den->data;
        for (int i = 0; i < n_nodes; i++) qsum += q[i];
        fprintf(stats, "%d %d %d %lld %d %.17g\n", OPP_main_loop_iter, before, particle_set->size, tot, mx, qsum);
        if (OPP_main_loop_iter == 0) {
            hops_first.assign(oppx_ref_hops.begin(), oppx_ref_hops.begin() + before);
            const int sz = particle_set->size;
            write_arr<double>(out + "/s1_pos.bin", (const double*)p_pos_d->data, 3 * (size_t)sz);
            write_arr<double>(out + "/s1_vel.bin", (const double*)p_vel_d->data, 3 * (size_t)sz);
            write_arr<double>(out + "/s1_lc.bin", (const double*)p_lc_d->data, 4 * (size_t)sz);
            write_arr<int>(out + "/s1_p2c.bin", (const int*)p2c_map->p2c_dat->data, (size_t)sz);
            write_arr<double>(out + "/s1_charge.bin", q, (size_t)n_nodes);
            write_arr<int>(out + "/s1_hops.bin", hops_first.data(), hops_first.size());
        }
        if (OPP_main_loop_iter == steps - 1) hops_last.assign(oppx_ref_hops.begin(), oppx_ref_hops.begin() + before);
    }
    fclose(stats);
    if (FILE* tf = fopen((out + "/loop_seconds.txt").c_str(), "w")) {
        fprintf(tf, "%.6f\n", loop_s);
        fclose(tf);
    }

    const int sz = particle_set->size;
    write_arr<double>(out + "/final_pos.bin", (const double*)p_pos_d->data, 3 * (size_t)sz);
    write_arr<double>(out + "/final_vel.bin", (const double*)p_vel_d->data, 3 * (size_t)sz);
    write_arr<double>(out + "/final_lc.bin", (const double*)p_lc_d->data, 4 * (size_t)sz);
    write_arr<int>(out + "/final_p2c.bin", (const int*)p2c_map->p2c_dat->data, (size_t)sz);
    write_arr<double>(out + "/final_charge.bin", (const double*)n_charge_den->data, (size_t)n_nodes);
    write_arr<int>(out + "/final_hops.bin", hops_last.data(), hops_last.size());

    opp_exit();
    return 0;
}
