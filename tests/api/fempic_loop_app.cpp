// FEM-PIC main loop minus the field solve, written only against the opp_lib API, as OP-PIC's
// generated main is (apps_codegen/app_fempic_cg/fempic_opp.cpp:52-231, with the PETSc solve replaced by
// a fixed node potential): boundary potential, injection through an inlet with
// opp_inc_part_count_with_distribution, push, move, deposit, node charge density, electric field and
// the loop's OPP_MAX globals. Built against op-pic_amd/include/opp_lib.h and linked with
// libopp_mi355x.so: the drop-in check of the injection and field-side loop symbols.
//
// usage: fempic_loop_app <params> <in_dir> <out_dir>   (same files as oracle/ref/fempic_loop_ref_main.cpp;
// with `STRING hdf_filename = <file>` in the params the mesh comes from that HDF5 file instead)

#include <cstdio>
#include <string>
#include <vector>

#include "opp_hdf5.h"
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
    const int n_nodes = opp_params->get<OPP_INT>("nnodes");
    const int n_cells = opp_params->get<OPP_INT>("ncells");
    const int n_faces = opp_params->get<OPP_INT>("nfaces");
    const int n_rand = opp_params->get<OPP_INT>("nrand");
    const int inject_count = opp_params->get<OPP_INT>("inject_count");
    const int steps = opp_params->get<OPP_INT>("num_steps");
    OPP_REAL dt = opp_params->get<OPP_REAL>("dt");
    OPP_REAL mass = opp_params->get<OPP_REAL>("mass");
    OPP_REAL charge = opp_params->get<OPP_REAL>("charge");
    OPP_REAL spwt = opp_params->get<OPP_REAL>("spwt");
    OPP_REAL ion_velocity = opp_params->get<OPP_REAL>("ion_velocity");
    OPP_REAL wall_potential = opp_params->get<OPP_REAL>("wall_potential");
    const OPP_REAL gs = opp_params->get<OPP_REAL>("grid_spacing");
    const opp_point bmin(opp_params->get<OPP_REAL>("bminx"), opp_params->get<OPP_REAL>("bminy"),
                         opp_params->get<OPP_REAL>("bminz"));
    const opp_point bmax(opp_params->get<OPP_REAL>("bmaxx"), opp_params->get<OPP_REAL>("bmaxy"),
                         opp_params->get<OPP_REAL>("bmaxz"));

    // the mesh and the particle sets: from an OP-PIC HDF5 file when hdf_filename is given, with the
    // declarations of the generated HDF5 main (apps_codegen/app_fempic_cg/fempic_hdf5_opp.cpp:79-118),
    // otherwise from the binary arrays in in_dir
    std::string h5 = opp_params->get<OPP_STRING>("hdf_filename", false);
    if (h5 == "NOT FOUND") h5.clear();  // the value Params::get gives for an absent STRING key
    opp_set node_set, cell_set, iface_set, particle_set, dummy_part_set;
    opp_map c2n_map, c2c_map, if2c_map, p2c_map;
    opp_dat c_det_d, c_vol_d, c_ef_d, c_sd_d, c_gbl_id, n_volume, n_potential, n_charge_den, n_type_d, n_bnd_pot;
    opp_dat if_v_norm, if_u_norm, if_norm_d, if_distrib, if_n_pos, p_pos, p_vel, p_lc, dp_rand;
    if (!h5.empty()) {
        const char* file = h5.c_str();
        node_set = opp_decl_set_hdf5(file, "mesh_nodes");
        cell_set = opp_decl_set_hdf5(file, "mesh_cells");
        iface_set = opp_decl_set_hdf5(file, "inlet_faces_cells");
        particle_set = opp_decl_particle_set_hdf5(file, "particles", cell_set);
        dummy_part_set = opp_decl_particle_set_hdf5(file, "dummy particles", cell_set);
        c2n_map = opp_decl_map_hdf5(cell_set, node_set, 4, file, "c_v_n_map");
        c2c_map = opp_decl_map_hdf5(cell_set, cell_set, 4, file, "c_v_c_map");
        if2c_map = opp_decl_map_hdf5(iface_set, cell_set, 1, file, "if_v_c_map");
        opp_decl_map_hdf5(iface_set, node_set, 3, file, "if_v_n_map");
        c_det_d = opp_decl_dat_hdf5(cell_set, 16, DT_REAL, file, "c_det");
        c_vol_d = opp_decl_dat_hdf5(cell_set, 1, DT_REAL, file, "c_volume");
        c_ef_d = opp_decl_dat_hdf5(cell_set, 3, DT_REAL, file, "c_ef");
        c_sd_d = opp_decl_dat_hdf5(cell_set, 12, DT_REAL, file, "c_shape_deri");
        c_gbl_id = opp_decl_dat_hdf5(cell_set, 1, DT_INT, file, "c_gbl_id");
        opp_decl_dat_hdf5(cell_set, 1, DT_INT, file, "c_colors");
        opp_decl_dat_hdf5(cell_set, 3, DT_REAL, file, "c_centroids");
        n_volume = opp_decl_dat_hdf5(node_set, 1, DT_REAL, file, "n_vol");
        n_potential = opp_decl_dat_hdf5(node_set, 1, DT_REAL, file, "n_potential");
        n_charge_den = opp_decl_dat_hdf5(node_set, 1, DT_REAL, file, "n_charge_den");
        opp_decl_dat_hdf5(node_set, 3, DT_REAL, file, "n_pos");
        n_type_d = opp_decl_dat_hdf5(node_set, 1, DT_INT, file, "n_type");
        n_bnd_pot = opp_decl_dat_hdf5(node_set, 1, DT_REAL, file, "n_bnd_pot");
        if_v_norm = opp_decl_dat_hdf5(iface_set, 3, DT_REAL, file, "iface_v_norm");
        if_u_norm = opp_decl_dat_hdf5(iface_set, 3, DT_REAL, file, "iface_u_norm");
        if_norm_d = opp_decl_dat_hdf5(iface_set, 3, DT_REAL, file, "iface_norm");
        opp_decl_dat_hdf5(iface_set, 1, DT_REAL, file, "iface_area");
        if_distrib = opp_decl_dat_hdf5(iface_set, 1, DT_INT, file, "iface_dist");
        if_n_pos = opp_decl_dat_hdf5(iface_set, 9, DT_REAL, file, "iface_n_pos");
        p_pos = opp_decl_dat_hdf5(particle_set, 3, DT_REAL, file, "part_position");
        p_vel = opp_decl_dat_hdf5(particle_set, 3, DT_REAL, file, "part_velocity");
        p_lc = opp_decl_dat_hdf5(particle_set, 4, DT_REAL, file, "part_lc");
        p2c_map = opp_decl_map(particle_set, cell_set, 1, nullptr, "part_mesh_rel");
        dp_rand = opp_decl_dat_hdf5(dummy_part_set, 2, DT_REAL, file, "dummy_part_rand");
        if (node_set->size != n_nodes || cell_set->size != n_cells || iface_set->size != n_faces)
            opp_abort("HDF5 set sizes differ from the parameters");
    } else {
        auto c2n = read_arr<int>(in + "/c2n.bin", 4 * (size_t)n_cells);
        auto c2c = read_arr<int>(in + "/c2c.bin", 4 * (size_t)n_cells);
        auto c_det = read_arr<double>(in + "/c_det.bin", 16 * (size_t)n_cells);
        auto c_vol = read_arr<double>(in + "/c_vol.bin", (size_t)n_cells);
        auto c_ef = read_arr<double>(in + "/c_ef.bin", 3 * (size_t)n_cells);
        auto c_sd = read_arr<double>(in + "/c_sd.bin", 12 * (size_t)n_cells);
        auto n_vol = read_arr<double>(in + "/n_vol.bin", (size_t)n_nodes);
        auto n_type = read_arr<int>(in + "/n_type.bin", (size_t)n_nodes);
        auto n_pot = read_arr<double>(in + "/n_pot.bin", (size_t)n_nodes);
        auto if2c = read_arr<int>(in + "/if2c.bin", (size_t)n_faces);
        auto if_u = read_arr<double>(in + "/if_u.bin", 3 * (size_t)n_faces);
        auto if_v = read_arr<double>(in + "/if_v.bin", 3 * (size_t)n_faces);
        auto if_norm = read_arr<double>(in + "/if_norm.bin", 3 * (size_t)n_faces);
        auto if_npos = read_arr<double>(in + "/if_npos.bin", 9 * (size_t)n_faces);
        auto if_dist = read_arr<int>(in + "/if_dist.bin", (size_t)n_faces);
        std::vector<int> c_id(n_cells);
        for (int i = 0; i < n_cells; i++) c_id[i] = i;
        std::vector<double> zeros(n_nodes, 0.0);

        node_set = opp_decl_set(n_nodes, "mesh_nodes");
        cell_set = opp_decl_set(n_cells, "mesh_cells");
        iface_set = opp_decl_set(n_faces, "inlet_faces_cells");
        particle_set = opp_decl_particle_set("particles", cell_set);
        dummy_part_set = opp_decl_particle_set("dummy particles", cell_set);
        c2n_map = opp_decl_map(cell_set, node_set, 4, c2n.data(), "c_v_n_map");
        c2c_map = opp_decl_map(cell_set, cell_set, 4, c2c.data(), "c_v_c_map");
        if2c_map = opp_decl_map(iface_set, cell_set, 1, if2c.data(), "if_v_c_map");
        c_det_d = opp_decl_dat(cell_set, 16, DT_REAL, c_det.data(), "c_det");
        c_vol_d = opp_decl_dat(cell_set, 1, DT_REAL, c_vol.data(), "c_volume");
        c_ef_d = opp_decl_dat(cell_set, 3, DT_REAL, c_ef.data(), "c_ef");
        c_sd_d = opp_decl_dat(cell_set, 12, DT_REAL, c_sd.data(), "c_shape_deri");
        c_gbl_id = opp_decl_dat(cell_set, 1, DT_INT, c_id.data(), "c_gbl_id");
        n_volume = opp_decl_dat(node_set, 1, DT_REAL, n_vol.data(), "n_vol");
        n_potential = opp_decl_dat(node_set, 1, DT_REAL, n_pot.data(), "n_potential");
        n_charge_den = opp_decl_dat(node_set, 1, DT_REAL, zeros.data(), "n_charge_den");
        n_type_d = opp_decl_dat(node_set, 1, DT_INT, n_type.data(), "n_type");
        n_bnd_pot = opp_decl_dat(node_set, 1, DT_REAL, zeros.data(), "n_bnd_pot");
        if_v_norm = opp_decl_dat(iface_set, 3, DT_REAL, if_v.data(), "iface_v_norm");
        if_u_norm = opp_decl_dat(iface_set, 3, DT_REAL, if_u.data(), "iface_u_norm");
        if_norm_d = opp_decl_dat(iface_set, 3, DT_REAL, if_norm.data(), "iface_norm");
        if_distrib = opp_decl_dat(iface_set, 1, DT_INT, if_dist.data(), "iface_dist");
        if_n_pos = opp_decl_dat(iface_set, 9, DT_REAL, if_npos.data(), "iface_n_pos");
        p_pos = opp_decl_dat(particle_set, 3, DT_REAL, nullptr, "p_position");
        p_vel = opp_decl_dat(particle_set, 3, DT_REAL, nullptr, "p_velocity");
        p_lc = opp_decl_dat(particle_set, 4, DT_REAL, nullptr, "p_lc");
        p2c_map = opp_decl_map(particle_set, cell_set, 1, nullptr, "p2c_map");
        dp_rand = opp_decl_dat(dummy_part_set, 2, DT_REAL, nullptr, "dummy_part_rand");
    }
    auto rnd = read_arr<double>(in + "/rand.bin", 2 * (size_t)n_rand);

    opp_decl_const<OPP_REAL>(1, &spwt, "CONST_spwt");
    opp_decl_const<OPP_REAL>(1, &ion_velocity, "CONST_ion_velocity");
    opp_decl_const<OPP_REAL>(1, &dt, "CONST_dt");
    opp_decl_const<OPP_REAL>(1, &mass, "CONST_mass");
    opp_decl_const<OPP_REAL>(1, &charge, "CONST_charge");
    opp_decl_const<OPP_REAL>(1, &wall_potential, "CONST_wall_potential");

    opp_par_loop_all__init_boundary_pot_kernel(node_set, opp_arg_dat(n_type_d, OPP_READ),
                                               opp_arg_dat(n_bnd_pot, OPP_WRITE));
    opp_download_dat(n_bnd_pot);
    write_arr<double>(out + "/bnd_pot.bin", (const double*)n_bnd_pot->data, (size_t)n_nodes);

    // the random pairs of the injection, as init_inject_distributions hands them over
    // (apps/app_fempic/fempic_misc.h:84,160-170)
    opp_increase_particle_count(dummy_part_set, n_rand);
    {
        double* r = opp_get_data<OPP_REAL>(dp_rand);
        for (size_t i = 0; i < 2 * (size_t)n_rand; i++) r[i] = rnd[i];
        dp_rand->dirty_hd = Dirty::Device;
    }

    opp_inc_part_count_with_distribution(particle_set, inject_count, if_distrib, false);
    opp_download_dat(p2c_map->p2c_dat);
    write_arr<int>(out + "/dist1_p2c.bin", (const int*)p2c_map->p2c_dat->data, (size_t)particle_set->size);

    opp::BoundingBox bbox(3, bmin, bmax);
    opp_init_direct_hop_cg(gs, c_gbl_id, bbox, c2c_map, p2c_map, opp_arg_dat(p_pos, OPP_READ),
                           opp_arg_dat(p_lc, OPP_WRITE), opp_arg_dat(c_vol_d, p2c_map, OPP_READ),
                           opp_arg_dat(c_det_d, p2c_map, OPP_READ));

    FILE* stats = fopen((out + "/steps.txt").c_str(), "w");
    if (!stats) opp_abort("cannot write steps.txt");
    for (OPP_main_loop_iter = 0; OPP_main_loop_iter < steps; OPP_main_loop_iter++) {
        if (OPP_main_loop_iter != 0) opp_inc_part_count_with_distribution(particle_set, inject_count, if_distrib, false);
        const int old_nparts = particle_set->size;
        opp_par_loop_injected__inject_ions_kernel(
            particle_set, opp_arg_dat(p_pos, OPP_WRITE), opp_arg_dat(p_vel, OPP_WRITE), opp_arg_dat(p2c_map, OPP_RW),
            opp_arg_dat(if2c_map, p2c_map, OPP_READ), opp_arg_dat(c_ef_d, 0, if2c_map, p2c_map, OPP_READ),
            opp_arg_dat(if_u_norm, p2c_map, OPP_READ), opp_arg_dat(if_v_norm, p2c_map, OPP_READ),
            opp_arg_dat(if_norm_d, p2c_map, OPP_READ), opp_arg_dat(if_n_pos, p2c_map, OPP_READ),
            opp_arg_dat(dp_rand, OPP_READ, false));
        if (OPP_main_loop_iter == 0) {
            opp_download_particle_set(particle_set);
            write_arr<double>(out + "/inj1_pos.bin", (const double*)p_pos->data, 3 * (size_t)old_nparts);
            write_arr<double>(out + "/inj1_vel.bin", (const double*)p_vel->data, 3 * (size_t)old_nparts);
            write_arr<int>(out + "/inj1_p2c.bin", (const int*)p2c_map->p2c_dat->data, (size_t)old_nparts);
        }
        opp_reset_dat(n_charge_den, opp_zero_double16);
        opp_par_loop_all__calculate_new_pos_vel_kernel(particle_set, opp_arg_dat(c_ef_d, p2c_map, OPP_READ),
                                                       opp_arg_dat(p_pos, OPP_WRITE), opp_arg_dat(p_vel, OPP_WRITE));
        opp_particle_move__move_kernel(particle_set, c2c_map, p2c_map, opp_arg_dat(p_pos, OPP_READ),
                                       opp_arg_dat(p_lc, OPP_WRITE), opp_arg_dat(c_vol_d, p2c_map, OPP_READ),
                                       opp_arg_dat(c_det_d, p2c_map, OPP_READ));
        opp_par_loop_all__deposit_charge_on_nodes_kernel(
            particle_set, opp_arg_dat(p_lc, OPP_READ), opp_arg_dat(n_charge_den, 0, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 1, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 2, c2n_map, p2c_map, OPP_INC),
            opp_arg_dat(n_charge_den, 3, c2n_map, p2c_map, OPP_INC));
        opp_par_loop_all__compute_node_charge_density_kernel(node_set, opp_arg_dat(n_charge_den, OPP_RW),
                                                             opp_arg_dat(n_volume, OPP_READ));
        opp_reset_dat(c_ef_d, opp_zero_double16);
        opp_par_loop_all__compute_electric_field_kernel(
            cell_set, opp_arg_dat(c_ef_d, OPP_INC), opp_arg_dat(c_sd_d, OPP_READ),
            opp_arg_dat(n_potential, 0, c2n_map, OPP_READ), opp_arg_dat(n_potential, 1, c2n_map, OPP_READ),
            opp_arg_dat(n_potential, 2, c2n_map, OPP_READ), opp_arg_dat(n_potential, 3, c2n_map, OPP_READ));
        OPP_REAL max_n_chg_den = 0.0, max_n_pot = 0.0, max_c_ef = 0.0;
        opp_par_loop_all__get_max_cef_kernel(cell_set, opp_arg_dat(c_ef_d, OPP_READ),
                                             opp_arg_gbl(&max_c_ef, 1, "double", OPP_MAX));
        opp_par_loop_all__get_final_max_values_kernel(
            node_set, opp_arg_dat(n_charge_den, OPP_READ), opp_arg_gbl(&max_n_chg_den, 1, "double", OPP_MAX),
            opp_arg_dat(n_potential, OPP_READ), opp_arg_gbl(&max_n_pot, 1, "double", OPP_MAX));

        fprintf(stats, "%d %d %d %d %a %a %a\n", OPP_main_loop_iter, old_nparts, particle_set->size,
                OPP_move_max_hops, max_c_ef, max_n_chg_den, max_n_pot);
        if (OPP_main_loop_iter == 0 || OPP_main_loop_iter == steps - 1) {
            const std::string t = OPP_main_loop_iter == 0 ? "/s1_" : "/final_";
            const int sz = particle_set->size;
            opp_download_particle_set(particle_set);
            opp_download_dat(n_charge_den);
            opp_download_dat(c_ef_d);
            write_arr<double>(out + t + "pos.bin", (const double*)p_pos->data, 3 * (size_t)sz);
            write_arr<double>(out + t + "vel.bin", (const double*)p_vel->data, 3 * (size_t)sz);
            write_arr<double>(out + t + "lc.bin", (const double*)p_lc->data, 4 * (size_t)sz);
            write_arr<int>(out + t + "p2c.bin", (const int*)p2c_map->p2c_dat->data, (size_t)sz);
            write_arr<double>(out + t + "charge.bin", (const double*)n_charge_den->data, (size_t)n_nodes);
            write_arr<double>(out + t + "ef.bin", (const double*)c_ef_d->data, 3 * (size_t)n_cells);
        }
    }
    fclose(stats);
    opp_exit();
    return 0;
}
