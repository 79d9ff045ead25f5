// Generated-loop entry points of the FEM-PIC and NESO-advection hot paths, served by gfx950
// kernels through the oppx C-ABI. Applications call them exactly as the translator-generated
// code does (apps_codegen/app_fempic_cg/fempic_opp.cpp:41-50, apps_codegen/app_neso_advection_mdir_cg/advec_opp.cpp:40-43).
//
// Each loop: uploads host-dirty argument dats, launches on the default stream, marks written dats
// device-newer, and (for moves) runs the fill policy and reads back the new particle count -- the
// one device->host transfer of opp_finalize_particle_move (opp_lib/src/hip/opp_particle_comm.cpp:115-226).

#include <cfloat>
#include <map>
#include <utility>

#include "opp_api_internal.h"

namespace oppx_api {

// derived per-cell records, rebuilt when the source dat/map pointers change
struct TetAuxCache {
    const void* vol_d = nullptr;
    const void* c2c_d = nullptr;
    int64_t n = 0;
    oppx_tet_aux* aux = nullptr;
};
struct QuadCache {
    const void* ll_d = nullptr;
    const void* c2c_d = nullptr;
    int64_t n = 0;
    oppx_quad_cell* cells = nullptr;
    int gnx = 0, gny = 0;  // > 0: the mesh is, bit for bit, this periodic grid
};
static TetAuxCache g_tet;
static QuadCache g_quad;
static int32_t* g_verify_count = nullptr;  // device int of the NESO verify loop

void free_field_loop_buffers();

void free_loop_caches() {
    free_field_loop_buffers();
    dev_free(g_tet.aux);
    dev_free(g_quad.cells);
    dev_free(g_verify_count);
    g_verify_count = nullptr;
    g_tet = TetAuxCache();
    g_quad = QuadCache();
}

static const oppx_tet_aux* tet_aux(opp_dat c_volume, opp_map c2c) {
    const int64_t n = c2c->from->size;
    if (g_tet.vol_d != c_volume->data_d || g_tet.c2c_d != c2c->map_d || g_tet.n != n || !g_tet.aux) {
        dev_free(g_tet.aux);
        g_tet.aux = (oppx_tet_aux*)dev_alloc((size_t)n * sizeof(oppx_tet_aux));
        check(oppx_tet_aux_build((const double*)c_volume->data_d, c2c->map_d, n, g_tet.aux, nullptr), "tet aux");
        g_tet.vol_d = c_volume->data_d;
        g_tet.c2c_d = c2c->map_d;
        g_tet.n = n;
    }
    return g_tet.aux;
}

static double const_real(const char* name) {
    int bytes = 0;
    const char* p = opp_const_data(name, &bytes);
    if (!p || bytes < (int)sizeof(double)) opp_abort(std::string("constant not declared: ") + name);
    double v;
    memcpy(&v, p, sizeof(v));
    return v;
}

static void particle_dat_descs(opp_set set, std::vector<oppx_dat_desc>& in, std::vector<oppx_dat_desc>& swap) {
    in.clear();
    swap.clear();
    for (opp_dat d : *set->particle_dats) {
        in.push_back(oppx_dat_desc{d->data_d, d->dim, elem_bytes(d)});
        swap.push_back(oppx_dat_desc{d->data_swap_d, d->dim, elem_bytes(d)});
    }
    if (in.size() > 16) opp_abort("particle sets with more than 16 dats are not supported");
}

static void all_particle_dats_device_newer(opp_set set) {
    for (opp_dat d : *set->particle_dats) d->dirty_hd = Dirty::Host;
}

void sort_particles_device(opp_set set, bool drop_removed) {
    ensure_capacity_scratch(set);
    ParticleDev& pd = pdev(set);
    sync_size_to_device(set);
    for (opp_dat d : *set->particle_dats)
        if (d->dirty_hd == Dirty::Device) opp_upload_dat(d);
    std::vector<oppx_dat_desc> in, out;
    particle_dat_descs(set, in, out);
    check(oppx_sort_by_cell(set->size, set->set_capacity, pd.counters, set->cells_set->size,
                            (const int32_t*)set->mesh_relation_dat->data_d, in.data(), out.data(), (int)in.size(),
                            nullptr, drop_removed ? 1 : 0, pd.scratch, pd.scratch_bytes, nullptr),
          "particle sort");
    for (opp_dat d : *set->particle_dats) std::swap(d->data_d, d->data_swap_d);
    all_particle_dats_device_newer(set);
    oppx_move_counters c;
    read_counters(set, &c);
    set->size = c.size;
}

// opp_finalize_particle_move with the fill policy of opp_lib/src/core/opp_particle_organize_core.cpp:49-95
void finalize_move_device(opp_set set) {
    ParticleDev& pd = pdev(set);
    oppx_move_counters c;
    read_counters(set, &c);
    set->particle_remove_count = c.removed;
    OPP_move_max_hops = c.max_hops;
    if (c.faults) opp_abort("particle move: " + std::to_string(c.faults) + " walks hit an invalid cell or the hop cap");
    if (c.removed <= 0) return;
    const bool periodic = (OPP_fill_type == OPP_Sort_Periodic || OPP_fill_type == OPP_Shuffle_Periodic);
    const bool holefill = OPP_fill_type == OPP_HoleFill_All ||
                          (periodic && (OPP_main_loop_iter % OPP_fill_period != 0 || OPP_comm_iteration != 0));
    if (holefill) {
        std::vector<oppx_dat_desc> in, unused;
        particle_dat_descs(set, in, unused);
        check(oppx_holefill(set->size, set->set_capacity, pd.counters, pd.hole_bits, in.data(), (int)in.size(),
                            pd.scratch, pd.scratch_bytes, nullptr),
              "hole fill");
        all_particle_dats_device_newer(set);
        read_counters(set, &c);
        set->size = c.size;
    } else {
        // Sort_* (and Shuffle_*, whose seq reference is a sort as well): order by cell, removed last
        sort_particles_device(set, true);
    }
}

}  // namespace oppx_api

using namespace oppx_api;

// =============================================================================================
// FEM-PIC
// =============================================================================================

// apps_codegen/app_fempic_cg/hip/calculate_new_pos_vel_kernel_loop.hpp; elemental apps/app_fempic/kernels.h:93-103
void opp_par_loop_all__calculate_new_pos_vel_kernel(opp_set set, opp_arg arg0, opp_arg arg1, opp_arg arg2) {
    opp_profiler->start("calculate_new_pos_vel_kernel");
    opp_arg args[3] = {arg0, arg1, arg2};
    args_to_device(3, args);
    if (set->mesh_relation_dat->dirty_hd == Dirty::Device) opp_upload_dat(set->mesh_relation_dat);
    sync_size_to_device(set);
    const double k = const_real("CONST_charge") / const_real("CONST_mass") * const_real("CONST_dt");
    oppx_tet_mesh mesh{};
    mesh.n_cells = arg0.dat->set->size;
    mesh.ef = (const double*)arg0.dat->data_d;
    oppx_fempic_parts parts{set->set_capacity, (double*)arg1.dat->data_d, (double*)arg2.dat->data_d, nullptr,
                            (int32_t*)set->mesh_relation_dat->data_d};
    check(oppx_fempic_push(set->size, &parts, &mesh, k, const_real("CONST_dt"), pdev(set).counters, nullptr),
          "calculate_new_pos_vel");
    args_written_on_device(3, args);
    opp_profiler->end("calculate_new_pos_vel_kernel");
}

// apps_codegen/app_fempic_cg/hip/move_kernel_loop.hpp:191-407; elemental apps/app_fempic/kernels.h:106-161
void opp_particle_move__move_kernel(opp_set set, opp_map c2c_map, opp_map p2c_map, opp_arg arg0, opp_arg arg1,
                                    opp_arg arg2, opp_arg arg3) {
    opp_profiler->start("move_kernel");
    opp_arg args[4] = {arg0, arg1, arg2, arg3};
    args_to_device(4, args);
    if (p2c_map->p2c_dat->dirty_hd == Dirty::Device) opp_upload_dat(p2c_map->p2c_dat);
    opp_init_particle_move(set, 4, args);
    sync_size_to_device(set);
    ensure_capacity_scratch(set);
    oppx_tet_mesh mesh{};
    mesh.n_cells = c2c_map->from->size;
    mesh.det = (const double*)arg3.dat->data_d;
    mesh.aux = tet_aux(arg2.dat, c2c_map);
    oppx_fempic_parts parts{set->set_capacity, (double*)arg0.dat->data_d, nullptr, (double*)arg1.dat->data_d,
                            (int32_t*)p2c_map->p2c_dat->data_d};
    oppx_dh_grid grid{};
    const oppx_dh_grid* dh = nullptr;
    if (useGlobalMove && cellMapper) {
        const opp_point& mn = cellMapper->boundingBox->getGlobalMin();
        grid.cell = cellMapper->structMeshToCellMapping_d;
        grid.dims[0] = (int64_t)cellMapper->globalGridDimsX;
        grid.dims[1] = (int64_t)cellMapper->globalGridDimsY;
        grid.dims[2] = (int64_t)cellMapper->globalGridDimsZ;
        grid.min[0] = mn.x;
        grid.min[1] = mn.y;
        grid.min[2] = mn.z;
        grid.inv_gs = cellMapper->oneOverGridSpacing;
        dh = &grid;
    }
    ParticleDev& pd = pdev(set);
    check(oppx_fempic_move(set->size, &parts, &mesh, dh, pd.counters, pd.hole_bits, pd.scratch, pd.scratch_bytes,
                           nullptr, nullptr),
          "move_kernel");
    arg1.dat->dirty_hd = Dirty::Host;
    p2c_map->p2c_dat->dirty_hd = Dirty::Host;
    opp_finalize_particle_move(set);
    opp_profiler->end("move_kernel");
}

// apps_codegen/app_fempic_cg/hip/deposit_charge_on_nodes_kernel_loop.hpp:39-95 (FP64 atomics)
void opp_par_loop_all__deposit_charge_on_nodes_kernel(opp_set set, opp_arg arg0, opp_arg arg1, opp_arg arg2,
                                                      opp_arg arg3, opp_arg arg4) {
    opp_profiler->start("deposit_charge_on_nodes_kernel");
    opp_arg args[5] = {arg0, arg1, arg2, arg3, arg4};
    args_to_device(5, args);
    if (set->mesh_relation_dat->dirty_hd == Dirty::Device) opp_upload_dat(set->mesh_relation_dat);
    sync_size_to_device(set);
    if (!arg1.map || arg1.map->dim != 4 || arg1.idx != 0 || arg4.idx != 3)
        opp_abort("deposit_charge_on_nodes: expects n_charge_den through c2n indices 0..3");
    oppx_tet_mesh mesh{};
    mesh.n_cells = arg1.map->from->size;
    mesh.c2n = arg1.map->map_d;
    oppx_fempic_parts parts{set->set_capacity, nullptr, nullptr, (double*)arg0.dat->data_d,
                            (int32_t*)set->mesh_relation_dat->data_d};
    check(oppx_fempic_deposit_atomic(set->size, &parts, &mesh, (double*)arg1.dat->data_d, pdev(set).counters, nullptr),
          "deposit_charge_on_nodes");
    arg1.dat->dirty_hd = Dirty::Host;
    opp_profiler->end("deposit_charge_on_nodes_kernel");
}

// ---------------------------------------------------------------------------------------------
// FEM-PIC injection and the loops either side of the field solve (fempic_opp.cpp:41-49)

// Device workspace of the OPP_MAX globals: scratch for the partials and the result.
static void* g_max_scratch = nullptr;
static size_t g_max_scratch_bytes = 0;
static double* g_max_result = nullptr;

// An OPP_MAX global (`gbl`, host double holding the initial value) folded over n device values.
static void max_global(const opp_arg& gbl, const double* values, int64_t n, int abs_values, const char* what) {
    if (gbl.argtype != OPP_ARG_GBL || !gbl.data || gbl.acc != OPP_MAX) opp_abort(std::string(what) + ": expects an OPP_MAX global");
    const size_t need = oppx_max_scratch_bytes(n);
    if (need > g_max_scratch_bytes) {
        dev_free(g_max_scratch);
        g_max_scratch = dev_alloc(need);
        g_max_scratch_bytes = need;
    }
    if (!g_max_result) g_max_result = (double*)dev_alloc(sizeof(double));
    double init;
    memcpy(&init, gbl.data, sizeof(double));
    check(oppx_max_reduce(n, values, abs_values, init, g_max_result, g_max_scratch, g_max_scratch_bytes, nullptr), what);
    check(oppx_memcpy_dtoh(gbl.data, g_max_result, sizeof(double), nullptr), what);
}

void oppx_api::free_field_loop_buffers() {
    dev_free(g_max_scratch);
    dev_free(g_max_result);
    g_max_scratch = nullptr;
    g_max_result = nullptr;
    g_max_scratch_bytes = 0;
}

// apps_codegen/app_fempic_cg/hip/init_boundary_pot_kernel_loop.hpp; elemental apps/app_fempic/kernels.h:48-59
void opp_par_loop_all__init_boundary_pot_kernel(opp_set set, opp_arg arg0, opp_arg arg1) {
    opp_profiler->start("init_boundary_pot_kernel");
    opp_arg args[2] = {arg0, arg1};
    args_to_device(2, args);
    check(oppx_fempic_boundary_potential(set->size, (const int32_t*)arg0.dat->data_d, const_real("CONST_wall_potential"),
                                         (double*)arg1.dat->data_d, nullptr),
          "init_boundary_pot");
    args_written_on_device(2, args);
    opp_profiler->end("init_boundary_pot_kernel");
}

// apps_codegen/app_fempic_cg/hip/inject_ions_kernel_loop.hpp; elemental apps/app_fempic/kernels.h:63-90. The
// injected particles [size - diff, size) hold their inlet face (opp_inc_part_count_with_distribution); the
// random pairs are read from the start of arg9's dat.
void opp_par_loop_injected__inject_ions_kernel(opp_set set, opp_arg arg0, opp_arg arg1, opp_arg arg2, opp_arg arg3,
                                               opp_arg arg4, opp_arg arg5, opp_arg arg6, opp_arg arg7, opp_arg arg8,
                                               opp_arg arg9) {
    opp_profiler->start("inject_ions_kernel");
    opp_arg args[10] = {arg0, arg1, arg2, arg3, arg4, arg5, arg6, arg7, arg8, arg9};
    args_to_device(10, args);
    if (set->mesh_relation_dat->dirty_hd == Dirty::Device) opp_upload_dat(set->mesh_relation_dat);
    const int64_t count = set->diff, first = (int64_t)set->size - set->diff;
    if (arg3.argtype != OPP_ARG_MAP || !arg3.data_d || arg3.dim != 1)
        opp_abort("inject_ions: expects the inlet-face -> cell map as argument 3");
    if (!arg4.map || arg4.map->dim != 1) opp_abort("inject_ions: expects c_ef through the inlet-face -> cell map");
    if (arg9.dat->set->size < count) opp_abort("inject_ions: fewer random pairs than injected particles");
    if (count > 0) {
        // every inlet face must name a tet: then the device never meets an index out of range
        const int n_faces = arg3.size, n_cells = arg4.dat->set->size;
        for (int f = 0; f < n_faces; f++)
            if (((const int*)arg3.data)[f] < 0 || ((const int*)arg3.data)[f] >= n_cells)
                opp_abort("inject_ions: inlet face " + std::to_string(f) + " names no tet");
        oppx_inlet in{};
        in.n_faces = n_faces;
        in.face_cell = (const int32_t*)arg3.data_d;
        in.u = (const double*)arg5.dat->data_d;
        in.v = (const double*)arg6.dat->data_d;
        in.normal = (const double*)arg7.dat->data_d;
        in.node_pos = (const double*)arg8.dat->data_d;
        oppx_fempic_parts parts{set->set_capacity, (double*)arg0.dat->data_d, (double*)arg1.dat->data_d, nullptr,
                                (int32_t*)set->mesh_relation_dat->data_d};
        check(oppx_fempic_inject(first, count, &parts, &in, n_cells, (const double*)arg4.dat->data_d,
                                 (const double*)arg9.dat->data_d, arg9.dat->set->set_capacity,
                                 const_real("CONST_ion_velocity"), const_real("CONST_charge"), const_real("CONST_mass"),
                                 const_real("CONST_dt"), nullptr, nullptr),
              "inject_ions");
    }
    args_written_on_device(10, args);
    set->mesh_relation_dat->dirty_hd = Dirty::Host;
    opp_profiler->end("inject_ions_kernel");
}

// apps_codegen/app_fempic_cg/hip/compute_node_charge_density_kernel_loop.hpp; elemental kernels.h:178-183
void opp_par_loop_all__compute_node_charge_density_kernel(opp_set set, opp_arg arg0, opp_arg arg1) {
    opp_profiler->start("compute_node_charge_density_kernel");
    opp_arg args[2] = {arg0, arg1};
    args_to_device(2, args);
    check(oppx_fempic_node_charge_density(set->size, (double*)arg0.dat->data_d, (const double*)arg1.dat->data_d,
                                          const_real("CONST_spwt"), nullptr),
          "compute_node_charge_density");
    args_written_on_device(2, args);
    opp_profiler->end("compute_node_charge_density_kernel");
}

// apps_codegen/app_fempic_cg/hip/compute_electric_field_kernel_loop.hpp; elemental kernels.h:186-202
void opp_par_loop_all__compute_electric_field_kernel(opp_set set, opp_arg arg0, opp_arg arg1, opp_arg arg2,
                                                     opp_arg arg3, opp_arg arg4, opp_arg arg5) {
    opp_profiler->start("compute_electric_field_kernel");
    opp_arg args[6] = {arg0, arg1, arg2, arg3, arg4, arg5};
    args_to_device(6, args);
    if (!arg2.map || arg2.map->dim != 4 || arg2.idx != 0 || arg3.idx != 1 || arg4.idx != 2 || arg5.idx != 3 ||
        arg2.dat != arg5.dat)
        opp_abort("compute_electric_field: expects the node potential through c2n indices 0..3");
    check(oppx_fempic_electric_field(set->size, (double*)arg0.dat->data_d, (const double*)arg1.dat->data_d,
                                     arg2.map->map_d, (const double*)arg2.dat->data_d, arg2.dat->set->size, nullptr,
                                     nullptr),
          "compute_electric_field");
    args_written_on_device(6, args);
    opp_profiler->end("compute_electric_field_kernel");
}

// apps_codegen/app_fempic_cg/hip/get_max_cef_kernel_loop.hpp; elemental kernels.h:216-224
void opp_par_loop_all__get_max_cef_kernel(opp_set set, opp_arg arg0, opp_arg arg1) {
    opp_profiler->start("get_max_cef_kernel");
    opp_arg args[2] = {arg0, arg1};
    args_to_device(2, args);
    max_global(arg1, (const double*)arg0.dat->data_d, (int64_t)set->size * arg0.dat->dim, 0, "get_max_cef");
    opp_profiler->end("get_max_cef_kernel");
}

// apps_codegen/app_fempic_cg/hip/get_final_max_values_kernel_loop.hpp; elemental kernels.h:205-214
void opp_par_loop_all__get_final_max_values_kernel(opp_set set, opp_arg arg0, opp_arg arg1, opp_arg arg2,
                                                   opp_arg arg3) {
    opp_profiler->start("get_final_max_values_kernel");
    opp_arg args[4] = {arg0, arg1, arg2, arg3};
    args_to_device(4, args);
    max_global(arg1, (const double*)arg0.dat->data_d, set->size, 1, "get_final_max_values");
    max_global(arg3, (const double*)arg2.dat->data_d, set->size, 1, "get_final_max_values");
    opp_profiler->end("get_final_max_values_kernel");
}

// apps_codegen/app_fempic_cg/hip/move_kernel_loop.hpp:409-484 (opp_init_direct_hop_cg); overlay per
// opp_lib/src/seq/opp_direct_hop_seq.cpp:134-265, built with the candidate-restricted search
void opp_init_direct_hop_cg(double grid_spacing, const opp_dat c_gbl_id, const opp::BoundingBox& b_box,
                            opp_map c2c_map, opp_map p2c_map, opp_arg arg0, opp_arg arg1, opp_arg arg2, opp_arg arg3) {
    (void)c2c_map;
    (void)p2c_map;
    (void)arg0;
    (void)arg1;
    opp_profiler->start("Setup_Mover");
    useGlobalMove = opp_params->get<OPP_BOOL>("opp_global_move", false);
    if (useGlobalMove) {
        boundingBox = std::make_shared<opp::BoundingBox>(b_box);
        cellMapper = std::make_shared<opp::CellMapper>(boundingBox, grid_spacing);
        const int64_t grid_n = (int64_t)cellMapper->globalGridDimsX * (int64_t)cellMapper->globalGridDimsY *
                               (int64_t)cellMapper->globalGridDimsZ;
        char name[256];
        const double ex[3] = {b_box.domain_expansion.x, b_box.domain_expansion.y, b_box.domain_expansion.z};
        check(oppx_dh_overlay_filename(c_gbl_id->set->size, grid_spacing, ex, name, sizeof(name)), "overlay name");
        if (!opp_params->get<OPP_BOOL>("opp_dh_data_generate", false)) {
            // generateStructuredMeshFromFile (opp_lib/src/seq/opp_direct_hop_seq.cpp:89-130)
            std::string path = name;
            const std::string folder = opp_params->get<OPP_STRING>("opp_dh_mesh_folder", false);
            if (folder != "NOT FOUND") path = folder + path;
            check(oppx_dh_overlay_load(path.c_str(), cellMapper->structMeshToCellMapping, grid_n),
                  ("reading direct-hop overlay " + path).c_str());
        } else {
            const opp_point& mn = boundingBox->getGlobalMin();
            const opp_point& mx = boundingBox->getGlobalMax();
            const double lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
            opp_dat c_vol = arg2.dat, c_det = arg3.dat;
            if (c_vol->dirty_hd == Dirty::Host) opp_download_dat(c_vol);
            if (c_det->dirty_hd == Dirty::Host) opp_download_dat(c_det);
            if (c_gbl_id->dirty_hd == Dirty::Host) opp_download_dat(c_gbl_id);
            check(oppx_dh_build_overlay_geom(c_det->set->size, (const double*)c_det->data, (const double*)c_vol->data,
                                             (const int*)c_gbl_id->data, lo, hi, grid_spacing,
                                             cellMapper->structMeshToCellMapping),
                  "direct-hop overlay");
            if (OPP_dh_data_dump)  // opp_lib/src/seq/opp_direct_hop_seq.cpp:277-298, to the working directory
                check(oppx_dh_overlay_save(name, cellMapper->structMeshToCellMapping, grid_n), "writing overlay");
        }
        cellMapper->hostToDeviceTransferMappings();
    }
    opp_profiler->end("Setup_Mover");
}

// =============================================================================================
// NESO-advection (mdir)
// =============================================================================================

static const oppx_quad_cell* quad_cells(opp_dat c_pos_ll, opp_map c2c) {
    const int64_t n = c2c->from->size;
    if (g_quad.ll_d != c_pos_ll->data_d || g_quad.c2c_d != c2c->map_d || g_quad.n != n || !g_quad.cells) {
        dev_free(g_quad.cells);
        g_quad.cells = (oppx_quad_cell*)dev_alloc((size_t)n * sizeof(oppx_quad_cell));
        check(oppx_quad_cells_build((const double*)c_pos_ll->data_d, c2c->map_d, n, g_quad.cells, nullptr), "quad cells");
        g_quad.ll_d = c_pos_ll->data_d;
        g_quad.c2c_d = c2c->map_d;
        g_quad.n = n;
        g_quad.gnx = g_quad.gny = 0;
        // NESO's grid numbers cell (x, y) as x + nx * y, so cell 0's x-1 neighbour is nx - 1
        // (apps/app_neso_advection/advec_misc.h:128-140); the device check then compares every record
        const int nx = (c2c->map && n > 0) ? c2c->map[0] + 1 : 0;
        if (nx > 0 && n % nx == 0) {
            const double cw = const_real("CONST_cell_width");
            int32_t* bad_d = (int32_t*)dev_alloc(sizeof(int32_t));
            int32_t bad = -1;
            check(oppx_quad_grid_check(g_quad.cells, n, nx, (int)(n / nx), cw, bad_d, nullptr), "quad grid check");
            check(oppx_memcpy_dtoh(&bad, bad_d, sizeof(int32_t), nullptr), "quad grid check");
            dev_free(bad_d);
            if (bad == 0) {
                g_quad.gnx = nx;
                g_quad.gny = (int)(n / nx);
            }
        }
    }
    return g_quad.cells;
}

static void neso_consts(double& dt, double& cw, double& ex, double& ey) {
    dt = const_real("CONST_dt");
    cw = const_real("CONST_cell_width");
    int bytes = 0;
    const char* p = opp_const_data("CONST_extents", &bytes);
    if (!p || bytes < 2 * (int)sizeof(double)) opp_abort("constant not declared: CONST_extents");
    memcpy(&ex, p, sizeof(double));
    memcpy(&ey, p + sizeof(double), sizeof(double));
}

// apps_codegen/app_neso_advection_mdir_cg/hip/update_pos_kernel_loop.hpp; elemental apps/app_neso_advection/kernels.h:41-54
void opp_par_loop_all__update_pos_kernel(opp_set set, opp_arg arg0, opp_arg arg1) {
    opp_profiler->start("update_pos_kernel");
    opp_arg args[2] = {arg0, arg1};
    args_to_device(2, args);
    sync_size_to_device(set);
    double dt, cw, ex, ey;
    neso_consts(dt, cw, ex, ey);
    oppx_neso_parts parts{set->set_capacity, (double*)arg1.dat->data_d, (const double*)arg0.dat->data_d, nullptr,
                          nullptr};
    check(oppx_neso_step(set->size, &parts, nullptr, set->cells_set->size, dt, cw, ex, ey, OPPX_NESO_UPDATE,
                         pdev(set).counters, nullptr, nullptr),
          "update_pos");
    args_written_on_device(2, args);
    opp_profiler->end("update_pos_kernel");
}

// apps_codegen/app_neso_advection_mdir_cg/hip/move_kernel_loop.hpp:213-420; elemental kernels.h:56-113
void opp_particle_move__move_kernel(opp_set set, opp_map c2c_map, opp_map p2c_map, opp_arg arg0, opp_arg arg1,
                                    opp_arg arg2) {
    opp_profiler->start("move_kernel");
    opp_arg args[3] = {arg0, arg1, arg2};
    args_to_device(3, args);
    if (p2c_map->p2c_dat->dirty_hd == Dirty::Device) opp_upload_dat(p2c_map->p2c_dat);
    opp_init_particle_move(set, 3, args);
    sync_size_to_device(set);
    double dt, cw, ex, ey;
    neso_consts(dt, cw, ex, ey);
    oppx_neso_parts parts{set->set_capacity, (double*)arg0.dat->data_d, nullptr, (int32_t*)arg1.dat->data_d,
                          (int32_t*)p2c_map->p2c_dat->data_d};
    const oppx_quad_cell* cells = quad_cells(arg2.dat, c2c_map);
    if (g_quad.gnx > 0)
        check(oppx_neso_step_grid(set->size, &parts, g_quad.gnx, g_quad.gny, dt, cw, ex, ey, OPPX_NESO_MOVE,
                                  pdev(set).counters, nullptr, nullptr),
              "move_kernel");
    else
        check(oppx_neso_step(set->size, &parts, cells, c2c_map->from->size, dt, cw, ex, ey, OPPX_NESO_MOVE,
                             pdev(set).counters, nullptr, nullptr),
              "move_kernel");
    arg1.dat->dirty_hd = Dirty::Host;
    p2c_map->p2c_dat->dirty_hd = Dirty::Host;
    oppx_move_counters c;
    read_counters(set, &c);
    if (c.faults) opp_abort("particle move: " + std::to_string(c.faults) + " walks hit an invalid cell or the hop cap");
    OPP_move_max_hops = c.max_hops;
    set->particle_remove_count = 0;
    opp_profiler->end("move_kernel");
}

// apps_codegen/app_neso_advection_mdir_cg/hip/verify_kernel_loop.hpp; elemental apps/app_neso_advection/kernels.h:115-150.
// arg0 p_pos (OPP_READ), arg1 the cells' global index through p2c (OPP_READ), arg2 the int OPP_INC global
// that receives the number of particles outside their cell.
void opp_par_loop_all__verify_kernel(opp_set set, opp_arg arg0, opp_arg arg1, opp_arg arg2) {
    opp_profiler->start("verify_kernel");
    if (arg2.argtype != OPP_ARG_GBL || !arg2.data || arg2.acc != OPP_INC)
        opp_abort("verify_kernel: expects an int OPP_INC global");
    opp_arg args[2] = {arg0, arg1};
    args_to_device(2, args);
    if (set->mesh_relation_dat->dirty_hd == Dirty::Device) opp_upload_dat(set->mesh_relation_dat);
    sync_size_to_device(set);
    int bytes = 0;
    const char* nd = opp_const_data("CONST_ndimcells", &bytes);
    if (!nd || bytes < (int)sizeof(int)) opp_abort("constant not declared: CONST_ndimcells");
    int nx = 0;
    memcpy(&nx, nd, sizeof(int));
    if (!g_verify_count) g_verify_count = (int32_t*)dev_alloc(sizeof(int32_t));
    check(oppx_memset_zero(g_verify_count, sizeof(int32_t), nullptr), "verify_kernel");
    check(oppx_neso_verify(set->size, set->set_capacity, (const double*)arg0.dat->data_d,
                           (const int32_t*)set->mesh_relation_dat->data_d, (const int32_t*)arg1.dat->data_d,
                           arg1.dat->set->size, nx, const_real("CONST_cell_width"), pdev(set).counters, g_verify_count,
                           nullptr),
          "verify_kernel");
    int32_t count = 0;
    check(oppx_memcpy_dtoh(&count, g_verify_count, sizeof(int32_t), nullptr), "verify_kernel");
    oppx_move_counters c;
    read_counters(set, &c);
    if (c.faults) opp_abort("verify_kernel: " + std::to_string(c.faults) + " particles hold a cell index outside the mesh");
    int total = 0;
    memcpy(&total, arg2.data, sizeof(int));
    total += count;
    memcpy(arg2.data, &total, sizeof(int));
    opp_profiler->end("verify_kernel");
}

// =============================================================================================
// CabanaPIC
// =============================================================================================

// apps_codegen/app_cabanapic_cg/hip/move_deposit_kernel_loop.hpp; elemental apps/app_cabanapic/kernels.h:109-353.
// c2c_map is the six-face non-ghost neighbour map (x-, y-, z-, x+, y+, z+).
void opp_particle_move__move_deposit_kernel(opp_set set, opp_map c2c_map, opp_map p2c_map, opp_arg arg0, opp_arg arg1,
                                            opp_arg arg2, opp_arg arg3, opp_arg arg4, opp_arg arg5) {
    opp_profiler->start("move_deposit_kernel");
    opp_arg args[6] = {arg0, arg1, arg2, arg3, arg4, arg5};
    args_to_device(6, args);
    if (p2c_map->p2c_dat->dirty_hd == Dirty::Device) opp_upload_dat(p2c_map->p2c_dat);
    if (c2c_map->dim != 6) opp_abort("move_deposit: expects the six-face neighbour map");
    opp_init_particle_move(set, 6, args);
    sync_size_to_device(set);
    int bytes = 0;
    const char* cd = opp_const_data("CONST_cdt_d", &bytes);
    if (!cd || bytes < 3 * (int)sizeof(double)) opp_abort("constant not declared: CONST_cdt_d");
    double cdt_d[3];
    memcpy(cdt_d, cd, sizeof(cdt_d));
    oppx_cabana_parts parts{set->set_capacity, (double*)arg1.dat->data_d, (double*)arg0.dat->data_d,
                            (double*)arg2.dat->data_d, (const double*)arg3.dat->data_d,
                            (int32_t*)p2c_map->p2c_dat->data_d};
    oppx_cabana_mesh mesh{c2c_map->from->size, c2c_map->map_d, (const double*)arg4.dat->data_d};
    check(oppx_cabana_move_deposit(set->size, &parts, &mesh, (double*)arg5.dat->data_d, const_real("CONST_qsp"),
                                   const_real("CONST_qdt_2mc"), cdt_d, pdev(set).counters, nullptr, nullptr),
          "move_deposit_kernel");
    args_written_on_device(6, args);
    p2c_map->p2c_dat->dirty_hd = Dirty::Host;
    oppx_move_counters c;
    read_counters(set, &c);
    if (c.faults) opp_abort("particle move: " + std::to_string(c.faults) + " walks hit an invalid cell or the hop cap");
    OPP_move_max_hops = c.max_hops;
    opp_finalize_particle_move(set);
    opp_profiler->end("move_deposit_kernel");
}

// NESO's direct-hop setup (advec_opp.cpp:43): the periodic NESO walk does not use an overlay here
void opp_init_direct_hop_cg(double grid_spacing, const opp_dat c_gbl_id, const opp::BoundingBox& b_box,
                            opp_map c2c_map, opp_map p2c_map, opp_arg arg0, opp_arg arg1, opp_arg arg2) {
    (void)grid_spacing; (void)c_gbl_id; (void)b_box; (void)c2c_map; (void)p2c_map; (void)arg0; (void)arg1; (void)arg2;
    useGlobalMove = opp_params->get<OPP_BOOL>("opp_global_move", false);
    if (useGlobalMove) opp_abort("NESO-advection direct hop (2-D overlay) is not supported; set opp_global_move = false");
}
