// opp_lib API of OP-PIC, served by the MI355X-native oppx library.
//
// Applications and translator-generated code written against OP-PIC's opp_lib.h compile against
// this header unchanged: same function names, signatures and argument meaning
// (opp_lib/include/opp_lib.h:38-136), the same constants and enums (opp_lib/include/opp_defs.h:58-259),
// and the record layouts generated code dereferences (opp_arg, opp_set_core, opp_map_core,
// opp_dat_core; opp_lib/include/opp_defs.h:274-360) with the same members in the same order.
//
// Device state: mesh dats/maps are mirrored on the GPU in the host (AoS) layout; particle dats are
// SoA with stride = set capacity. The hot-path loops (push, direct-hop move, deposit for FEM-PIC;
// drift + move for NESO-advection) run as gfx950 kernels through include/oppx.h.
#pragma once

// The standard headers the reference's opp_defs.h and opp_profiler.h include
// (opp_lib/include/opp_defs.h:35-56, opp_lib/include/opp_profiler.h:35-42):
// generated code and application headers use std::mt19937, PRId64, std::ostringstream, ... without
// including them.
#include <inttypes.h>
#include <sys/time.h>
#include <unistd.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stack>
#include <string>
#include <vector>

#define OPP_INT int
#define OPP_REAL double
#define OPP_BOOL bool
#define OPP_STRING std::string

// single-process build: the MPI types of the reference collapse to placeholders
#define Comm int
#define GlobalParticleMover int
#define THRUST_REAL void
#define THRUST_INT void

namespace opp {
class BoundingBox;
class CellMapper;
class Params;
class Profiler;
}  // namespace opp

constexpr double MAX_REAL = std::numeric_limits<double>::max();
constexpr double MIN_REAL = std::numeric_limits<double>::min();
constexpr int MAX_INT = std::numeric_limits<int>::max();
constexpr int MIN_INT = std::numeric_limits<int>::min();

#define OPP_DBG false
#define OPP_ROOT 0
#define OPP_RUN_ON_ROOT(command) if (OPP_rank == OPP_ROOT command)
#define MAX_CELL_INDEX INT_MAX
#define OPP_OUT_OF_SAVED_DOMAIN ULONG_MAX
#define X_HOPS 5
#define UNUSED_VAR(expr) { (void)(expr); }

#define OPP_READ 0
#define OPP_WRITE 1
#define OPP_RW 2
#define OPP_INC 3
#define OPP_MIN 4
#define OPP_MAX 5

#define OPP_ARG_GBL 0
#define OPP_ARG_DAT 1
#define OPP_ARG_MAP 2

#define OPP_DEFAULT_GPU_THREADS_PER_BLOCK 32
#define LOG_STR_LEN 100000
#define OPP_IS_VALID_PROCESS (OPP_rank >= 0)
#define OPP_RETURN_IF_INVALID_PROCESS if (OPP_rank < 0) return;
#define OPP_RETURN_NULL_IF_INVALID_PROCESS if (OPP_rank < 0) return nullptr;
#define OPP_RETURN_ZERO_IF_INVALID_PROCESS if (OPP_rank < 0) return 0;
#define OPP_RETURN_EMPTY_IF_INVALID_PROCESS if (OPP_rank < 0) return "";
#define OPP_RETURN_ARG_IF_INVALID_PROCESS if (OPP_rank < 0) return opp_arg();
#define UNUSED(expr) do { (void)(expr); } while (0)
#define ZERO_double 0.0
#define ZERO_float 0.0f
#define ZERO_int 0
#define ZERO_uint 0
#define ZERO_ll 0
#define ZERO_ull 0
#define ZERO_bool 0
#define OPP_REAL_ZERO 0.0
#define OPP_INT_ZERO 0
#define OPP_BOOL_ZERO 0

#ifndef MIN
#define MIN(a, b) ((a < b) ? (a) : (b))
#endif
#ifndef MAX
#define MAX(a, b) ((a > b) ? (a) : (b))
#endif
#define FMOD(a, b) ((a) - (b) * trunc((a) / (b)))

#include "opp_util.h"

enum opp_iterate_type { OPP_ITERATE_ALL = 1, OPP_ITERATE_INJECTED };
enum opp_move_status { OPP_MOVE_DONE = 0, OPP_NEED_MOVE, OPP_NEED_REMOVE };
enum opp_data_type { DT_INT = 0, DT_REAL };
enum DeviceType { Device_CPU = 1, Device_GPU = 2 };
enum Dirty { NotDirty = 0, Device = 1, Host = 2 };  // Device: device copy stale; Host: host copy stale
enum opp_mapping { OPP_Map_Default = 0, OPP_Map_from_Mesh_Rel, OPP_Map_from_Inj_part, OPP_Map_to_Mesh_Rel };
enum opp_reset { OPP_Reset_Core = 0, OPP_Reset_Set, OPP_Reset_ieh, OPP_Reset_inh, OPP_Reset_All };
enum opp_reduc_comm { OPP_Reduc_NO_Comm = 0, OPP_Reduc_SUM_Comm, OPP_Reduc_MIN_Comm, OPP_Reduc_MAX_Comm };
enum opp_fill_type { OPP_HoleFill_All = 0, OPP_Sort_All, OPP_Shuffle_All, OPP_Sort_Periodic, OPP_Shuffle_Periodic };

struct part_index {
    int start;
    int end;
};

struct opp_set_core;
typedef struct opp_set_core* opp_set;
struct opp_map_core;
typedef struct opp_map_core* opp_map;
struct opp_dat_core;
typedef struct opp_dat_core* opp_dat;
typedef int opp_access;
typedef int opp_arg_type;

// Loop argument record (members and order as generated code expects them).
struct opp_arg {
    int index;
    opp_dat dat;
    opp_map map;
    opp_dat p2c_map;
    int dim;
    int idx;
    int size;
    char* data;        // host
    char* data_d;      // device
    int* map_data;     // host
    int* map_data_d;   // device
    char const* type;
    opp_access acc;
    opp_arg_type argtype;
    int sent;
    int opt;
    opp_mapping mesh_mapping;
};

struct opp_set_core {
    int index;
    int size;            // live elements
    char const* name;
    int core_size;
    int exec_size;
    int nonexec_size;
    bool is_particle;
    int set_capacity;    // allocated elements (SoA stride of the device particle dats)
    int diff;
    int particle_size;
    std::vector<int>* indexes_to_remove;
    opp_dat mesh_relation_dat = NULL;
    std::vector<opp_dat>* particle_dats;
    std::map<int, part_index>* cell_index_v_part_index_map;
    int particle_remove_count;
    int* particle_remove_count_d;
    void* mpi_part_buffers;
    opp_set cells_set;
};

struct opp_map_core {
    int index;
    opp_set from;
    opp_set to;
    int dim;
    int* map;     // host, from->size x dim
    int* map_d;   // device, same layout
    char const* name;
    int user_managed;
    opp_dat p2c_dat;
};

struct opp_dat_core {
    int index;
    opp_set set;
    int dim;
    int size;            // bytes per element
    char* data;          // host, AoS
    char* data_d;        // device: AoS for mesh sets, SoA (stride set_capacity) for particle sets
    char* data_swap_d;   // device: second buffer of a particle dat (sort target)
    char const* type;
    char const* name;
    char* buffer_d;
    char* buffer_d_r;
    int dirtybit;
    Dirty dirty_hd;
    int user_managed;
    void* mpi_buffer;
    void* mpi_reduc_buffer;
    opp_reduc_comm reduc_comm;
    std::vector<char*>* thread_data;
    char** thread_data_d;
    bool is_cell_index;
    THRUST_INT* thrust_int;
    THRUST_REAL* thrust_real;
    THRUST_INT* thrust_int_sort;
    THRUST_REAL* thrust_real_sort;
    opp_map p2c_map;
};

struct opp_point {
    opp_point(double _x, double _y, double _z) : x(_x), y(_y), z(_z) {}
    opp_point() {}
    double x = 0.0;
    double y = 0.0;
    double z = 0.0;
};

struct opp_ipoint {
    opp_ipoint(int _x, int _y, int _z) : x(_x), y(_y), z(_z) {}
    opp_ipoint() {}
    int x = 0, y = 0, z = 0;
};

struct opp_uipoint {
    opp_uipoint(size_t _x, size_t _y, size_t _z) : x(_x), y(_y), z(_z) {}
    opp_uipoint() {}
    size_t x = 0, y = 0, z = 0;
};

constexpr double opp_zero_double16[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
constexpr int opp_zero_int16[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

namespace opp {

// Runtime configuration read from the params file named by argv[1]: lines
// "<BOOL|STRING|INT|REAL> key = value" (opp_lib/src/core/opp_params.cpp:11-64); a missing key gives
// INT_MAX / DBL_MAX / false / "NOT FOUND" (opp_lib/include/opp_params.h:62-99).
class Params {
   public:
    explicit Params(const std::string& file_name);
    template <typename T>
    T get(std::string key, bool error = true);
    template <typename T>
    void add(std::string key, T value);
    void write(std::ostream& out) const;

   private:
    std::map<std::string, OPP_INT> int_params;
    std::map<std::string, OPP_REAL> real_params;
    std::map<std::string, OPP_STRING> str_params;
    std::map<std::string, OPP_BOOL> bool_params;
};

template <> OPP_INT Params::get<OPP_INT>(std::string key, bool error);
template <> OPP_REAL Params::get<OPP_REAL>(std::string key, bool error);
template <> OPP_BOOL Params::get<OPP_BOOL>(std::string key, bool error);
template <> OPP_STRING Params::get<OPP_STRING>(std::string key, bool error);
template <> void Params::add<OPP_INT>(std::string key, OPP_INT value);
template <> void Params::add<OPP_REAL>(std::string key, OPP_REAL value);
template <> void Params::add<OPP_BOOL>(std::string key, OPP_BOOL value);
template <> void Params::add<OPP_STRING>(std::string key, OPP_STRING value);

// Named wall-clock regions (opp_lib/include/opp_profiler.h:66-349); device work inside a region
// is synchronised at end() so GPU time is attributed to it.
class Profiler {
   public:
    Profiler();
    void reg(std::string name);
    void start(std::string name);
    double end(std::string name);
    double getTime(std::string name) const;
    int getCount(std::string name) const;
    void printProfile(bool fine = false) const;

   private:
    struct Entry {
        double total = 0.0;
        int count = 0;
        double started = -1.0;
    };
    std::map<std::string, Entry> entries_;
    std::vector<std::string> order_;
};

// Axis-aligned bounding box of the direct-hop overlay (opp_lib/include/opp_direct_hop_core.h:42-209).
class BoundingBox {
   public:
    BoundingBox(int dim, opp_point minCoordinate, opp_point maxCoordinate);
    // From node positions, widened by `expansion` with the reference's component use and a 1e-12 offset.
    BoundingBox(const opp_dat node_pos_dat, int dim, const opp_point expansion);
    BoundingBox(const BoundingBox& other) = default;
    const opp_point& getLocalMin() const { return box_[0]; }
    const opp_point& getLocalMax() const { return box_[1]; }
    const opp_point& getGlobalMin() const { return box_[0]; }
    const opp_point& getGlobalMax() const { return box_[1]; }
    int getDim() const { return dim_; }
    bool isCoordinateInBoundingBox(const opp_point& p) const;
    bool isCoordinateInGlobalBoundingBox(const opp_point& p) const { return isCoordinateInBoundingBox(p); }
    opp_point domain_expansion;

   private:
    int dim_ = 0;
    opp_point box_[2];
};

// Direct-hop overlay (structured grid -> owning cell) and its device copy.
class CellMapper {
   public:
    CellMapper(const std::shared_ptr<BoundingBox> boundingBox, const double gridSpacing,
               const std::shared_ptr<Comm> comm = nullptr);
    ~CellMapper();
    const std::shared_ptr<BoundingBox> boundingBox;
    const double gridSpacing;
    const double oneOverGridSpacing;
    const int dim;
    size_t globalGridDimsX = 0, globalGridDimsY = 0, globalGridDimsZ = 0, globalGridDimsXY = 0;
    size_t globalGridSize = 0;
    OPP_INT* structMeshToCellMapping = nullptr;    // host
    OPP_INT* structMeshToCellMapping_d = nullptr;  // device
    size_t findStructuredCellIndex3D(const opp_point& position) const;
    size_t findStructuredCellIndex2D(const opp_point& position) const;
    int findClosestCellIndex(const size_t& structCellIdx) const { return structMeshToCellMapping[structCellIdx]; }
    void hostToDeviceTransferMappings();
};

}  // namespace opp

// ---------------------------------------------------------------------------------------------
// process-wide state (opp_lib/include/opp_defs.h:404-450)
extern std::shared_ptr<opp::BoundingBox> boundingBox;
extern std::shared_ptr<opp::CellMapper> cellMapper;
extern bool useGlobalMove;
extern std::shared_ptr<Comm> comm;
extern std::unique_ptr<GlobalParticleMover> globalMover;
extern double OPP_hybrid_balance;
extern int OPP_maps_base_index;
extern int OPP_auto_soa;
extern int OPP_gpu_direct;
extern double OPP_part_alloc_mult;
extern int OPP_fill_period;
extern opp_fill_type OPP_fill_type;
extern int OPP_mpi_part_alloc_mult;
extern int OPP_rank;
extern int OPP_comm_size;
extern int OPP_comm_iteration;
extern int OPP_max_comm_iteration;
extern int OPP_iter_start;
extern int OPP_iter_end;
extern int* OPP_mesh_relation_data;
extern int* OPP_mesh_relation_data_d;
extern int OPP_main_loop_iter;
extern int OPP_gpu_threads_per_block;
extern size_t OPP_gpu_shared_mem_per_block;
extern int OPP_part_cells_set_size;
extern int OPP_part_comm_count_per_iter;
extern int OPP_move_max_hops;
extern int OPP_move_moreX_hops;
extern bool OPP_dh_data_dump;
extern std::vector<opp_set> opp_sets;
extern std::vector<opp_map> opp_maps;
extern std::vector<opp_dat> opp_dats;
extern std::unique_ptr<opp::Params> opp_params;
extern std::unique_ptr<opp::Profiler> opp_profiler;
extern size_t opp_maxSavedDHGrid[3];
extern size_t opp_minSavedDHGrid[3];

// ---------------------------------------------------------------------------------------------
// API (opp_lib/include/opp_lib.h:38-136)
void opp_init(int argc, char** argv);
void opp_exit();
void opp_abort(std::string s = "");

opp_set opp_decl_set(int size, char const* name);
opp_set opp_decl_particle_set(int size, char const* name, opp_set cells_set);
opp_set opp_decl_particle_set(char const* name, opp_set cells_set);
opp_map opp_decl_map(opp_set from, opp_set to, int dim, int* imap, char const* name);
opp_dat opp_decl_dat(opp_set set, int dim, opp_data_type dtype, void* data, char const* name);
opp_map opp_decl_map_txt(opp_set from, opp_set to, int dim, const char* file_name, char const* name);
opp_dat opp_decl_dat_txt(opp_set set, int dim, opp_data_type dtype, const char* file_name, char const* name);

void opp_decl_const_impl(int dim, int size, char* data, const char* name);
template <typename T>
void opp_decl_const(int dim, T* data, const char* name) {
    opp_decl_const_impl(dim, sizeof(T), (char*)data, name);
}
// Bytes registered by opp_decl_const under `name` (nullptr if unknown); the library's loops read
// the app constants (CONST_dt, CONST_charge, ...) through this.
const char* opp_const_data(const char* name, int* bytes = nullptr);

opp_arg opp_arg_gbl_core(void* data, int dim, char const* typ, int elem_bytes, opp_access acc);
inline opp_arg opp_arg_gbl(double* data, int dim, char const* typ, opp_access acc) {
    return opp_arg_gbl_core(data, dim, typ, sizeof(double), acc);
}
inline opp_arg opp_arg_gbl(int* data, int dim, char const* typ, opp_access acc) {
    return opp_arg_gbl_core(data, dim, typ, sizeof(int), acc);
}
inline opp_arg opp_arg_gbl(const bool* data, int dim, char const* typ, opp_access acc) {
    return opp_arg_gbl_core((void*)data, dim, typ, sizeof(bool), acc);
}

opp_arg opp_arg_dat(opp_dat dat, int idx, opp_map map, int dim, const char* typ, opp_map p2c_map, opp_access acc,
                    bool offset = true);
opp_arg opp_arg_dat(opp_dat dat, int idx, opp_map map, opp_map p2c_map, opp_access acc, bool offset = true);
opp_arg opp_arg_dat(opp_dat dat, int idx, opp_map map, opp_access acc, bool offset = true);
opp_arg opp_arg_dat(opp_dat dat, opp_map p2c_map, opp_access acc, bool offset = true);
opp_arg opp_arg_dat(opp_dat dat, opp_access acc, bool offset = true);
opp_arg opp_arg_dat(opp_map data_map, opp_access acc, bool offset = true);
opp_arg opp_arg_dat(opp_map data_map, opp_map p2c_map, opp_access acc, bool offset = true);
opp_arg opp_arg_dat(opp_map data_map, int idx, opp_map map, opp_access acc, bool offset = true);
opp_arg opp_arg_dat(opp_map data_map, int idx, opp_map map, opp_map p2c_map, opp_access acc, bool offset = true);

void opp_increase_particle_count(opp_set part_set, int insert_count);
void opp_inc_part_count_with_distribution(opp_set part_set, int insert_count, opp_dat part_dist, bool calc_new = true);
void opp_reset_num_particles_to_insert(opp_set set);

void opp_init_particle_move(opp_set set, int nargs, opp_arg* args);
bool opp_finalize_particle_move(opp_set set);
void opp_particle_sort(opp_set set);

void opp_reset_dat_impl(opp_dat dat, char* val, opp_reset reset = OPP_Reset_All);
template <typename T>
void opp_reset_dat(opp_dat dat, T* val, opp_reset reset = OPP_Reset_All) {
    opp_reset_dat_impl(dat, (char*)val, reset);
}

void opp_upload_dat(opp_dat dat);
void opp_download_dat(opp_dat dat);
void opp_download_particle_set(opp_set set, bool force_download = false);
void opp_upload_particle_set(opp_set set, bool realloc = false);

template <typename T>
T* opp_get_data(opp_dat dat) {
    return (T*)dat->data;
}
inline OPP_INT* opp_get_data(opp_map map) {
    if (map->from->is_particle) return (OPP_INT*)map->p2c_dat->data;
    return map->map;
}

void opp_print_dat_to_txtfile(opp_dat dat, const char* file_prefix, const char* file_suffix);
void opp_print_map_to_txtfile(opp_map map, const char* file_prefix, const char* file_suffix);
void opp_dump_dat(opp_dat data);
opp_dat opp_fetch_data(opp_dat dat);

void opp_set_dirtybit(int nargs, opp_arg* args);
void opp_set_dirtybit_device(int nargs, opp_arg* args);

inline void opp_printf(const char* function, const char* format, ...) {
    char buf[4096];
    va_list args;
    va_start(args, format);
    vsnprintf(buf, sizeof(buf), format, args);
    va_end(args);
    printf("%s[%d][%d] - %s\n", function, OPP_rank, OPP_main_loop_iter, buf);
    fflush(stdout);
}

inline void opp_get_global_values(const int64_t value, int64_t& gbl_value, int64_t& gbl_max, int64_t& gbl_min) {
    gbl_value = gbl_max = gbl_min = value;
}
inline int64_t opp_get_global_value(int64_t value) { return value; }

// ---------------------------------------------------------------------------------------------
// Generated-loop entry points served by the library (translator naming rule:
// opp_translator/translator/cpp/translator/program.py:54-96; prototypes as in
// apps_codegen/app_fempic_cg/fempic_opp.cpp:41-50 and apps_codegen/app_neso_advection_mdir_cg/advec_opp.cpp:40-43).
// FEM-PIC
void opp_par_loop_all__calculate_new_pos_vel_kernel(opp_set, opp_arg, opp_arg, opp_arg);
void opp_particle_move__move_kernel(opp_set, opp_map, opp_map, opp_arg, opp_arg, opp_arg, opp_arg);
void opp_par_loop_all__deposit_charge_on_nodes_kernel(opp_set, opp_arg, opp_arg, opp_arg, opp_arg, opp_arg);
void opp_init_direct_hop_cg(double, const opp_dat, const opp::BoundingBox&, opp_map, opp_map, opp_arg, opp_arg,
                            opp_arg, opp_arg);
void opp_par_loop_all__init_boundary_pot_kernel(opp_set, opp_arg, opp_arg);
void opp_par_loop_injected__inject_ions_kernel(opp_set, opp_arg, opp_arg, opp_arg, opp_arg, opp_arg, opp_arg, opp_arg,
                                               opp_arg, opp_arg, opp_arg);
void opp_par_loop_all__compute_node_charge_density_kernel(opp_set, opp_arg, opp_arg);
void opp_par_loop_all__compute_electric_field_kernel(opp_set, opp_arg, opp_arg, opp_arg, opp_arg, opp_arg, opp_arg);
void opp_par_loop_all__get_max_cef_kernel(opp_set, opp_arg, opp_arg);
void opp_par_loop_all__get_final_max_values_kernel(opp_set, opp_arg, opp_arg, opp_arg, opp_arg);
// CabanaPIC (apps_codegen/app_cabanapic_cg/cabana_opp.cpp)
void opp_particle_move__move_deposit_kernel(opp_set, opp_map, opp_map, opp_arg, opp_arg, opp_arg, opp_arg, opp_arg,
                                            opp_arg);
// NESO-advection
void opp_par_loop_all__update_pos_kernel(opp_set, opp_arg, opp_arg);
void opp_par_loop_all__verify_kernel(opp_set, opp_arg, opp_arg, opp_arg);
void opp_particle_move__move_kernel(opp_set, opp_map, opp_map, opp_arg, opp_arg, opp_arg);
void opp_init_direct_hop_cg(double, const opp_dat, const opp::BoundingBox&, opp_map, opp_map, opp_arg, opp_arg, opp_arg);
