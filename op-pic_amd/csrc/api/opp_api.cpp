// opp_lib API runtime over the oppx C-ABI: registry of sets/maps/dats, host<->device mirrors,
// particle-count growth, fill policies, params, profiler, direct-hop overlay setup.
//
// Behaviour follows the reference core + hip backend (opp_lib/src/core/opp_lib_core.cpp:75-444,
// opp_lib/src/core/opp_increase_part_count_core.cpp:35-146, opp_lib/src/hip/opp_hip.cpp:221-577,
// opp_lib/src/hip/opp_increase_part_count.cpp:39-209) with these device-side choices: particle dats
// are SoA with stride = set capacity and grow by device-to-device copies; mesh dats/maps keep the
// host (AoS) layout on the device; move counters live on the device.

#include <algorithm>
#include <chrono>
#include <fstream>
#include <iostream>
#include <regex>
#include <sstream>
#include <unordered_map>

#include "opp_api_internal.h"

// ---------------------------------------------------------------------------------------------
// process-wide state
std::shared_ptr<opp::BoundingBox> boundingBox;
std::shared_ptr<opp::CellMapper> cellMapper;
bool useGlobalMove = false;
std::shared_ptr<Comm> comm;
std::unique_ptr<GlobalParticleMover> globalMover;
double OPP_hybrid_balance = 0.0;
int OPP_maps_base_index = 0;
int OPP_auto_soa = 0;
int OPP_gpu_direct = 0;
double OPP_part_alloc_mult = 1.0;
int OPP_fill_period = 4;
opp_fill_type OPP_fill_type = OPP_HoleFill_All;
int OPP_mpi_part_alloc_mult = 1;
int OPP_rank = 0;
int OPP_comm_size = 1;
int OPP_comm_iteration = 0;
int OPP_max_comm_iteration = 0;
int OPP_iter_start = 0;
int OPP_iter_end = 0;
int* OPP_mesh_relation_data = nullptr;
int* OPP_mesh_relation_data_d = nullptr;
int OPP_main_loop_iter = 0;
int OPP_gpu_threads_per_block = 256;
size_t OPP_gpu_shared_mem_per_block = 0;
int OPP_part_cells_set_size = 0;
int OPP_part_comm_count_per_iter = 0;
int OPP_move_max_hops = 0;
int OPP_move_moreX_hops = 0;
bool OPP_dh_data_dump = false;
std::vector<opp_set> opp_sets;
std::vector<opp_map> opp_maps;
std::vector<opp_dat> opp_dats;
std::unique_ptr<opp::Params> opp_params;
std::unique_ptr<opp::Profiler> opp_profiler;
size_t opp_maxSavedDHGrid[3] = {0, 0, 0};
size_t opp_minSavedDHGrid[3] = {0, 0, 0};

namespace oppx_api {

std::unordered_map<opp_set, ParticleDev> g_pdev;
std::map<std::string, std::vector<char>> g_consts;
std::vector<opp_map> g_particle_maps;  // maps out of particle sets (not in opp_maps)

void check(int rc, const char* what) {
    if (rc != OPPX_OK) opp_abort(std::string(what) + ": " + oppx_last_error());
}

void* dev_alloc(size_t bytes) {
    void* p = nullptr;
    check(oppx_device_malloc(&p, bytes), "device allocation");
    return p;
}

void dev_free(void* p) {
    if (p) oppx_device_free(p);
}

int elem_bytes(opp_dat dat) { return dat->size / dat->dim; }

ParticleDev& pdev(opp_set set) {
    ParticleDev& d = g_pdev[set];
    if (!d.counters) {
        d.counters = (oppx_move_counters*)dev_alloc(sizeof(oppx_move_counters));
        check(oppx_memset_zero(d.counters, sizeof(oppx_move_counters), nullptr), "counters");
        d.size_on_device = 0;
    }
    return d;
}

void sync_size_to_device(opp_set set) {
    ParticleDev& d = pdev(set);
    if (d.size_on_device != set->size) {
        const int32_t s = set->size;
        check(oppx_memcpy_htod(&d.counters->size, &s, sizeof(s), nullptr), "size upload");
        d.size_on_device = set->size;
    }
}

void read_counters(opp_set set, oppx_move_counters* out) {
    ParticleDev& d = pdev(set);
    check(oppx_memcpy_dtoh(out, d.counters, sizeof(*out), nullptr), "counters download");
    d.size_on_device = out->size;
}

void ensure_capacity_scratch(opp_set set) {
    ParticleDev& d = pdev(set);
    const int64_t cap = set->set_capacity;
    if (d.cap_scratch == cap) return;
    dev_free(d.hole_bits);
    dev_free(d.scratch);
    d.hole_bits = (uint64_t*)dev_alloc((size_t)oppx_hole_bits_words(cap) * sizeof(uint64_t));
    // one workspace serves the move's worklist, then the hole fill or the sort (never at once)
    d.scratch_bytes = std::max({oppx_holefill_scratch_bytes(cap), oppx_sort_scratch_bytes(cap),
                                oppx_fempic_work_bytes(cap)});
    d.scratch = dev_alloc(d.scratch_bytes);
    d.cap_scratch = cap;
}

// device arrays of a dat, sized for the set (capacity for particle sets)
void create_device_arrays(opp_dat dat) {
    const size_t n = dat->set->is_particle ? (size_t)dat->set->set_capacity : (size_t)dat->set->size;
    const size_t bytes = n * (size_t)dat->size;
    dev_free(dat->data_d);
    dev_free(dat->data_swap_d);
    dat->data_d = bytes ? (char*)dev_alloc(bytes) : nullptr;
    dat->data_swap_d = (bytes && dat->set->is_particle) ? (char*)dev_alloc(bytes) : nullptr;
}

// a loop is about to read/write the device copies of these args
void args_to_device(int nargs, opp_arg* args) {
    for (int i = 0; i < nargs; i++) {
        opp_arg& a = args[i];
        if (a.argtype != OPP_ARG_DAT || !a.dat) continue;
        if (a.dat->dirty_hd == Dirty::Device) opp_upload_dat(a.dat);
        a.data_d = a.dat->data_d;
    }
}

// after a device loop: written dats are newer on the device
void args_written_on_device(int nargs, opp_arg* args) {
    for (int i = 0; i < nargs; i++) {
        opp_arg& a = args[i];
        if (a.argtype == OPP_ARG_DAT && a.dat && (a.acc == OPP_WRITE || a.acc == OPP_RW || a.acc == OPP_INC))
            a.dat->dirty_hd = Dirty::Host;
    }
}

double now_seconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace oppx_api

using namespace oppx_api;

// ---------------------------------------------------------------------------------------------
// Params
namespace opp {

Params::Params(const std::string& file_name) {
    std::ifstream in(file_name);
    if (!in.is_open()) {
        std::cerr << "opp_init: cannot open params file '" << file_name << "'" << std::endl;
        return;
    }
    const std::regex line_re("(BOOL|STRING|INT|REAL)(\\s+)([aA-zZ]*)(\\s+)(\\=)(\\s+)(.*)");
    const std::regex truthy("^t(rue)?$", std::regex_constants::icase);
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty()) continue;
        std::smatch m;
        if (!std::regex_match(line, m, line_re)) continue;
        const std::string type = m[1], key = m[3];
        std::string value = m[7];
        while (!value.empty() && (value.back() == '\r' || value.back() == ' ')) value.pop_back();
        if (type == "STRING") add<OPP_STRING>(key, value);
        else if (type == "INT") add<OPP_INT>(key, std::stoi(value));
        else if (type == "REAL") add<OPP_REAL>(key, std::stod(value));
        else add<OPP_BOOL>(key, std::regex_match(value, truthy) || value == "1");
    }
}

template <> OPP_INT Params::get<OPP_INT>(std::string key, bool error) {
    auto it = int_params.find(key);
    if (it != int_params.end()) return it->second;
    if (error) std::cerr << "ERROR : " << key << " (OPP_INT) not found in configs" << std::endl;
    return INT_MAX;
}
template <> OPP_REAL Params::get<OPP_REAL>(std::string key, bool error) {
    auto it = real_params.find(key);
    if (it != real_params.end()) return it->second;
    if (error) std::cerr << "ERROR : " << key << " (OPP_REAL) not found in configs" << std::endl;
    return DBL_MAX;
}
template <> OPP_BOOL Params::get<OPP_BOOL>(std::string key, bool error) {
    auto it = bool_params.find(key);
    if (it != bool_params.end()) return it->second;
    if (error) std::cerr << "ERROR : " << key << " (BOOL) not found in configs" << std::endl;
    return false;
}
template <> OPP_STRING Params::get<OPP_STRING>(std::string key, bool error) {
    auto it = str_params.find(key);
    if (it != str_params.end()) return it->second;
    if (error) std::cerr << "ERROR : " << key << " (OPP_STRING) not found in configs" << std::endl;
    return "NOT FOUND";
}
template <> void Params::add<OPP_INT>(std::string key, OPP_INT v) { int_params[key] = v; }
template <> void Params::add<OPP_REAL>(std::string key, OPP_REAL v) { real_params[key] = v; }
template <> void Params::add<OPP_BOOL>(std::string key, OPP_BOOL v) { bool_params[key] = v; }
template <> void Params::add<OPP_STRING>(std::string key, OPP_STRING v) { str_params[key] = v; }

void Params::write(std::ostream& out) const {
    for (auto& kv : int_params) out << "INT " << kv.first << " = " << kv.second << "\n";
    for (auto& kv : real_params) out << "REAL " << kv.first << " = " << kv.second << "\n";
    for (auto& kv : bool_params) out << "BOOL " << kv.first << " = " << (kv.second ? "true" : "false") << "\n";
    for (auto& kv : str_params) out << "STRING " << kv.first << " = " << kv.second << "\n";
}

// ---------------------------------------------------------------------------------------------
// Profiler
Profiler::Profiler() {}

void Profiler::reg(std::string name) {
    if (!entries_.count(name)) {
        entries_[name] = Entry();
        order_.push_back(name);
    }
}

void Profiler::start(std::string name) {
    reg(name);
    entries_[name].started = now_seconds();
}

double Profiler::end(std::string name) {
    auto it = entries_.find(name);
    if (it == entries_.end() || it->second.started < 0) return 0.0;
    oppx_stream_sync(nullptr);  // attribute the region's device work to it
    const double dt = now_seconds() - it->second.started;
    it->second.total += dt;
    it->second.count += 1;
    it->second.started = -1.0;
    return dt;
}

double Profiler::getTime(std::string name) const {
    auto it = entries_.find(name);
    return it == entries_.end() ? 0.0 : it->second.total;
}

int Profiler::getCount(std::string name) const {
    auto it = entries_.find(name);
    return it == entries_.end() ? 0 : it->second.count;
}

void Profiler::printProfile(bool fine) const {
    (void)fine;
    printf("oppx profile: %-40s %8s %12s\n", "region", "count", "seconds");
    for (const auto& n : order_) {
        const Entry& e = entries_.at(n);
        if (e.count) printf("oppx profile: %-40s %8d %12.6f\n", n.c_str(), e.count, e.total);
    }
    fflush(stdout);
}

// ---------------------------------------------------------------------------------------------
// BoundingBox / CellMapper
BoundingBox::BoundingBox(int dim, opp_point minCoordinate, opp_point maxCoordinate) : dim_(dim) {
    box_[0] = minCoordinate;
    box_[1] = maxCoordinate;
}

BoundingBox::BoundingBox(const opp_dat node_pos_dat, int dim, const opp_point expansion)
    : domain_expansion(expansion), dim_(dim) {
    // node-based box (opp_lib/include/opp_direct_hop_core.h:70-124): min/max over the nodes, then
    // min.{x,y} -= (expansion.x + 1e-12), max.{x,y} += (expansion.y - 1e-12), z uses expansion.z
    if (dim != 2 && dim != 3) opp_abort("BoundingBox: only 2D/3D supported");
    if (node_pos_dat->dirty_hd == Dirty::Host) opp_download_dat(node_pos_dat);
    const double* p = (const double*)node_pos_dat->data;
    const int n = node_pos_dat->set->size;
    opp_point mn(MAX_REAL, MAX_REAL, MAX_REAL), mx(MIN_REAL, MIN_REAL, MIN_REAL);
    if (dim == 2) mn.z = mx.z = 0.0;
    for (int i = 0; i < n; i++) {
        mn.x = std::min(p[i * dim + 0], mn.x);
        mx.x = std::max(p[i * dim + 0], mx.x);
        mn.y = std::min(p[i * dim + 1], mn.y);
        mx.y = std::max(p[i * dim + 1], mx.y);
        if (dim == 3) {
            mn.z = std::min(p[i * dim + 2], mn.z);
            mx.z = std::max(p[i * dim + 2], mx.z);
        }
    }
    const double kOffset = 1e-12;
    if (n != 0) {
        mn.x -= (expansion.x + kOffset);
        mn.y -= (expansion.x + kOffset);
        mx.x += (expansion.y - kOffset);
        mx.y += (expansion.y - kOffset);
        if (dim == 3) {
            mn.z -= (expansion.z + kOffset);
            mx.z += (expansion.z - kOffset);
        }
    }
    box_[0] = mn;
    box_[1] = mx;
}

bool BoundingBox::isCoordinateInBoundingBox(const opp_point& p) const {
    return !(p.x < box_[0].x || p.x > box_[1].x || p.y < box_[0].y || p.y > box_[1].y || p.z < box_[0].z ||
             p.z > box_[1].z);
}

CellMapper::CellMapper(const std::shared_ptr<BoundingBox> bb, const double gs, const std::shared_ptr<Comm> c)
    : boundingBox(bb), gridSpacing(gs), oneOverGridSpacing(1.0 / gs), dim(bb->getDim()) {
    (void)c;
    const opp_point& mn = bb->getGlobalMin();
    const opp_point& mx = bb->getGlobalMax();
    const double lo[3] = {mn.x, mn.y, mn.z}, hi[3] = {mx.x, mx.y, mx.z};
    int64_t dims[3];
    check(oppx_dh_grid_dims(lo, hi, gs, dims), "CellMapper grid dims");
    globalGridDimsX = dims[0];
    globalGridDimsY = dims[1];
    globalGridDimsZ = dims[2];
    globalGridDimsXY = dims[0] * dims[1];
    globalGridSize = globalGridDimsXY * dims[2];
    opp_minSavedDHGrid[0] = opp_minSavedDHGrid[1] = opp_minSavedDHGrid[2] = 0;
    opp_maxSavedDHGrid[0] = globalGridDimsX;
    opp_maxSavedDHGrid[1] = globalGridDimsY;
    opp_maxSavedDHGrid[2] = globalGridDimsZ;
    structMeshToCellMapping = new OPP_INT[globalGridSize];
    for (size_t i = 0; i < globalGridSize; i++) structMeshToCellMapping[i] = MAX_CELL_INDEX;
}

CellMapper::~CellMapper() {
    delete[] structMeshToCellMapping;
    dev_free(structMeshToCellMapping_d);
}

size_t CellMapper::findStructuredCellIndex3D(const opp_point& p) const {
    const opp_point& mn = boundingBox->getGlobalMin();
    const double t[3] = {(p.x - mn.x) * oneOverGridSpacing, (p.y - mn.y) * oneOverGridSpacing,
                         (p.z - mn.z) * oneOverGridSpacing};
    const size_t dims[3] = {globalGridDimsX, globalGridDimsY, globalGridDimsZ};
    size_t idx[3];
    for (int d = 0; d < 3; d++) {
        if (!(t[d] > -1.0) || !(t[d] < (double)dims[d])) return OPP_OUT_OF_SAVED_DOMAIN;
        idx[d] = (size_t)(t[d] < 0 ? 0.0 : t[d]);
    }
    return idx[0] + idx[1] * globalGridDimsX + idx[2] * globalGridDimsXY;
}

size_t CellMapper::findStructuredCellIndex2D(const opp_point& p) const {
    const opp_point& mn = boundingBox->getGlobalMin();
    const size_t x = (size_t)((p.x - mn.x) * oneOverGridSpacing);
    const size_t y = (size_t)((p.y - mn.y) * oneOverGridSpacing);
    const size_t i = x + y * globalGridDimsX;
    return i >= globalGridSize ? MAX_CELL_INDEX : i;
}

void CellMapper::hostToDeviceTransferMappings() {
    if (!structMeshToCellMapping_d) structMeshToCellMapping_d = (OPP_INT*)dev_alloc(globalGridSize * sizeof(OPP_INT));
    check(oppx_memcpy_htod(structMeshToCellMapping_d, structMeshToCellMapping, globalGridSize * sizeof(OPP_INT),
                           nullptr),
          "overlay upload");
}

}  // namespace opp

// ---------------------------------------------------------------------------------------------
// init / exit
void opp_abort(std::string s) {
    std::cerr << "opp_abort: " << s << std::endl;
    std::exit(-1);
}

void opp_init(int argc, char** argv) {
    opp_sets.clear();
    opp_maps.clear();
    opp_dats.clear();
    if (argc < 2) opp_abort("opp_init: the params file must be argv[1]");
    opp_params = std::make_unique<opp::Params>(argv[1]);
    opp_profiler = std::make_unique<opp::Profiler>();
    const double mult = opp_params->get<OPP_REAL>("opp_allocation_multiple", false);
    OPP_part_alloc_mult = (mult == DBL_MAX) ? 1.0 : mult;
    const std::string fill = opp_params->get<OPP_STRING>("opp_fill", false);
    if (fill == "Sort_All") OPP_fill_type = OPP_Sort_All;
    else if (fill == "Shuffle_All") OPP_fill_type = OPP_Shuffle_All;
    else if (fill == "Sort_Periodic") OPP_fill_type = OPP_Sort_Periodic;
    else if (fill == "Shuffle_Periodic") OPP_fill_type = OPP_Shuffle_Periodic;
    else OPP_fill_type = OPP_HoleFill_All;
    const int period = opp_params->get<OPP_INT>("opp_fill_period", false);
    if ((OPP_fill_type == OPP_Sort_Periodic || OPP_fill_type == OPP_Shuffle_Periodic) && period != INT_MAX &&
        period >= 0)
        OPP_fill_period = period;
    const int tpb = opp_params->get<OPP_INT>("opp_threads_per_block", false);
    if (tpb != INT_MAX) OPP_gpu_threads_per_block = tpb;
    for (int n = 1; n < argc; n++) {  // argv overrides (opp_lib/src/core/opp_lib_core.cpp:149-184)
        const char* a = argv[n];
        if (const char* p = strstr(a, "OPP_ALLOC_MULT=")) OPP_part_alloc_mult = atof(p + 15);
        if (const char* p = strstr(a, "OPP_DH_DATA_DUMP=")) OPP_dh_data_dump = atoi(p + 17) != 0;
    }
    check(oppx_init(0), "opp_init device");
    opp_profiler->reg("move_kernel");
}

void opp_exit() {
    oppx_stream_sync(nullptr);
    if (opp_profiler && opp_params && opp_params->get<OPP_BOOL>("opp_profile_all", false))
        opp_profiler->printProfile(true);
    for (auto& kv : g_pdev) {
        dev_free(kv.second.counters);
        dev_free(kv.second.hole_bits);
        dev_free(kv.second.scratch);
        dev_free(kv.second.hops);
    }
    g_pdev.clear();
    free_loop_caches();
    for (auto& m : opp_maps) {
        free(m->map);
        dev_free(m->map_d);
        free((char*)m->name);
        free(m);
    }
    for (auto& m : g_particle_maps) {
        free((char*)m->name);
        free(m);
    }
    for (auto& d : opp_dats) {
        free(d->data);
        dev_free(d->data_d);
        dev_free(d->data_swap_d);
        delete d->thread_data;
        free((char*)d->name);
        free((char*)d->type);
        free(d);
    }
    for (auto& s : opp_sets) {
        delete s->indexes_to_remove;
        delete s->particle_dats;
        delete s->cell_index_v_part_index_map;
        free((char*)s->name);
        free(s);
    }
    opp_maps.clear();
    g_particle_maps.clear();
    opp_dats.clear();
    opp_sets.clear();
    cellMapper.reset();
    boundingBox.reset();
    g_consts.clear();
}

// ---------------------------------------------------------------------------------------------
// declarations (opp_lib/src/core/opp_lib_core.cpp:187-351; hip opp_lib/src/hip/opp_hip.cpp:332-373)
static char* copy_cstr(const char* s) {
    const size_t n = strlen(s) + 1;
    char* d = (char*)malloc(n);
    memcpy(d, s, n);
    return d;
}

opp_set opp_decl_set(int size, char const* name) {
    if (size < 0) opp_abort(std::string("opp_decl_set: negative size for ") + name);
    opp_set set = (opp_set)calloc(1, sizeof(opp_set_core));
    set->index = (int)opp_sets.size();
    set->size = size;
    set->name = copy_cstr(name);
    set->core_size = size;
    set->set_capacity = size;
    set->indexes_to_remove = new std::vector<int>();
    set->particle_dats = new std::vector<opp_dat>();
    set->cell_index_v_part_index_map = new std::map<int, part_index>();
    opp_sets.push_back(set);
    return set;
}

opp_set opp_decl_particle_set(int size, char const* name, opp_set cells_set) {
    opp_set set = opp_decl_set(size, name);
    set->is_particle = true;
    set->cells_set = cells_set;
    pdev(set);
    return set;
}

opp_set opp_decl_particle_set(char const* name, opp_set cells_set) { return opp_decl_particle_set(0, name, cells_set); }

static opp_dat decl_dat_core(opp_set set, int dim, const char* type, int elem, const char* data, const char* name) {
    if (!set) opp_abort(std::string("opp_decl_dat: invalid set for ") + name);
    if (dim <= 0) opp_abort(std::string("opp_decl_dat: non-positive dim for ") + name);
    opp_dat dat = (opp_dat)calloc(1, sizeof(opp_dat_core));
    dat->index = (int)opp_dats.size();
    dat->set = set;
    dat->dim = dim;
    dat->size = dim * elem;
    const size_t n = set->is_particle ? (size_t)set->set_capacity : (size_t)set->size;
    if (n > 0) {
        dat->data = (char*)calloc(n, (size_t)dat->size);
        if (data) memcpy(dat->data, data, (size_t)set->size * (size_t)dat->size);
    }
    dat->name = copy_cstr(name);
    dat->type = copy_cstr(type);
    dat->user_managed = 1;
    dat->dirty_hd = Dirty::NotDirty;
    dat->thread_data = new std::vector<char*>();
    set->particle_size += dat->size;
    opp_dats.push_back(dat);
    if (set->is_particle) set->particle_dats->push_back(dat);
    create_device_arrays(dat);
    opp_upload_dat(dat);
    return dat;
}

opp_dat opp_decl_dat(opp_set set, int dim, opp_data_type dtype, void* data, char const* name) {
    const bool real = (dtype == DT_REAL);
    return decl_dat_core(set, dim, real ? "double" : "int", real ? (int)sizeof(double) : (int)sizeof(int),
                         (const char*)data, name);
}

opp_map opp_decl_map(opp_set from, opp_set to, int dim, int* imap, char const* name) {
    if (!from || !to) opp_abort(std::string("opp_decl_map: invalid set for ") + name);
    if (dim <= 0) opp_abort(std::string("opp_decl_map: non-positive dim for ") + name);
    if (to->is_particle) opp_abort(std::string("opp_decl_map: cannot map to a particle set: ") + name);
    opp_map map = (opp_map)calloc(1, sizeof(opp_map_core));
    map->from = from;
    map->to = to;
    map->dim = dim;
    map->name = copy_cstr(name);
    map->user_managed = 1;
    if (!from->is_particle) {
        map->index = (int)opp_maps.size();
        const size_t n = (size_t)from->size * (size_t)dim;
        map->map = (int*)malloc(std::max<size_t>(n, 1) * sizeof(int));
        if (n) memcpy(map->map, imap, n * sizeof(int));
        if (OPP_maps_base_index == 1)
            for (size_t i = 0; i < n; i++) map->map[i]--;
        for (size_t i = 0; i < n; i++)
            if (map->map[i] < -1 || map->map[i] >= to->size)
                opp_abort(std::string("opp_decl_map: entry out of range in ") + name);
        map->map_d = n ? (int*)dev_alloc(n * sizeof(int)) : nullptr;
        if (n) check(oppx_memcpy_htod(map->map_d, map->map, n * sizeof(int), nullptr), "map upload");
        opp_maps.push_back(map);
    } else {
        // a particle -> cell map is the set's mesh relation dat (opp_lib/src/core/opp_lib_core.cpp:272-286)
        map->index = -1;
        map->p2c_dat = decl_dat_core(from, dim, "int", sizeof(int), (const char*)imap, name);
        from->mesh_relation_dat = map->p2c_dat;
        map->p2c_dat->is_cell_index = true;
        map->p2c_dat->p2c_map = map;
        g_particle_maps.push_back(map);
    }
    return map;
}

static void* load_txt(const char* file, int n, int dim, bool real) {
    std::ifstream in(file);
    if (!in.is_open()) opp_abort(std::string("cannot open ") + file);
    int rows = 0, cols = 0;
    in >> rows >> cols;  // "%d %d" header of opp_print_*_to_txtfile
    std::string dashes;
    int a, b;
    in >> dashes >> a >> b;
    if (rows != n || cols != dim) opp_abort(std::string("size mismatch in ") + file);
    const size_t cnt = (size_t)n * dim;
    void* out = malloc(std::max<size_t>(cnt, 1) * (real ? sizeof(double) : sizeof(int)));
    for (size_t i = 0; i < cnt; i++) {
        if (real) in >> ((double*)out)[i];
        else in >> ((int*)out)[i];
    }
    return out;
}

opp_map opp_decl_map_txt(opp_set from, opp_set to, int dim, const char* file_name, char const* name) {
    int* d = (int*)load_txt(file_name, from->size, dim, false);
    opp_map m = opp_decl_map(from, to, dim, d, name);
    free(d);
    return m;
}

opp_dat opp_decl_dat_txt(opp_set set, int dim, opp_data_type dtype, const char* file_name, char const* name) {
    void* d = load_txt(file_name, set->size, dim, dtype == DT_REAL);
    opp_dat dat = opp_decl_dat(set, dim, dtype, d, name);
    free(d);
    return dat;
}

void opp_decl_const_impl(int dim, int size, char* data, const char* name) {
    g_consts[name].assign(data, data + (size_t)dim * (size_t)size);
}

const char* opp_const_data(const char* name, int* bytes) {
    auto it = g_consts.find(name);
    if (it == g_consts.end()) return nullptr;
    if (bytes) *bytes = (int)it->second.size();
    return it->second.data();
}

// ---------------------------------------------------------------------------------------------
// args (opp_lib/src/core/opp_lib_core.cpp:354-405 and the hip overloads of opp_lib/src/hip/opp_hip.cpp)
opp_arg opp_arg_gbl_core(void* data, int dim, char const* typ, int eb, opp_access acc) {
    opp_arg a{};
    a.argtype = OPP_ARG_GBL;
    a.dim = dim;
    a.idx = -1;
    a.size = dim * eb;
    a.data = (char*)data;
    a.type = typ;
    a.acc = acc;
    a.opt = 1;
    return a;
}

static opp_arg arg_of_dat(opp_dat dat, int idx, opp_map map, opp_dat p2c, opp_access acc) {
    if (!dat) opp_abort("opp_arg_dat: dat is NULL");
    opp_arg a{};
    a.index = -1;
    a.argtype = OPP_ARG_DAT;
    a.dat = dat;
    a.map = map;
    a.p2c_map = p2c;
    a.dim = dat->dim;
    a.idx = idx;
    a.size = map ? (map->from->size + map->from->exec_size + map->from->nonexec_size) : -1;
    a.data = dat->data;
    a.data_d = dat->data_d;
    a.map_data = (idx == -1 || idx == -2 || !map) ? nullptr : map->map;
    a.map_data_d = (idx == -1 || idx == -2 || !map) ? nullptr : map->map_d;
    a.type = dat->type;
    a.acc = acc;
    a.opt = 1;
    return a;
}

static opp_arg arg_of_map(opp_map data_map, int idx, opp_map map, opp_dat p2c, opp_access acc) {
    if (!data_map) opp_abort("opp_arg_dat: map is NULL");
    if (data_map->from->is_particle) return arg_of_dat(data_map->p2c_dat, idx, map, p2c, acc);
    opp_arg a{};
    a.argtype = OPP_ARG_MAP;
    a.map = map;
    a.p2c_map = p2c;
    a.dim = data_map->dim;
    a.idx = idx;
    a.size = data_map->from->size;
    a.data = (char*)data_map->map;
    a.data_d = (char*)data_map->map_d;
    a.map_data = map ? map->map : nullptr;
    a.map_data_d = map ? map->map_d : nullptr;
    a.type = "int";
    a.acc = acc;
    a.opt = 1;
    return a;
}

opp_arg opp_arg_dat(opp_dat dat, int idx, opp_map map, int, const char*, opp_map p2c_map, opp_access acc, bool) {
    return arg_of_dat(dat, idx, map, p2c_map ? p2c_map->p2c_dat : nullptr, acc);
}
opp_arg opp_arg_dat(opp_dat dat, int idx, opp_map map, opp_map p2c_map, opp_access acc, bool) {
    return arg_of_dat(dat, idx, map, p2c_map ? p2c_map->p2c_dat : nullptr, acc);
}
opp_arg opp_arg_dat(opp_dat dat, int idx, opp_map map, opp_access acc, bool) {
    return arg_of_dat(dat, idx, map, nullptr, acc);
}
opp_arg opp_arg_dat(opp_dat dat, opp_map p2c_map, opp_access acc, bool) {
    return arg_of_dat(dat, -1, nullptr, p2c_map ? p2c_map->p2c_dat : nullptr, acc);
}
opp_arg opp_arg_dat(opp_dat dat, opp_access acc, bool) { return arg_of_dat(dat, -1, nullptr, nullptr, acc); }
opp_arg opp_arg_dat(opp_map data_map, opp_access acc, bool) { return arg_of_map(data_map, -1, nullptr, nullptr, acc); }
opp_arg opp_arg_dat(opp_map data_map, opp_map p2c_map, opp_access acc, bool) {
    return arg_of_map(data_map, -1, nullptr, p2c_map ? p2c_map->p2c_dat : nullptr, acc);
}
opp_arg opp_arg_dat(opp_map data_map, int idx, opp_map map, opp_access acc, bool) {
    return arg_of_map(data_map, idx, map, nullptr, acc);
}
opp_arg opp_arg_dat(opp_map data_map, int idx, opp_map map, opp_map p2c_map, opp_access acc, bool) {
    return arg_of_map(data_map, idx, map, p2c_map ? p2c_map->p2c_dat : nullptr, acc);
}

void opp_set_dirtybit(int nargs, opp_arg* args) {
    for (int n = 0; n < nargs; n++)
        if (args[n].opt == 1 && args[n].argtype == OPP_ARG_DAT &&
            (args[n].acc == OPP_INC || args[n].acc == OPP_WRITE || args[n].acc == OPP_RW))
            args[n].dat->dirty_hd = Dirty::Device;
}

void opp_set_dirtybit_device(int nargs, opp_arg* args) { args_written_on_device(nargs, args); }

// ---------------------------------------------------------------------------------------------
// transfers (AoS host <-> device; particle dats are SoA on the device)
void opp_upload_dat(opp_dat dat) {
    opp_set set = dat->set;
    const size_t n = (size_t)set->size;
    if (n && dat->data && dat->data_d) {
        if (set->is_particle) {
            void* tmp = dev_alloc(n * dat->size);
            check(oppx_memcpy_htod(tmp, dat->data, n * dat->size, nullptr), "upload");
            check(oppx_aos_to_soa(tmp, dat->data_d, (int64_t)n, dat->dim, set->set_capacity, elem_bytes(dat), nullptr),
                  "aos_to_soa");
            oppx_stream_sync(nullptr);
            dev_free(tmp);
        } else {
            check(oppx_memcpy_htod(dat->data_d, dat->data, n * dat->size, nullptr), "upload");
        }
    }
    dat->dirty_hd = Dirty::NotDirty;
}

void opp_download_dat(opp_dat dat) {
    opp_set set = dat->set;
    const size_t n = (size_t)set->size;
    if (n && dat->data && dat->data_d) {
        if (set->is_particle) {
            void* tmp = dev_alloc(n * dat->size);
            check(oppx_soa_to_aos(dat->data_d, tmp, (int64_t)n, dat->dim, set->set_capacity, elem_bytes(dat), nullptr),
                  "soa_to_aos");
            check(oppx_memcpy_dtoh(dat->data, tmp, n * dat->size, nullptr), "download");
            dev_free(tmp);
        } else {
            check(oppx_memcpy_dtoh(dat->data, dat->data_d, n * dat->size, nullptr), "download");
        }
    }
    dat->dirty_hd = Dirty::NotDirty;
}

void opp_download_particle_set(opp_set set, bool force) {
    for (opp_dat d : *set->particle_dats)
        if (force || d->dirty_hd == Dirty::Host) opp_download_dat(d);
}

void opp_upload_particle_set(opp_set set, bool realloc) {
    for (opp_dat d : *set->particle_dats) {
        if (realloc) create_device_arrays(d);
        opp_upload_dat(d);
    }
}

// ---------------------------------------------------------------------------------------------
// particle count growth (opp_lib/src/core/opp_increase_part_count_core.cpp:35-110): capacity grows to
// size + insert * alloc_mult, and never below the new size (an opp_allocation_multiple below 1 would
// otherwise leave the new particles past the allocation); device arrays grow by device-to-device copies
// of the live range; new mesh-relation slots hold MAX_CELL_INDEX.
void opp_increase_particle_count(opp_set set, int insert_count) {
    if (!set->is_particle) opp_abort("opp_increase_particle_count: not a particle set");
    if (insert_count <= 0) return;
    const int new_size = set->size + insert_count;
    if (set->set_capacity < new_size) {
        size_t new_cap = (size_t)set->size + (size_t)((size_t)insert_count * OPP_part_alloc_mult);
        if (new_cap < (size_t)new_size) new_cap = (size_t)new_size;
        if (new_cap > (size_t)INT_MAX - 1) opp_abort("opp_increase_particle_count: capacity beyond int range");
        for (opp_dat d : *set->particle_dats) {
            char* h = (char*)realloc(d->data, new_cap * d->size);
            if (!h) opp_abort("opp_increase_particle_count: host allocation failed");
            d->data = h;
            if (d->is_cell_index)
                for (size_t i = (size_t)set->size; i < new_cap; i++) ((int*)d->data)[i] = MAX_CELL_INDEX;
            char* nd = (char*)dev_alloc(new_cap * d->size);
            const int eb = elem_bytes(d);
            for (int c = 0; c < d->dim && set->size > 0; c++)
                check(oppx_memcpy_dtod(nd + (size_t)c * new_cap * eb, d->data_d + (size_t)c * set->set_capacity * eb,
                                       (size_t)set->size * eb, nullptr),
                      "grow copy");
            dev_free(d->data_d);
            dev_free(d->data_swap_d);
            d->data_d = nd;
            d->data_swap_d = (char*)dev_alloc(new_cap * d->size);
        }
        set->set_capacity = (int)new_cap;
        ensure_capacity_scratch(set);
    }
    // new mesh-relation slots on the device: MAX_CELL_INDEX
    if (set->mesh_relation_dat) {
        opp_dat r = set->mesh_relation_dat;
        const int32_t pat = MAX_CELL_INDEX;
        check(oppx_fill_pattern((char*)r->data_d + (size_t)set->size * sizeof(int), insert_count, 1, 4, &pat, 0, nullptr),
              "mesh relation fill");
        for (int i = set->size; i < new_size; i++) ((int*)r->data)[i] = MAX_CELL_INDEX;
    }
    set->size = new_size;
    set->diff = insert_count;
}

void opp_inc_part_count_with_distribution(opp_set set, int insert_count, opp_dat part_dist, bool calc_new) {
    // the new particles' relation holds their face of part_dist's set, filled as the seq core does
    // (opp_lib/src/core/opp_increase_part_count_core.cpp:122-146: j advances by one whenever i reaches
    // dist[j]); only the new slice travels to the device
    (void)calc_new;
    opp_increase_particle_count(set, insert_count);
    if (part_dist->dirty_hd == Dirty::Host) opp_download_dat(part_dist);
    const int* dist = (const int*)part_dist->data;
    const int n_faces = part_dist->set->size;
    const int n = set->diff, start = set->size - set->diff;
    if (n <= 0) return;
    std::vector<int> rel((size_t)n);
    int j = 0;
    for (int i = 0; i < n; i++) {
        if (j < n_faces && i >= dist[j]) j++;
        rel[(size_t)i] = j;
    }
    if (rel.back() >= n_faces)
        opp_abort("opp_inc_part_count_with_distribution: insert_count exceeds the distribution's total");
    opp_dat r = set->mesh_relation_dat;
    if (r->dirty_hd != Dirty::Device)  // device copy current: the slice goes up now
        check(oppx_memcpy_htod((int*)r->data_d + start, rel.data(), (size_t)n * sizeof(int), nullptr), "relation slice");
    if (r->dirty_hd != Dirty::Host)  // host copy current (or the newer one): keep it so
        memcpy((int*)r->data + start, rel.data(), (size_t)n * sizeof(int));
}

void opp_reset_num_particles_to_insert(opp_set set) { set->diff = 0; }

// ---------------------------------------------------------------------------------------------
// reset (opp_lib/include/opp_lib.h:98-102): every element takes the per-component value `val`
// (the reference hip backend zeroes regardless of val, opp_lib/src/hip/opp_hip.cpp:548-577)
void opp_reset_dat_impl(opp_dat dat, char* val, opp_reset reset) {
    (void)reset;
    const int eb = elem_bytes(dat);
    const int64_t n = dat->set->is_particle ? dat->set->set_capacity : dat->set->size;
    if (n > 0 && dat->data_d) {
        const int64_t stride = dat->set->is_particle ? dat->set->set_capacity : 0;
        check(oppx_fill_pattern(dat->data_d, n, dat->dim, eb, val, stride, nullptr), "reset fill");
    }
    if (dat->data)
        for (int64_t i = 0; i < dat->set->size; i++) memcpy(dat->data + i * dat->size, val, (size_t)dat->size);
    dat->dirty_hd = Dirty::NotDirty;
}

// ---------------------------------------------------------------------------------------------
// move bookkeeping entry points (loops call them internally; apps rarely do)
void opp_init_particle_move(opp_set set, int nargs, opp_arg* args) {
    (void)nargs;
    (void)args;
    OPP_iter_start = 0;
    OPP_iter_end = set->size;
    OPP_part_cells_set_size = set->cells_set ? set->cells_set->size : 0;
    set->particle_remove_count = 0;
}

bool opp_finalize_particle_move(opp_set set) {
    finalize_move_device(set);
    return false;  // single process: nothing to communicate
}

void opp_particle_sort(opp_set set) { sort_particles_device(set, false); }

// ---------------------------------------------------------------------------------------------
// text dumps (format of opp_lib/src/core/opp_lib_file_core.cpp:35-102)
static void print_array(const char* path, const void* data, int n, int dim, bool real) {
    FILE* f = fopen(path, "w");
    if (!f) opp_abort(std::string("cannot write ") + path);
    fprintf(f, "%d %d -- %d %d\n", n, dim, 0, 0);
    for (int i = 0; i < n; i++) {
        for (int d = 0; d < dim; d++) {
            if (real) {
                double v = ((const double*)data)[(size_t)i * dim + d];
                if (v == 0.0) v = 0.0;  // -0 printed as +0
                fprintf(f, " %+2.25lE", v);
            } else {
                fprintf(f, " %d", ((const int*)data)[(size_t)i * dim + d]);
            }
        }
        fprintf(f, "\n");
    }
    fclose(f);
}

void opp_print_dat_to_txtfile(opp_dat dat, const char* prefix, const char* suffix) {
    if (dat->dirty_hd == Dirty::Host) opp_download_dat(dat);
    const std::string p = std::string(prefix) + "_" + suffix;
    print_array(p.c_str(), dat->data, dat->set->size, dat->dim, strcmp(dat->type, "double") == 0);
}

void opp_print_map_to_txtfile(opp_map map, const char* prefix, const char* suffix) {
    const std::string p = std::string(prefix) + "_" + suffix;
    print_array(p.c_str(), opp_get_data(map), map->from->size, map->dim, false);
}

void opp_dump_dat(opp_dat dat) {
    if (dat->dirty_hd == Dirty::Host) opp_download_dat(dat);
    print_array("/dev/stdout", dat->data, dat->set->size, dat->dim, strcmp(dat->type, "double") == 0);
}

opp_dat opp_fetch_data(opp_dat dat) {
    if (dat->set->is_particle) opp_abort("opp_fetch_data: cannot rearrange particle dats");
    if (dat->dirty_hd == Dirty::Host) opp_download_dat(dat);
    return dat;
}
