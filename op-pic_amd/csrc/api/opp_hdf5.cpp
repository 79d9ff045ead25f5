// opp_decl_*_hdf5 of the opp_lib API over the library's HDF5 reader (oppx_h5_dataset_info, oppx_h5_read):
// the checks and messages of opp_lib/src/hdf5/opp_mpi_hdf5.cpp:126-632 for one process (the local size
// of every set is its global size, so every read takes the whole dataset).

#include "opp_hdf5.h"

#include <sys/stat.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "oppx.h"

namespace {

void need_file(const char* file, const char* who) {
    struct stat st;
    if (!file || stat(file, &st) != 0)
        opp_abort(std::string("File ") + (file ? file : "(null)") + " does not exist .... aborting " + who);
}

// OPPX_OK with the dataset's shape; OPPX_ERR_NOTFOUND; any other failure aborts
int info(const char* file, const char* name, int32_t& type, int64_t& rows, int64_t& dim) {
    const int rc = oppx_h5_dataset_info(file, name, &type, &rows, &dim);
    if (rc != OPPX_OK && rc != OPPX_ERR_NOTFOUND)
        opp_abort(std::string("HDF5 read of '") + name + "' failed: " + oppx_last_error());
    return rc;
}

void read(const char* file, const char* name, int32_t type, int64_t rows, int64_t dim, void* out) {
    if (oppx_h5_read(file, name, type, rows, dim, out) != OPPX_OK)
        opp_abort(std::string("HDF5 read of '") + name + "' failed: " + oppx_last_error());
}

bool is_int(int32_t t) { return t == OPPX_H5_INT32 || t == OPPX_H5_INT64; }

int set_size(const char* file, const char* name) {
    need_file(file, "opp_decl_set_hdf5()");
    int32_t t = 0;
    int64_t rows = 0, dim = 0;
    if (info(file, name, t, rows, dim) == OPPX_ERR_NOTFOUND)
        opp_abort(std::string("Could not open dataset '") + name + "' in file '" + file + "'");
    if (rows < 1 || dim < 1) opp_abort(std::string("dataset '") + name + "' in file '" + file + "' is empty");
    std::vector<int32_t> v((size_t)dim);
    read(file, name, OPPX_H5_INT32, 1, dim, v.data());
    return v[0];
}

}  // namespace

opp_set opp_decl_set_hdf5(char const* file, char const* name) { return opp_decl_set(set_size(file, name), name); }

opp_set opp_decl_particle_set_hdf5(char const* file, char const* name, opp_set cells_set) {
    return opp_decl_particle_set(set_size(file, name), name, cells_set);
}

opp_map opp_decl_map_hdf5(opp_set from, opp_set to, int dim, char const* file, char const* name) {
    need_file(file, "op_decl_map_hdf5()");
    int32_t t = 0;
    int64_t rows = 0, fdim = 0;
    if (info(file, name, t, rows, fdim) == OPPX_ERR_NOTFOUND) {
        printf("opp_map with name : %s not found in file : %s \n", name, file);
        return nullptr;
    }
    if (from->size != rows)
        opp_abort("map from set size " + std::to_string(rows) + " in file " + file + " and size " +
                  std::to_string(from->size) + " do not match on rank 0");
    if (fdim != dim)
        opp_abort("map.dim " + std::to_string(fdim) + " in file " + file + " and dim " + std::to_string(dim) +
                  " do not match");
    if (!is_int(t)) opp_abort("unknown type");
    std::vector<int32_t> m((size_t)rows * (size_t)dim + 1);
    read(file, name, OPPX_H5_INT32, rows, dim, m.data());
    return opp_decl_map(from, to, dim, m.data(), name);
}

opp_dat opp_decl_dat_hdf5(opp_set set, int dim, opp_data_type dtype, char const* file, char const* name) {
    need_file(file, "op_decl_dat_hdf5()");
    int32_t t = 0;
    int64_t rows = 0, fdim = 0;
    if (info(file, name, t, rows, fdim) == OPPX_ERR_NOTFOUND) {
        if (!set->is_particle) printf("opp_dat with name : %s not found in file : %s \n", name, file);
        return opp_decl_dat(set, dim, dtype, nullptr, name);
    }
    if (fdim != dim)
        opp_abort("dat.dim " + std::to_string(fdim) + " in file " + file + " and dim " + std::to_string(dim) +
                  " do not match");
    const bool real = dtype == DT_REAL;
    const size_t n = (size_t)set->size * (size_t)dim;
    std::vector<char> buf((n + 1) * (real ? sizeof(double) : sizeof(int32_t)));
    read(file, name, real ? OPPX_H5_FLOAT64 : OPPX_H5_INT32, set->size, dim, buf.data());
    return opp_decl_dat(set, dim, dtype, buf.data(), name);
}

void opp_get_const_hdf5(char const* name, int dim, char const* type, char* const_data, char const* file_name) {
    need_file(file_name, "op_get_const_hdf5()");
    int32_t t = 0;
    int64_t rows = 0, fdim = 0;
    if (info(file_name, name, t, rows, fdim) == OPPX_ERR_NOTFOUND) {
        printf("dataset with '%s' not found in file '%s' \n", name, file_name);
        return;
    }
    if (fdim != 1) opp_abort(std::string("constant '") + name + "' in file " + file_name + " is not one-dimensional");
    if (rows != dim)
        opp_abort("dim of constant " + std::to_string(rows) + " in file " + file_name + " and requested dim " +
                  std::to_string(dim) + " do not match");
    const std::string ty(type ? type : "");
    int32_t as = 0;
    if (ty == "int" || ty == "int(4)" || ty == "integer" || ty == "integer(4)") as = OPPX_H5_INT32;
    else if (ty == "long" || ty == "long long") as = OPPX_H5_INT64;
    else if (ty == "float" || ty == "real(4)" || ty == "real") as = OPPX_H5_FLOAT32;
    else if (ty == "double" || ty == "double precision" || ty == "real(8)") as = OPPX_H5_FLOAT64;
    else opp_abort(std::string("Unknown type in file ") + file_name + " for constant " + name);
    read(file_name, name, as, rows, fdim, const_data);
}
