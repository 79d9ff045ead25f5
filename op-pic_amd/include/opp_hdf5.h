// HDF5 declarations of the opp_lib API (opp_lib/include/opp_hdf5.h:39-45): sets, maps, dats and
// constants read from an OP-PIC HDF5 file, here without the HDF5 library (the reader of
// op-pic_amd/csrc/host/h5_file.cpp: superblock 0/1, symbol-table groups, contiguous or compact
// datasets of 4- or 8-byte integers and floats, no filters). One process reads the whole of every
// dataset, as the reference's MPI reader does on one rank. Writing HDF5 (opp_dump_to_hdf5,
// opp_fetch_data_hdf5_file*) is not provided.
#pragma once

#include "opp_lib.h"

// size of the set: the first element of the integer dataset `name` (opp_mpi_hdf5.cpp:126-140)
opp_set opp_decl_set_hdf5(char const* file, char const* name);
opp_set opp_decl_particle_set_hdf5(char const* file, char const* name, opp_set cells_set);

// a (from->size x dim) integer dataset; NULL, with a message, when the file has no such dataset
// (opp_mpi_hdf5.cpp:224-362)
opp_map opp_decl_map_hdf5(opp_set from, opp_set to, int dim, char const* file, char const* name);

// the first set->size rows of a (rows x dim) dataset, converted to the dat's type; a dataset the file
// lacks gives a zero-filled dat (and a message for a mesh set) (opp_mpi_hdf5.cpp:367-517)
opp_dat opp_decl_dat_hdf5(opp_set set, int dim, opp_data_type dtype, char const* file, char const* name);

// `dim` values of type `type` ("int", "long", "long long", "float", "double", and the reference's
// Fortran spellings) from dataset `name` into const_data (opp_mpi_hdf5.cpp:519-632)
void opp_get_const_hdf5(char const* name, int dim, char const* type, char* const_data, char const* file_name);
