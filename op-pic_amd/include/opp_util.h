// Helpers of OP-PIC's opp_util.h that application and generated code call directly
// (opp_lib/include/opp_util.h:48-240): printf-style formatting of log fields, the samplers the
// particle initialisers use, and small index utilities. Included by opp_lib.h, as the reference's
// opp_lib.h reaches its opp_util.h through opp_lib_core.h.
//
// The samplers draw their random numbers in the same order as the reference's (dimension-major,
// particle-minor; per cell for the cell-uniform sampler), so an application seeded the same way
// places the same particles.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#ifndef ROUND_UP
#define ROUND_UP(bytes) (((bytes) + 15) & ~15)
#endif
#ifndef ROUND_UP_64
#define ROUND_UP_64(bytes) (((bytes) + 63) & ~63)
#endif

// One value formatted with a printf format (opp_lib/include/opp_util.h:88-94). The result is cut
// to 31 characters, as the reference's 32-byte buffer cuts it, so application logs match.
template <typename T>
std::string str(T value, std::string format) {
    char out[32];
    snprintf(out, sizeof(out), format.c_str(), value);
    return std::string(out);
}

// ndim arrays of N normal samples N(mu, sigma), filled one dimension after the other
// (opp_lib/include/opp_util.h:107-122).
template <typename RNG>
inline std::vector<std::vector<double>> get_normal_distribution(const int N, const int ndim, const double mu,
                                                                const double sigma, RNG& rng) {
    std::normal_distribution<> normal{mu, sigma};
    std::vector<std::vector<double>> out((size_t)ndim, std::vector<double>((size_t)(N > 0 ? N : 0)));
    for (auto& axis : out)
        for (double& v : axis) v = normal(rng);
    return out;
}

// ndim arrays of N samples uniform in [0, extents[d]), one dimension after the other
// (opp_lib/include/opp_util.h:125-142).
template <typename RNG>
inline std::vector<std::vector<double>> get_uniform_within_extents(const int N, const int ndim, const double* extents,
                                                                   RNG& rng) {
    std::uniform_real_distribution<double> unit(0.0, 1.0);
    std::vector<std::vector<double>> out((size_t)ndim, std::vector<double>((size_t)(N > 0 ? N : 0)));
    for (int d = 0; d < ndim; d++)
        for (double& v : out[(size_t)d]) v = extents[d] * unit(rng);
    return out;
}

// npart_per_cell particles per cell, uniform inside each cell's box [ll, ll + extents): cells are
// visited in index order and each draws its particles' offsets with get_uniform_within_extents;
// `cells` receives the owning cell of every particle (opp_lib/include/opp_util.h:145-178). The
// generator is taken by value, as in the reference, so the caller's copy does not advance.
template <typename T>
inline void uniform_within_cartesian_cells(int ndim, const double* extents, const T* cell_pos_ll,
                                           const int64_t cells_set_size, const int64_t npart_per_cell,
                                           std::vector<std::vector<double>>& positions, std::vector<int>& cells,
                                           std::mt19937 rng) {
    const int64_t total = npart_per_cell * cells_set_size;
    cells.assign((size_t)(total > 0 ? total : 0), 0);
    positions.assign((size_t)ndim, std::vector<double>((size_t)(total > 0 ? total : 0)));
    for (int64_t c = 0; c < cells_set_size; c++) {
        const auto offs = get_uniform_within_extents((int)npart_per_cell, ndim, extents, rng);
        for (int64_t k = 0; k < npart_per_cell; k++) {
            const int64_t p = c * npart_per_cell + k;
            cells[(size_t)p] = (int)c;
            for (int d = 0; d < ndim; d++)
                positions[(size_t)d][(size_t)p] = cell_pos_ll[c * ndim + d] + offs[(size_t)d][(size_t)k];
        }
    }
}

// Indices that order `array` by descending value (opp_lib/include/opp_util.h:181-192).
template <typename T>
inline std::vector<T> reverse_argsort(const std::vector<T>& array) {
    std::vector<T> idx(array.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&array](int a, int b) { return array[a] > array[b]; });
    return idx;
}

// Block decomposition of N_work_items over N_compute_units: unit `work_unit` gets [*rstart, *rend),
// the first (N_work_items mod N_compute_units) units one item more (opp_lib/include/opp_util.h:195-210).
template <typename T>
inline void get_decomp_1d(const T N_compute_units, const T N_work_items, const T work_unit, T* rstart, T* rend) {
    const T base = N_work_items / N_compute_units, extra = N_work_items % N_compute_units;
    const T longer = work_unit < extra ? work_unit : extra;
    *rstart = longer * (base + 1) + (work_unit > extra ? (work_unit - extra) * base : 0);
    *rend = *rstart + (work_unit < extra ? base + 1 : base);
}

// Indices that order `keys` ascending (opp_lib/include/opp_util.h:213-222). Stable here: equal keys
// keep their index order (the reference's std::sort leaves their order unspecified).
template <typename T>
std::vector<T> sort_iota_by_key(const T* keys, size_t size) {
    std::vector<T> idx(size);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [keys](T a, T b) { return keys[a] < keys[b]; });
    return idx;
}
