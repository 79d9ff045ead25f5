// Host-side builders for the NESO-advection hot path: the 2-D periodic quad mesh of
// apps/app_neso_advection/advec_misc.h:115-143 (cell i = x + nx*y, lower-left corner x*w, y*w,
// neighbours [x-1, x+1, y-1, y+1] wrapped periodically as VOXEL_MAP of advec_defs.h) and a
// seeded particle population (uniform positions, uniform velocities in [-vmax, vmax]^2).
// Compiled with -ffp-contract=off like the other builders.

#include <cmath>
#include <cstdint>

#include "oppx.h"

namespace {
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
inline double u01(uint64_t seed, uint64_t ctr) {
    const uint64_t r = mix64(mix64(seed + 0x9E3779B97F4A7C15ULL) ^ (ctr * 0xD1B54A32D192ED03ULL + 0x632BE59BD9B4E019ULL));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}
inline int wrap(int v, int n) { return v < 0 ? n - 1 : (v >= n ? 0 : v); }
}  // namespace

extern "C" {

int oppx_neso_mesh(int nx, int ny, double cell_width, double* c_pos_ll, int* c2c) {
    if (nx <= 0 || ny <= 0 || !(cell_width > 0) || !c_pos_ll || !c2c) return OPPX_ERR_ARG;
    if ((int64_t)nx * ny >= OPPX_MAX_CELL_INDEX) return OPPX_ERR_ARG;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
            const int64_t i = x + (int64_t)nx * y;
            c_pos_ll[2 * i + 0] = x * cell_width;
            c_pos_ll[2 * i + 1] = y * cell_width;
            c2c[4 * i + 0] = wrap(x - 1, nx) + nx * y;  // xd_y
            c2c[4 * i + 1] = wrap(x + 1, nx) + nx * y;  // xu_y
            c2c[4 * i + 2] = x + nx * wrap(y - 1, ny);  // x_yd
            c2c[4 * i + 3] = x + nx * wrap(y + 1, ny);  // x_yu
        }
    return OPPX_OK;
}

int oppx_neso_particles(int64_t n_part, uint64_t seed, int nx, int ny, double cell_width, double vmax, double* pos,
                        double* vel, int* p2c) {
    if (n_part < 0 || nx <= 0 || ny <= 0 || !(cell_width > 0) || !pos || !vel || !p2c) return OPPX_ERR_ARG;
    const double ext[2] = {nx * cell_width, ny * cell_width};
    const int nd[2] = {nx, ny};
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n_part; p++) {
        int c[2];
        for (int d = 0; d < 2; d++) {
            const double x = u01(seed, (uint64_t)p * 4 + d) * ext[d];
            pos[2 * p + d] = x;
            int ci = (int)(x / cell_width);
            c[d] = ci >= nd[d] ? nd[d] - 1 : ci;
            vel[2 * p + d] = (2.0 * u01(seed, (uint64_t)p * 4 + 2 + d) - 1.0) * vmax;
        }
        p2c[p] = c[0] + nx * c[1];
    }
    return OPPX_OK;
}

}  // extern "C"
