/* Oracle: CPU restatement of OP-PIC's particle hot path, used ONLY as the checker of the
 * GPU path (tests, smoke check, the bench's CPU-baseline leg). Never linked by the product.
 *
 * Arrays use the reference host layout (AoS: component d of element i at [i*dim + d]).
 * Parity status of every function is stated next to it; "pinned" means checked against the
 * reference build's outputs (tests/golden) or against a closed form in tests/. */
#ifndef OPPX_ORACLE_H
#define OPPX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* calculate_new_pos_vel over particles [0,n) (pinned: golden fixtures). */
void orc_fempic_push(int64_t n, const double* c_ef, const int* p2c, double* pos, double* vel,
                     double charge, double mass, double dt, int threads);

/* Direct-hop relocation + multi-hop walk over [0,n) (pinned: golden fixtures).
 * overlay may be NULL (no direct hop). Returns the number of removed particles; hops (nullable)
 * receives the elemental-kernel evaluations per particle. */
int64_t orc_fempic_move(int64_t n, const int* c2c, const double* c_det, const double* c_vol,
                        const int* overlay, const int64_t dims[3], const double bmin[3], double inv_gs,
                        const double* pos, double* lc, int* p2c, int* hops, int threads);

/* Deposit lc into node_charge through c2n[p2c] in particle order (threads == 1, pinned: golden
 * fixtures) or with per-thread node arrays reduced in thread order (threads > 1, the reference
 * omp scheme; pinned to 1e-12 relative). node_charge is accumulated, not cleared. */
void orc_fempic_deposit(int64_t n, int64_t n_nodes, const int* c2n, const int* p2c, const double* lc,
                        double* node_charge, int threads);

/* Hole-fill compaction over [0,n) with `removed` holes marked p2c == MAX_CELL_INDEX
 * (pinned: golden fixtures + closed form). dats/dims/elem sizes describe every particle dat
 * (AoS), p2c included. Returns the new size. */
int64_t orc_holefill(int64_t n, int64_t removed, const int* p2c_marks, void** dats, const int* dims,
                     const int* elem_bytes, int n_dats);

/* Stable sort of the particle records by p2c (ties keep index order). Permutation written to
 * perm (nullable). (Sorted p2c pinned by invariant; tie order is this build's contract.) */
void orc_sort_by_cell(int64_t n, const int* p2c, void** dats, const int* dims, const int* elem_bytes,
                      int n_dats, int64_t* perm);

/* One full FEM-PIC hot-path step: push, move (+direct hop), hole-fill, deposit (node_charge
 * cleared first). Returns the new particle count; stats (nullable, 4 entries) =
 * {removed, total_hops, max_hops, 0}. */
int64_t orc_fempic_step(int64_t n, int64_t n_nodes, const int* c2n, const int* c2c, const double* c_det,
                        const double* c_vol, const double* c_ef, const int* overlay, const int64_t dims[3],
                        const double bmin[3], double inv_gs, double* pos, double* vel, double* lc, int* p2c,
                        double* node_charge, double charge, double mass, double dt, int threads,
                        int64_t* stats);

/* Brute-force direct-hop overlay of the reference (every box scans all cells in index
 * order; opp_lib/src/seq/opp_direct_hop_seq.cpp:134-265). Parity pinned: golden overlay. */
int orc_dh_overlay_bruteforce(int64_t n_cells, const int* c2c, const double* c_det, const double* c_vol,
                              const double bmin[3], const double bmax[3], double gs, int* overlay, int threads);

/* NESO-advection (apps/app_neso_advection/kernels.h:41-113): periodic drift then the
 * axis-wise walk. Returns total hops. (pinned: golden fixtures from the reference seq build) */
int64_t orc_neso_step(int64_t n, int nx, int ny, double cell_width, double dt, double ext_x, double ext_y,
                      const int* c2c, const double* c_pos_ll, double* pos, const double* vel, int* mdir,
                      int* p2c, int* hops, int threads);

/* opp_inc_part_count_with_distribution's relation fill (seq core): rel[start + i] = inlet face of
 * new particle i of [0, count) from the cumulative counts dist (pinned: reference loop fixture). */
void orc_inc_part_distribution(int64_t start, int64_t count, const int* dist, int n_faces, int* rel);

/* inject_ions over particles [first, first + count) (pinned: reference loop fixture). */
void orc_fempic_inject(int64_t first, int64_t count, const int* if2c, const double* c_ef, const double* if_u,
                       const double* if_v, const double* if_norm, const double* if_npos, const double* rnd,
                       double* pos, double* vel, int* p2c, double ion_velocity, double charge, double mass,
                       double dt);

/* compute_node_charge_density (pinned: reference loop fixture). */
void orc_node_charge_density(int64_t n, double* q, const double* vol, double spwt);

/* compute_electric_field, c_ef accumulated (pinned: reference loop fixture). */
void orc_electric_field(int64_t n_cells, double* ef, const double* sd, const int* c2n, const double* pot);

/* init_boundary_pot (pinned: reference loop fixture). */
void orc_boundary_potential(int64_t n, const int* type, double wall, double* bnd);

/* OPP_MAX fold in index order (pinned: reference loop fixture). */
double orc_max_fold(int64_t n, const double* x, int abs_values, double init);

/* CabanaPIC move_deposit loop over particles [0, n) in order (apps/app_cabanapic/kernels.h:109-353;
 * seq loop apps_codegen/app_cabanapic_cg/seq/move_deposit_kernel_loop.hpp): Boris push on the first
 * evaluation, then streak by streak with the current of each streak added to acc (12 per cell) and
 * a hop through the six face neighbours c2c6 (x-, y-, z-, x+, y+, z+). Returns total evaluations
 * (pinned: reference fixtures). */
int64_t orc_cabana_move_deposit(int64_t n, const int* c2c6, const double* interp, double* acc, double* pos,
                                double* vel, double* streak, const double* weight, int* p2c, double qsp,
                                double qdt_2mc, const double* cdt_d, int* hops);

#ifdef __cplusplus
}
#endif

#endif
