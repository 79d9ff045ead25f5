/* oppx: MI355X-native (gfx950) particle hot path behind OP-PIC's opp_lib API.
 *
 * This is the C-ABI layer: extern "C", plain pointers and sizes. Device pointers are
 * HIP device allocations; `stream` is a hipStream_t passed as void* (NULL = default
 * stream). Every entry point returns an int status (OPPX_OK or an OPPX_ERR_* code);
 * oppx_last_error() gives the message of the last failure on the calling thread.
 *
 * Data layout on the device (chosen for HBM3E streaming + L2-friendly gathers):
 *   particle dats : SoA, component d of particle i at base[d * stride + i], stride = set capacity
 *   mesh dats     : AoS, cell c's record contiguous (c_det = 16 doubles = one 128 B line)
 *   tet aux       : per cell {c2 = (1/6)/volume, c2c[4]} packed in 32 B (oppx_tet_aux)
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * OP-PIC source tree).
 */
#ifndef OPPX_H
#define OPPX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OPPX_OK 0
#define OPPX_ERR_ARG 1    /* invalid argument (null pointer, size out of range, ...) */
#define OPPX_ERR_HIP 2    /* a HIP runtime call failed */
#define OPPX_ERR_NOMEM 3  /* device or host allocation failed, or admission budget exceeded */
#define OPPX_ERR_STATE 4  /* call out of order (library not initialised, handle busy, ...) */
#define OPPX_ERR_NOTFOUND 5  /* a named item (an HDF5 dataset) is not in the file */

/* MAX_CELL_INDEX of opp_lib/include/opp_defs.h:150: the p2c value of a removed particle. */
#define OPPX_MAX_CELL_INDEX 2147483647

/* Fill policy of opp_lib/include/opp_defs.h:252-258 (opp_fill_type). */
#define OPPX_FILL_HOLEFILL_ALL 0
#define OPPX_FILL_SORT_ALL 1
#define OPPX_FILL_SHUFFLE_ALL 2
#define OPPX_FILL_SORT_PERIODIC 3
#define OPPX_FILL_SHUFFLE_PERIODIC 4

/* ---------------------------------------------------------------------------------------
 * Device-side records
 * ------------------------------------------------------------------------------------- */

/* Per-tet data the multi-hop walk reads besides c_det: the mover's coefficient
 * (1/6)/volume (apps/app_fempic/kernels.h:110) and the face neighbours c2c (negative = boundary;
 * an entry at or past n_cells is reported as a fault). */
typedef struct oppx_tet_aux {
    double c2;
    int32_t c2c[4];
    int32_t pad[2];
} oppx_tet_aux;

/* Tet mesh as the hot path sees it (all device pointers). */
typedef struct oppx_tet_mesh {
    int64_t n_cells;
    int64_t n_nodes;
    const double* det;         /* [16*n_cells] alpha,beta,gamma,delta per vertex (c_det) */
    const oppx_tet_aux* aux;   /* [n_cells] */
    const int32_t* c2n;        /* [4*n_cells] cell -> node map */
    const double* ef;          /* [3*n_cells] cell electric field (c_ef), AoS */
} oppx_tet_mesh;

/* Direct-hop overlay (opp_lib/include/opp_direct_hop_core.h:264-454, CellMapper). */
typedef struct oppx_dh_grid {
    const int32_t* cell;       /* device [dims[0]*dims[1]*dims[2]] owning cell or MAX_CELL_INDEX */
    int64_t dims[3];
    double min[3];             /* global minimum corner of the bounding box */
    double inv_gs;             /* 1/grid_spacing */
} oppx_dh_grid;

/* Particle state of one FEM-PIC instance (device pointers, SoA with `stride`). */
typedef struct oppx_fempic_parts {
    int64_t stride;            /* set capacity */
    double* pos;               /* [3*stride] */
    double* vel;               /* [3*stride] */
    double* lc;                /* [4*stride] */
    int32_t* p2c;              /* [stride]   */
} oppx_fempic_parts;

/* Device counters a move pass maintains (one per particle set). */
typedef struct oppx_move_counters {
    int32_t size;              /* live particles before the move (input), after finalize (output) */
    int32_t removed;           /* particles removed by this move */
    int32_t holes_low;         /* holes below the new size = records moved by hole-fill */
    int32_t max_hops;          /* longest walk of the pass */
    int64_t total_hops;        /* kernel evaluations of the pass */
    int32_t faults;            /* particles whose walk hit the hop cap or an invalid cell index */
    int32_t pad;
    int32_t queue[8];          /* per XCD: next block tile of the persistent walk, or worklist length of
                                  the two-pass move (reset per launch) */
} oppx_move_counters;

/* Walks longer than this are stopped and counted in oppx_move_counters.faults (the
 * reference would loop forever on a cycling walk). */
#define OPPX_MAX_HOPS 4096

/* Launch shape of the wavefront-cooperative walk and the sort: each wave owns `tile` consecutive
 * particles (multiple of 64, <= 512) and refills its free lanes once `refill_min` of them are idle;
 * `split` = 1 runs push + overlay relocation as a separate coalesced pass before the walk;
 * `persistent` = 1 launches a resident grid whose blocks take block tiles from per-XCD queues;
 * `sort_items` (4, 8, 16 or 32) is the elements per thread of the radix sort passes (a block sorts
 * 256 * sort_items keys); `inline_hops` >= 1 selects the two-pass move (walk evaluations done in the
 * first, one-lane-per-particle pass before a particle goes to the worklist walk), 0 the
 * cooperative tile walk. Process-wide; chosen by the autotuner (op-pic_amd/autotune.py) or
 * oppx_set_tuning. Results do not depend on it. */
typedef struct oppx_tuning {
    int32_t tile;
    int32_t refill_min;
    int32_t split;
    int32_t persistent;
    int32_t sort_items;
    int32_t inline_hops;
    int32_t first_block;       /* threads per block of the first-hop pass: 256 or 1024 */
    int32_t lanes;             /* lanes per particle in the first-hop pass: 1, or 4 (a quad walks one
                                  particle, lane r owning tet vertex r) */
} oppx_tuning;

/* One instance of a batched FEM-PIC ensemble. The instances' meshes are concatenated into one
 * mesh (cells of instance i are [cell_base, cell_base + n_cells), its nodes
 * [node_base, node_base + n_nodes); c2c, c2n and p2c hold concatenated indices) and their particles
 * into one particle set. dh_cell (device, nullable) is the instance's overlay holding concatenated
 * cell indices or MAX_CELL_INDEX. 112 B. */
typedef struct oppx_batch_inst {
    int64_t cell_base;
    int64_t n_cells;
    int64_t node_base;
    int64_t n_nodes;
    const int32_t* dh_cell;
    int64_t dims[3];
    double min[3];
    double inv_gs;
    double k;                  /* charge / mass * dt */
    double dt;
} oppx_batch_inst;

/* One particle dat for compaction/permutation: SoA base, components, element bytes (4 or 8). */
typedef struct oppx_dat_desc {
    void* base;
    int32_t dim;
    int32_t elem_bytes;
} oppx_dat_desc;

/* ---------------------------------------------------------------------------------------
 * Library / device
 * ------------------------------------------------------------------------------------- */

/* opp_init's device part (opp_lib/src/hip/opp_hip.cpp:221-271, opp_device_init). */
int oppx_init(int device);
const char* oppx_last_error(void);
int oppx_device_malloc(void** ptr, size_t bytes);
int oppx_device_free(void* ptr);
int oppx_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int oppx_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int oppx_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int oppx_memset_zero(void* dst, size_t bytes, void* stream);
int oppx_stream_sync(void* stream);

/* Fill n elements of dim components (elem_bytes 4 or 8, dim*elem_bytes <= 256) with the per-component
 * values in `pattern` (host memory): AoS when stride == 0, else SoA with `stride`. A kernel, so it can be
 * graph-captured. Serves opp_reset_dat (opp_lib/include/opp_lib.h:98-102). */
int oppx_fill_pattern(void* dst, int64_t n, int dim, int elem_bytes, const void* pattern, int64_t stride, void* stream);

/* Adds `offset` to the index fields [first, first+count) of n_records records of rec_ints int32
 * each (device), leaving negative entries and MAX_CELL_INDEX alone: rebasing the maps of an instance
 * placed into a concatenated mesh (no reference counterpart). */
int oppx_offset_indices(int32_t* base, int64_t n_records, int rec_ints, int first, int count, int32_t offset,
                        void* stream);

/* AoS (host layout, doc/api.rst:115-118) <-> SoA (device layout) transposes on the device,
 * replacing the host transposes of opp_lib/src/hip/opp_hip_utils.cpp:35-153. */
int oppx_aos_to_soa(const void* aos, void* soa, int64_t n, int dim, int64_t stride, int elem_bytes, void* stream);
int oppx_soa_to_aos(const void* soa, void* aos, int64_t n, int dim, int64_t stride, int elem_bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * Host builders (CPU only; no device needed). Inputs of the hot path.
 * ------------------------------------------------------------------------------------- */

/* Box mesh of nx*ny*nz cubes x 6 tets; geometry per apps/app_fempic/minifempic_funcs.h:284-443.
 * order 0 numbers cubes and nodes x fastest; order 1 in Morton order (spatial neighbours close in
 * index, so cell records of neighbouring tets share cache and HBM pages). Cube r owns tets 6r..6r+5. */
int oppx_box_mesh_sizes(int nx, int ny, int nz, int64_t* n_nodes, int64_t* n_cells);
int oppx_box_mesh_build(int nx, int ny, int nz, double lx, double ly, double lz, int order,
                        double* node_pos, int* c2n, int* c2c, double* c_det, double* c_vol);
/* c_ef = -grad of the P1 interpolant of sin(pi x) sin(pi y) sin(pi z). */
int oppx_box_field_sines(int64_t n_nodes, int64_t n_cells, const double* node_pos, const int* c2n,
                         const double* c_det, const double* c_vol, double* c_ef);
/* Seeded uniform positions, N(0,sigma^2) velocities, containing tet (AoS host arrays). */
int oppx_box_particles_uniform(int64_t n_part, uint64_t seed, double sigma, int nx, int ny, int nz,
                               double lx, double ly, double lz, int order, double* pos, double* vel, int* p2c);
/* node -> (cell, vertex) incidence: node_start [n_nodes+1], node_ref [4*n_cells] = c*4+k ascending per node. */
int oppx_node_cell_csr(int64_t n_cells, int64_t n_nodes, const int* c2n, int* node_start, int* node_ref);

/* FEM-PIC text mesh files (mesh.dat: node count, element count, "index x y z" in mm, "index 304 n1 n2 n3 n4"
 * tets with 1-based nodes; inlet.dat: inlet nodes, then "index 203 n1 n2 n3" triangles; wall.dat: fixed nodes).
 * oppx_box12_write writes the 12-tet box mesh of nx x ny x nz cubes of edge cube_length mm into dir in that
 * format (node and tet numbering, face rows and node lists of scripts/batch/fempic/box_mesh_gen.py:19-178):
 * corners numbered x fastest, then one centre node per cube; tets 12q .. 12q+11 belong to cube q.
 * oppx_mesh_file_load reads the files as apps/app_fempic/minifempic_funcs.h:230-588 (LoadVolumeMesh,
 * LoadSurfaceMesh) does: positions in m, volume sign flip, c_det [16 n_cells] (alpha, beta, gamma, delta per
 * vertex, packed as apps_codegen/app_fempic_cg/fempic_misc_mesh_loader.h:119-127), c_vol, c2c (-1 on the
 * boundary), node volumes (centroid weights x tet volume, scattered in tet order), node types (0 normal,
 * 1 open, 2 inlet, 3 fixed) and the inlet faces (owning tet, nodes, u, v, unit normal, area). inlet_path and
 * wall_path may be NULL. Host arrays sized by oppx_mesh_file_sizes. */
int oppx_box12_write(int nx, int ny, int nz, double cube_length, const char* dir);
/* n_part particles uniform in [0,lx]x[0,ly]x[0,lz] (the streams of oppx_box_particles_uniform) on a 12-tet box
 * mesh read back from oppx_box12_write's files, each tagged with the first of its cube's 12 tets that the
 * mover's inside test accepts (OPPX_ERR_STATE if one is in none). */
int oppx_box12_particles(int64_t n_part, uint64_t seed, double sigma, int nx, int ny, int nz, double lx, double ly,
                         double lz, const double* c_det, const double* c_vol, double* pos, double* vel, int* p2c);
int oppx_mesh_file_sizes(const char* mesh_path, const char* inlet_path, int64_t* n_nodes, int64_t* n_cells,
                         int64_t* n_ifaces);
int oppx_mesh_file_load(const char* mesh_path, const char* inlet_path, const char* wall_path, double* node_pos,
                        int32_t* node_type, double* node_vol, int32_t* c2n, int32_t* c2c, double* c_det,
                        double* c_vol, int32_t* if2c, int32_t* if2n, double* if_u, double* if_v, double* if_normal,
                        double* if_area);

/* HDF5 files as OP-PIC's opp_decl_set_hdf5 / opp_decl_map_hdf5 / opp_decl_dat_hdf5 read them
 * (opp_lib/src/hdf5/opp_mpi_hdf5.cpp:126-517), without an HDF5 library: superblock version 0 or 1,
 * symbol-table groups, version-1 object headers, contiguous or compact datasets of little-endian 4- or
 * 8-byte integers or floats, no filters (anything else returns OPPX_ERR_ARG with the reason in
 * oppx_last_error). `name` is a path from the root group ("c_det", "group/name"); a missing dataset
 * returns OPPX_ERR_NOTFOUND. oppx_h5_dataset_info gives the stored element type and the extents as
 * the reference takes them (rows = first extent, dim = second extent or 1; a scalar is 1 x 1).
 * oppx_h5_read copies the first `rows` rows (dim must match) into `out` as `type`, converting like
 * HDF5's conversion to a memory type (integers to floats round, floats to integers truncate; a value
 * out of the target's range is an error). Host-side; thread-safe. */
#define OPPX_H5_INT32 1
#define OPPX_H5_INT64 2
#define OPPX_H5_FLOAT32 3
#define OPPX_H5_FLOAT64 4
int oppx_h5_dataset_info(const char* path, const char* name, int32_t* type, int64_t* rows, int64_t* dim);
int oppx_h5_read(const char* path, const char* name, int32_t type, int64_t rows, int64_t dim, void* out);
/* Overlay dims of CellMapper::init_host (opp_lib/include/opp_direct_hop_core.h:317-336). */
int oppx_dh_grid_dims(const double bmin[3], const double bmax[3], double gs, int64_t dims[3]);
/* Overlay contents of CellMapper::generateStructuredMesh (opp_lib/src/seq/opp_direct_hop_seq.cpp:134-265),
 * built with a candidate-restricted search that returns the reference's first-in-index-order owner. */
int oppx_dh_build_overlay_nodes(int64_t n_cells, const double* c_det, const double* c_vol, const int* c_gbl_id,
                                const int* c2n, const double* node_pos, const double bmin[3],
                                const double bmax[3], double gs, int* overlay);
/* The same overlay from c_det / c_volume alone (tet vertices recovered from the barycentric planes). */
int oppx_dh_build_overlay_geom(int64_t n_cells, const double* c_det, const double* c_vol, const int* c_gbl_id,
                               const double bmin[3], const double bmax[3], double gs, int* overlay);

/* Overlay files in the reference's format: the zlib stream of the int32 entries
 * (opp_lib/src/core/opp_util.cpp:442-496), named as opp_lib/src/seq/opp_direct_hop_seq.cpp:107-119 and
 * :288-298 name them (dh_c<cells>_gs<spacing>_ex<expansion x>_<y>_<z>.bin). load fails unless the file
 * holds exactly n entries. */
int oppx_dh_overlay_filename(int64_t n_cells, double grid_spacing, const double expansion[3], char* out, size_t cap);
int oppx_dh_overlay_save(const char* path, const int32_t* overlay, int64_t n);
int oppx_dh_overlay_load(const char* path, int32_t* overlay, int64_t n);

/* NESO-advection 2-D periodic quad mesh (apps/app_neso_advection/advec_misc.h:115-143):
 * c_pos_ll [2*nx*ny] lower-left corners, c2c [4*nx*ny] = {x-1, x+1, y-1, y+1} wrapped. */
int oppx_neso_mesh(int nx, int ny, double cell_width, double* c_pos_ll, int* c2c);
/* Seeded uniform positions in the periodic box, velocities uniform in [-vmax, vmax]^2, containing cell. */
int oppx_neso_particles(int64_t n_part, uint64_t seed, int nx, int ny, double cell_width, double vmax, double* pos,
                        double* vel, int* p2c);

/* ---------------------------------------------------------------------------------------
 * FEM-PIC hot path (device)
 * ------------------------------------------------------------------------------------- */

/* Build the per-tet aux records from device c_volume [n] and c2c [4n]. */
int oppx_tet_aux_build(const double* c_vol, const int32_t* c2c, int64_t n_cells, oppx_tet_aux* aux, void* stream);

/* opp_par_loop_all__calculate_new_pos_vel_kernel
 * (apps_codegen/app_fempic_cg/hip/calculate_new_pos_vel_kernel_loop.hpp:31-100; elemental
 * apps/app_fempic/kernels.h:93-103): vel += k*E[p2c]; pos += vel*dt with k = (charge/mass)*dt,
 * over the counters->size live particles (n_max bounds the grid). */
int oppx_fempic_push(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                     double k, double dt, oppx_move_counters* counters, void* stream);

/* Device workspace of the two-pass move (first-hop pass + worklist walk) for n_max particles. */
size_t oppx_fempic_work_bytes(int64_t n_max);

/* opp_particle_move__move_kernel (apps_codegen/app_fempic_cg/hip/move_kernel_loop.hpp:142-407):
 * optional direct-hop relocation (dh may be NULL; opp_lib/include/device_kernels/hip_inline_kernels.h:330-393)
 * then the multi-hop barycentric walk (apps/app_fempic/kernels.h:106-161). Reads counters->size,
 * writes counters->removed/max_hops/total_hops/faults and the removal bitmap hole_bits
 * [oppx_hole_bits_words(n_max)] used by oppx_holefill. `hops` (nullable) receives per-particle hop counts.
 * work (nullable, oppx_fempic_work_bytes(n_max) bytes) selects the two-pass move when
 * tuning.inline_hops >= 1; without it the cooperative tile walk runs. Results are the same.
 * n_max bounds the live count (grid size); the live count itself is read on the device. */
int oppx_fempic_move(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                     const oppx_dh_grid* dh, oppx_move_counters* counters, uint64_t* hole_bits,
                     void* work, size_t work_bytes, int32_t* hops, void* stream);

/* opp_par_loop_all__deposit_charge_on_nodes_kernel
 * (apps_codegen/app_fempic_cg/hip/deposit_charge_on_nodes_kernel_loop.hpp:39-95, 208-222;
 * elemental apps/app_fempic/kernels.h:164-175): node_charge[c2n[p2c][k]] += lc[k], FP64 atomics.
 * Reads the live count from counters->size. */
int oppx_fempic_deposit_atomic(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                               double* node_charge, oppx_move_counters* counters, void* stream);

/* Start of a fused step: zeroes node_charge [n_nodes] (nullable) and the per-step counters
 * (removed, holes_low, max_hops, total_hops) in one launch -- opp_reset_dat of the deposit
 * target (opp_lib/src/hip/opp_hip.cpp:548-577) plus opp_init_particle_move's counter reset. */
int oppx_fempic_step_begin(oppx_move_counters* counters, double* node_charge, int64_t n_nodes, void* stream);

/* Fused step: push + (direct hop) + walk + FP64 atomic deposit in one pass over the particles (plus
 * the worklist walk of the particles still moving after tuning.inline_hops evaluations when `work`
 * is given, as for oppx_fempic_move). Same per-particle results as the three loops above in
 * sequence; node_charge (nullable) is accumulated and the hop statistics add to counters, so a
 * step starts with oppx_fempic_step_begin. */
int oppx_fempic_step_fused(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                           const oppx_dh_grid* dh, double k, double dt, oppx_move_counters* counters,
                           uint64_t* hole_bits, void* work, size_t work_bytes, int32_t* hops, double* node_charge,
                           void* stream);

/* The fused step with a pending sort applied on the fly: particle i reads its pos/vel/p2c (and, for
 * fields the step does not produce, lc) at perm[i] of `src` and writes every field at i of `parts`
 * (separate buffers, same stride). A Sort_All / Sort_Periodic finalize can then compute only the
 * permutation (oppx_sort_by_cell with perm_out and at most the p2c dat) and leave the record gather to
 * the next step, which reads the records anyway (no reference counterpart: the reference gathers every
 * dat in the finalize). src->p2c == NULL: the sort already wrote the sorted p2c into parts->p2c, which
 * is then read in place. Needs the two-pass move (tuning.inline_hops >= 1 and its workspace). */
int oppx_fempic_step_fused_gather(int64_t n_max, const int32_t* perm, const oppx_fempic_parts* src,
                                  const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh, const oppx_dh_grid* dh,
                                  double k, double dt, oppx_move_counters* counters, uint64_t* hole_bits, void* work,
                                  size_t work_bytes, int32_t* hops, double* node_charge, void* stream);

/* The fused step with each particle's position and velocity as one 48-byte record {x, y, z, vx, vy, vz}
 * on either side (no reference counterpart: a layout for the step before a lazy sort and the step
 * after it). rec_out (nullable): particle i's new position and velocity go to rec_out[6i .. 6i+5]
 * instead of parts->pos/vel (left as they were); every particle in [0, counters->size) gets a record.
 * perm (nullable): a pending sort permutation -- particle i then reads its position and velocity at
 * src_rec[6 perm[i] ..] (one 48-byte read in place of six 8-byte SoA reads at permuted addresses), its
 * lc at src_lc[d*stride + perm[i]] and its p2c in place (parts->p2c, written sorted by the sort), and
 * writes every field at i of parts (pos/vel, or rec_out when given). perm NULL: src_rec and src_lc NULL
 * too, and the step reads parts in place. Per-particle results equal oppx_fempic_step_fused /
 * oppx_fempic_step_fused_gather. Needs the two-pass move (tuning.inline_hops >= 1 and its workspace). */
int oppx_fempic_step_fused_rec(int64_t n_max, const int32_t* perm, const double* src_rec, const double* src_lc,
                               const oppx_fempic_parts* parts, double* rec_out, const oppx_tet_mesh* mesh,
                               const oppx_dh_grid* dh, double k, double dt, oppx_move_counters* counters,
                               uint64_t* hole_bits, void* work, size_t work_bytes, int32_t* hops, double* node_charge,
                               void* stream);

/* pos[d*stride + i] = rec[6 perm[i] + d], vel[d*stride + i] = rec[6 perm[i] + 3 + d] for i < counters->size
 * (perm NULL: record i): materialises the records oppx_fempic_step_fused_rec left, e.g. before the state is
 * read back, hole-filled or sorted eagerly. A perm entry outside [0, stride) is counted in counters->faults. */
int oppx_gather_records(int64_t n_max, int64_t stride, oppx_move_counters* counters, const int32_t* perm,
                        const double* rec, double* pos, double* vel, void* stream);

/* Batched fused step over an ensemble of independent instances concatenated as described at
 * oppx_batch_inst: one launch walks every instance's particles (no reference counterpart: OP-PIC
 * runs one instance per process). cell_inst (device [mesh->n_cells]) is the instance of every cell,
 * insts (device [n_insts]) the instance table. Per-particle results are those of running each
 * instance alone through oppx_fempic_step_fused. */
int oppx_fempic_batch_step_fused(int64_t n_max, const oppx_fempic_parts* parts, const oppx_tet_mesh* mesh,
                                 const int32_t* cell_inst, const oppx_batch_inst* insts, int32_t n_insts,
                                 oppx_move_counters* counters, uint64_t* hole_bits, void* work, size_t work_bytes,
                                 int32_t* hops, double* node_charge, void* stream);

/* oppx_fempic_batch_step_fused with 48-byte position/velocity records on either side, as
 * oppx_fempic_step_fused_rec describes them (no reference counterpart): the batched step before a lazy
 * sort of the whole ensemble and the step after it. Per-particle results equal
 * oppx_fempic_batch_step_fused. Needs the two-pass move (tuning.inline_hops >= 1 and its workspace). */
int oppx_fempic_batch_step_fused_rec(int64_t n_max, const int32_t* perm, const double* src_rec, const double* src_lc,
                                     const oppx_fempic_parts* parts, double* rec_out, const oppx_tet_mesh* mesh,
                                     const int32_t* cell_inst, const oppx_batch_inst* insts, int32_t n_insts,
                                     oppx_move_counters* counters, uint64_t* hole_bits, void* work, size_t work_bytes,
                                     int32_t* hops, double* node_charge, void* stream);

/* ---------------------------------------------------------------------------------------
 * FEM-PIC injection and the loops either side of the field solve (device)
 * ------------------------------------------------------------------------------------- */

/* Inlet faces of the injection (the reference app's iface set and its dats; device, AoS). */
typedef struct oppx_inlet {
    int64_t n_faces;
    const int32_t* face_cell;  /* if2c_map [n_faces]: the tet behind the face */
    const double* u;           /* iface_u_norm [n_faces][3] */
    const double* v;           /* iface_v_norm [n_faces][3] */
    const double* normal;      /* iface_norm [n_faces][3] */
    const double* node_pos;    /* iface_n_pos [n_faces][9]; the first node is the origin */
} oppx_inlet;

/* opp_par_loop_injected__inject_ions_kernel (apps/app_fempic/kernels.h:63-90; loops
 * apps_codegen/app_fempic_cg/{seq,hip}/inject_ions_kernel_loop.hpp) over particles
 * [first, first + count), whose p2c holds their inlet face (opp_inc_part_count_with_distribution):
 * position on the face from the k-th random pair (rand[k], rand[rand_stride + k], SoA), velocity
 * ion_velocity * normal - (charge/mass) * E(cell) * dt/2, p2c = the face's cell. A face or cell
 * index out of range counts in counters->faults (nullable) and leaves the particle as it was. */
int oppx_fempic_inject(int64_t first, int64_t count, const oppx_fempic_parts* parts, const oppx_inlet* inlet,
                       int64_t n_cells, const double* c_ef, const double* rand, int64_t rand_stride,
                       double ion_velocity, double charge, double mass, double dt, oppx_move_counters* counters,
                       void* stream);

/* opp_par_loop_all__compute_node_charge_density_kernel (kernels.h:178-183):
 * node_charge[i] *= spwt / node_volume[i]. */
int oppx_fempic_node_charge_density(int64_t n_nodes, double* node_charge, const double* node_volume, double spwt,
                                    void* stream);

/* opp_par_loop_all__compute_electric_field_kernel (kernels.h:186-202): per tet and axis,
 * c_ef -= sum_k c_sd[k][axis] * node_potential[c2n[k]] (c_sd: 12 doubles per tet, AoS). A c2n
 * entry out of range counts in counters->faults (nullable) and leaves that tet's field. */
int oppx_fempic_electric_field(int64_t n_cells, double* c_ef, const double* c_sd, const int32_t* c2n,
                               const double* node_potential, int64_t n_nodes, oppx_move_counters* counters,
                               void* stream);

/* opp_par_loop_all__init_boundary_pot_kernel (kernels.h:48-59): node_bnd_pot = -1 * wall_potential
 * on FIXED nodes (type 3), 0 elsewhere. */
int oppx_fempic_boundary_potential(int64_t n_nodes, const int32_t* node_type, double wall_potential,
                                   double* node_bnd_pot, void* stream);

/* Scratch needed by oppx_max_reduce over n values (bytes). */
size_t oppx_max_scratch_bytes(int64_t n);

/* An OPP_MAX global over a dat (get_max_cef_kernel, get_final_max_values_kernel, kernels.h:205-224;
 * MAX of opp_lib/include/opp_util.h:45): *result (device) = the fold m = init; m = MAX(v, m) over
 * values[0..n) in order (|v| when abs_values), i.e. the first value attaining the maximum. */
int oppx_max_reduce(int64_t n, const double* values, int abs_values, double init, double* result, void* scratch,
                    size_t scratch_bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * CabanaPIC move with per-hop current deposit (device)
 * ------------------------------------------------------------------------------------- */

/* CabanaPIC particle state (device pointers, SoA with `stride`); positions are cell-local in [-1, 1]. */
typedef struct oppx_cabana_parts {
    int64_t stride;
    double* pos;               /* [3*stride] p_pos */
    double* vel;               /* [3*stride] p_vel */
    double* streak;            /* [3*stride] p_streak_mid */
    const double* weight;      /* [stride]   p_weight */
    int32_t* p2c;              /* [stride] */
} oppx_cabana_parts;

/* Structured-mesh data the move reads (device, AoS). */
typedef struct oppx_cabana_mesh {
    int64_t n_cells;
    const int32_t* c2c6;       /* [n_cells][6] face neighbours x-, y-, z-, x+, y+, z+ (c2c_non_ghost_map) */
    const double* interp;      /* [n_cells][18] c_interp */
} oppx_cabana_mesh;

/* opp_particle_move__move_deposit_kernel (apps/app_cabanapic/kernels.h:109-353; loops
 * apps_codegen/app_cabanapic_cg/{seq,hip}/move_deposit_kernel_loop.hpp) over counters->size
 * particles: Boris push on the first evaluation, then streak by streak with each streak's current
 * added to acc ([n_cells][12], accumulated, not cleared) and a hop to the face neighbour. cdt_d is a
 * host array of 3. The counters' hop statistics are reset first; `hops` (nullable) receives the
 * evaluations per particle; a neighbour index out of range or OPPX_MAX_HOPS counts in faults. */
int oppx_cabana_move_deposit(int64_t n_max, const oppx_cabana_parts* parts, const oppx_cabana_mesh* mesh, double* acc,
                             double qsp, double qdt_2mc, const double cdt_d[3], oppx_move_counters* counters,
                             int32_t* hops, void* stream);

/* ---------------------------------------------------------------------------------------
 * NESO-advection hot path (device)
 * ------------------------------------------------------------------------------------- */

/* Per-quad record of the 2-D walk: lower-left corner (c_pos_ll) and neighbours
 * {x-1, x+1, y-1, y+1} (c2c), 32 B. */
typedef struct oppx_quad_cell {
    double llx, lly;
    int32_t c2c[4];
} oppx_quad_cell;

/* NESO particle state (device pointers, SoA with `stride`). */
typedef struct oppx_neso_parts {
    int64_t stride;
    double* pos;               /* [2*stride] */
    const double* vel;         /* [2*stride] */
    int32_t* mdir;             /* [2*stride] move direction (OPP_RW dat p_mdir) */
    int32_t* p2c;              /* [stride] */
} oppx_neso_parts;

/* Pack c_pos_ll [2n] and c2c [4n] (device) into quad records. */
int oppx_quad_cells_build(const double* c_pos_ll, const int32_t* c2c, int64_t n_cells, oppx_quad_cell* cells,
                          void* stream);

#define OPPX_NESO_UPDATE 1
#define OPPX_NESO_MOVE 2
/* flags & OPPX_NESO_UPDATE: opp_par_loop_all__update_pos_kernel (apps/app_neso_advection/kernels.h:41-54);
 * flags & OPPX_NESO_MOVE: opp_particle_move__move_kernel (apps/app_neso_advection/kernels.h:56-113;
 * apps_codegen/app_neso_advection_mdir_cg/hip/move_kernel_loop.hpp:166-420); both in one pass when both are set,
 * over counters->size particles. The counters' hop statistics are reset first; `hops` (nullable) receives
 * per-particle hop counts. */
int oppx_neso_step(int64_t n_max, const oppx_neso_parts* parts, const oppx_quad_cell* cells, int64_t n_cells,
                   double dt, double cell_width, double ext_x, double ext_y, int flags,
                   oppx_move_counters* counters, int32_t* hops, void* stream);

/* opp_par_loop_all__verify_kernel (apps/app_neso_advection/kernels.h:115-150;
 * apps_codegen/app_neso_advection_mdir_cg/hip/verify_kernel_loop.hpp): adds to *count (device int) the
 * number of the counters->size particles that are not inside their cell: the cell's global index
 * c_gbl_idx[p2c] (device [n_cells]) gives (ix, iy) = (g mod nx, g div nx), and a particle counts when
 * ix or iy is negative or a coordinate lies outside [i * cell_width, i * cell_width + cell_width].
 * A p2c outside [0, n_cells) is counted in counters->faults instead. pos is SoA with `stride`. */
int oppx_neso_verify(int64_t n_max, int64_t stride, const double* pos, const int32_t* p2c, const int32_t* c_gbl_idx,
                     int64_t n_cells, int nx, double cell_width, oppx_move_counters* counters, int32_t* count,
                     void* stream);

/* *mismatches (device) = the number of quad records that differ, bit for bit, from those of the
 * nx x ny periodic grid of apps/app_neso_advection/advec_misc.h:115-143 (corner (x, y) * cell_width,
 * neighbours wrapped on both axes; cell = x + nx * y). Zero lets oppx_neso_step_grid replace the
 * record gathers. */
int oppx_quad_grid_check(const oppx_quad_cell* cells, int64_t n_cells, int nx, int ny, double cell_width,
                         int32_t* mismatches, void* stream);

/* oppx_neso_step on a mesh that oppx_quad_grid_check found to be the nx x ny periodic grid: the walk
 * visits the same cells with the same hop counts, each visited record computed in registers instead
 * of gathered. */
int oppx_neso_step_grid(int64_t n_max, const oppx_neso_parts* parts, int nx, int ny, double dt, double cell_width,
                        double ext_x, double ext_y, int flags, oppx_move_counters* counters, int32_t* hops,
                        void* stream);

/* ---------------------------------------------------------------------------------------
 * Compaction
 * ------------------------------------------------------------------------------------- */

/* Launch shape of the walk kernels (no reference counterpart: the reference fixes its block size
 * by opp_threads_per_block / OPP_BLOCK_SIZE_4, apps_codegen/app_fempic_cg/hip/move_kernel_loop.hpp:221-225). */
int oppx_set_tuning(const oppx_tuning* t);
int oppx_get_tuning(oppx_tuning* t);

/* 64-bit words the removal bitmap of a move over n_max particles needs (any tile size). */
int64_t oppx_hole_bits_words(int64_t n_max);

/* Scratch needed by oppx_holefill for n_max particles (bytes). */
size_t oppx_holefill_scratch_bytes(int64_t n_max);

/* Hole-fill of opp_finalize_particle_move (opp_lib/src/core/opp_particle_organize_core.cpp:124-218;
 * omp opp_lib/src/omp/opp_particle_mover.cpp:153-218; hip opp_lib/src/hip/opp_particle_sorter.cpp:271-359):
 * the k-th hole (ascending) below the new size receives the k-th live record (descending)
 * above it, for every dat in `dats` (host array of n_dats <= 16 SoA descriptors sharing `stride`). Updates
 * counters->size -= removed and counters->holes_low. */
int oppx_holefill(int64_t n_max, int64_t stride, oppx_move_counters* counters, const uint64_t* hole_bits,
                  const oppx_dat_desc* dats, int n_dats, void* scratch, size_t scratch_bytes, void* stream);

/* No reference counterpart: oppx_holefill for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same arguments and results as
 * oppx_holefill without the workspace: the scratch is a static device array of
 * OPPX_HOLEFILL_NOALLOC_SCRATCH_BYTES bytes that every GPU the library runs on holds until the process
 * ends; n_max above OPPX_HOLEFILL_NOALLOC_MAX_PARTICLES is rejected (OPPX_ERR_ARG). Allocates nothing.
 * All calls on one GPU share its array: calls that overlap on one GPU (two streams, two threads) race. */
#define OPPX_HOLEFILL_NOALLOC_MAX_PARTICLES 134217728
#define OPPX_HOLEFILL_NOALLOC_SCRATCH_BYTES 8454144
int oppx_holefill_noalloc(int64_t n_max, int64_t stride, oppx_move_counters* counters, const uint64_t* hole_bits,
                          const oppx_dat_desc* dats, int n_dats, void* stream);

/* FP32 compensated deposit over cell-sorted particles (p2c ascending, e.g. after oppx_sort_by_cell):
 * per-cell Kahan sums of lc in index order into cell_partials [4*n_cells] (scratch, zeroed here), then per
 * node a Kahan sum of its (cell, vertex) partials in node_ref order: entries [node_start[v], node_end[v])
 * of node_ref (4*cell + vertex), or [node_start[v], node_start[v+1]) when node_end is NULL. Writes
 * node_charge (double, nullable) and/or node_charge_f32 (nullable). The OPP_INC deposit of
 * apps/app_fempic/kernels.h:164-175 in the segmented-reduction form of
 * apps_codegen/app_fempic_cg/hip/deposit_charge_on_nodes_kernel_loop.hpp:98-158; every addition order
 * is fixed by the data (deterministic). scratch: oppx_fempic_deposit_scratch_bytes(n_max) bytes. A p2c that
 * breaks the precondition (a valid cell right after a larger one, or after a removed particle) is reported:
 * each such run start adds one to counters->faults, and the charge is then not the deposit's. */
size_t oppx_fempic_deposit_scratch_bytes(int64_t n_max);
int oppx_fempic_deposit_sorted_f32(int64_t n_max, int64_t stride, oppx_move_counters* counters, const int32_t* p2c,
                                   const double* lc, int64_t n_cells, const int32_t* node_start, const int32_t* node_end,
                                   const int32_t* node_ref, int64_t n_nodes, float* cell_partials, double* node_charge,
                                   float* node_charge_f32, void* scratch, size_t scratch_bytes, void* stream);

/* oppx_fempic_deposit_sorted_f32 over a lazy sort (the same deposit bit for bit): p2c is sorted, while lc is
 * still in the order before the sort, and perm[q] (oppx_sort_by_cell's perm_out) is the index in lc of
 * sorted slot q. The run sums read lc through perm, so the sorted lc is never written. lc_f32 (nullable,
 * 16-byte aligned, 4 * n_max floats of the caller's free memory) receives lc in FP32 in sorted order, from
 * which runs longer than 64 particles are summed instead of gathering through perm again. A perm entry
 * outside [0, stride) reads particle 0 and adds one to counters->faults. */
int oppx_fempic_deposit_sorted_f32_perm(int64_t n_max, int64_t stride, oppx_move_counters* counters,
                                        const int32_t* p2c, const double* lc, const int32_t* perm, float* lc_f32,
                                        int64_t n_cells, const int32_t* node_start, const int32_t* node_end,
                                        const int32_t* node_ref, int64_t n_nodes, float* cell_partials,
                                        double* node_charge, float* node_charge_f32, void* scratch,
                                        size_t scratch_bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * Sort
 * ------------------------------------------------------------------------------------- */

/* Scratch needed by oppx_sort_by_cell for n_max particles (bytes). */
size_t oppx_sort_scratch_bytes(int64_t n_max);

/* opp_particle_sort and the Sort_All / Sort_Periodic finalize (opp_lib/src/core/opp_particle_organize_core.cpp:98-121;
 * opp_lib/src/omp/opp_particle_mover.cpp:223-265; hip opp_lib/src/hip/opp_particle_sorter.cpp:188-267): orders the
 * counters->size particles by p2c, removed particles (MAX_CELL_INDEX) last, stable by (p2c, index). Each dat in
 * `dats` is gathered into the matching separate buffer of `out_dats` (same dims; the caller swaps them in);
 * a dat whose base is `p2c` itself is written from the sorted keys instead of gathered. perm_out
 * (nullable) receives the source index of every output slot. drop_removed != 0 also applies the
 * finalize: counters->removed = number removed, counters->size -= removed. */
int oppx_sort_by_cell(int64_t n_max, int64_t stride, oppx_move_counters* counters, int64_t n_cells,
                      const int32_t* p2c, const oppx_dat_desc* dats, const oppx_dat_desc* out_dats, int n_dats,
                      int32_t* perm_out, int drop_removed, void* scratch, size_t scratch_bytes, void* stream);

/* out_dats[q][i] = dats[q][perm[i]] for i < counters->size: materialises a permutation left pending
 * by oppx_fempic_step_fused_gather's caller (e.g. before the state is read back). */
int oppx_gather_dats(int64_t n_max, int64_t stride, const oppx_move_counters* counters, const int32_t* perm,
                     const oppx_dat_desc* dats, const oppx_dat_desc* out_dats, int n_dats, void* stream);

/* Range lookup in cell-sorted particles: out[j] = the first index in [0, counters->size) whose p2c is
 * >= keys[j] (device arrays, k keys). With keys = {first cell, end cell} of a mesh region this gives
 * the region's particle range (no reference counterpart: used to hand back one instance of a
 * batched set). */
int oppx_lower_bound(int64_t n_max, const oppx_move_counters* counters, const int32_t* p2c, const int32_t* keys,
                     int64_t k, int64_t* out, void* stream);

/* ---------------------------------------------------------------------------------------
 * Batched ensembles and continuous-batching serving for C++ hosts (no reference counterpart:
 * OP-PIC runs one instance per process). Host code over the calls above (op-pic_amd/csrc/host/arena.cpp).
 * ------------------------------------------------------------------------------------- */

/* One FEM-PIC instance as host arrays, indices local to the instance (read during oppx_arena_admit only). */
typedef struct oppx_instance {
    int64_t n_cells, n_nodes, n_part;
    const double* c_det;       /* [16*n_cells] */
    const double* c_vol;       /* [n_cells] */
    const int32_t* c2c;        /* [4*n_cells], -1 = boundary face */
    const int32_t* c2n;        /* [4*n_cells] */
    const double* c_ef;        /* [3*n_cells] */
    const double* pos;         /* [3*n_part], x y z per particle */
    const double* vel;         /* [3*n_part] */
    const int32_t* p2c;        /* [n_part] */
    const int32_t* overlay;    /* nullable: direct-hop overlay [dh_dims[0]*dh_dims[1]*dh_dims[2]] */
    int64_t dh_dims[3];
    double dh_min[3];
    double dh_inv_gs;
    double charge, mass, dt;   /* push coefficient k = charge / mass * dt */
} oppx_instance;

/* A resident arena of instances stepped together (opaque). Capacities: concatenated cells, nodes,
 * particles and overlay words, and instance slots. Device memory is allocated here, once. */
typedef struct oppx_arena oppx_arena;
int oppx_arena_create(int64_t max_cells, int64_t max_nodes, int64_t max_particles, int64_t max_overlay,
                      int32_t max_insts, void* stream, oppx_arena** out);
int oppx_arena_destroy(oppx_arena* arena);

/* Admit an instance between steps: its mesh and particles are copied into free ranges (indices
 * rebased) and it joins the next step. *slot identifies it. OPPX_ERR_NOMEM: no room at the moment
 * (nothing changed; retry after a retirement). OPPX_ERR_ARG: an index of the instance out of range. */
int oppx_arena_admit(oppx_arena* arena, const oppx_instance* inst, int32_t* slot);

/* One step of every resident instance: batched push + direct hop + walk
 * (oppx_fempic_batch_step_fused), Sort_All finalize (oppx_sort_by_cell), FP32 compensated deposit
 * (oppx_fempic_deposit_sorted_f32). Returns when the step is complete. */
int oppx_arena_step(oppx_arena* arena);

/* Hand an instance back and free its ranges: *n_out particles remain; pos/vel (3 per particle),
 * lc (4), p2c (local cell indices) receive them in the instance's order alone (each nullable,
 * capacity `cap` particles), charge_f32 (nullable) its n_nodes FP32 node charges of the last step.
 * Call after a step (OPPX_ERR_STATE when an instance was admitted since the last step). */
int oppx_arena_retire(oppx_arena* arena, int32_t slot, int64_t cap, int64_t* n_out, double* pos, double* vel,
                      double* lc, int32_t* p2c, float* charge_f32);

/* Host copy of the arena's move counters (faults and hop statistics of the last step). */
int oppx_arena_counters(const oppx_arena* arena, oppx_move_counters* out);

#ifdef __cplusplus
}
#endif

#endif /* OPPX_H */
