/*
 * pfem2_amd.h — C ABI of the MI355X particle stage of PFEM2.
 *
 * One handle holds one simulation: the mesh (uploaded once), the nodal
 * velocity fields V and V_old, and the particles, all resident in HBM.
 * Every entry point returns 0 on success or a negative PF2_E* code; the
 * message of the last failure on the calling thread is pf2_last_error().
 * All work of a handle is issued on one HIP stream owned by the handle;
 * calls on one handle must not overlap (the reference is single-threaded
 * per process as well).  Calls on different handles are independent.
 *
 * Each entry point names the member of the reference interface it replaces
 * (paths relative to the andreyypopov/pfem2 repository).  The deal.II
 * objects behind those members (Triangulation, DoFHandler, solution
 * vectors) become plain arrays here:
 *   vertices       n_vertices x dim doubles, row-major
 *   cell_vertices  n_cells x 2^dim ints, deal.II vertex order
 *                  (lexicographic: bit d of the local index = unit coordinate d)
 *   cell_dofs      n_cells x 2^dim ints: cell->vertex_dof_index(v, 0)
 *   nodal fields   dim x n_dofs doubles, component-major
 *                  (solutionVx, solutionVy[, solutionVz] back to back)
 *   particles      cell-major, ascending id inside a cell;
 *                  per-particle vectors n x dim doubles, row-major
 */
#ifndef PFEM2_AMD_H
#define PFEM2_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF2_OK 0
#define PF2_E_ARG (-1)      /* invalid argument */
#define PF2_E_HIP (-2)      /* HIP runtime failure */
#define PF2_E_STATE (-3)    /* call not valid in the current state */
#define PF2_E_CAPACITY (-4) /* particle capacity exceeded */

#define PF2_FP64 0          /* FP64 everywhere (the reference's arithmetic) */
#define PF2_FP32_MIXED 1    /* cell + FP32 reference coordinates, FP32 velocities */

typedef struct pf2_handle_s* pf2_handle;

/* Triangulation + DoFHandler of the reference's pfem2Solver
 * (cylinder2D/pfem2particle.h:170-175; TurekBenchmark/pfem2particle.h:196-205).
 * The reference's velocity field is FE_Q<dim>(1) (cylinder2D/pfem2particle.cpp:473);
 * degree 2 (FE_Q<dim>(2) on the same Q1 geometry) extends it: cell_dofs then
 * lists 3^dim DoFs per cell in deal.II's order (vertices, lines, [faces,]
 * interior; line and face numbering of deal.II's GeometryInfo). */
typedef struct pf2_mesh {
    int dim;                    /* 2 or 3 */
    int n_cells;
    int n_vertices;
    int n_dofs;                 /* velocity DoFs per component */
    const double* vertices;
    const int* cell_vertices;
    const int* cell_dofs;       /* degree 1: NULL means the same as cell_vertices;
                                   degree 2: n_cells x 3^dim, required */
    int degree;                 /* velocity field degree: 1 (0 also means 1) or 2;
                                   degree-2 handles take no walls */
} pf2_mesh;

/* Compile-time knobs of the reference made run-time
 * (cylinder2D/pfem2particle.h:4-5: PARTICLES_MOVEMENT_STEPS,
 * MAX_PARTICLES_PER_CELL_PART), the seeding lattice
 * (seed_particles' `quantities`, cylinder2D/pfem2particle.h:132) and the
 * time step (pfem2Solver::time_step, cylinder2D/pfem2particle.h:161). */
typedef struct pf2_params {
    double dt;
    int substeps;
    int max_per_part;
    int quantities[3];
    int precision;              /* PF2_FP64 or PF2_FP32_MIXED */
    int keep_velocity_ext;      /* store the last advection velocity per particle
                                   (needed by pf2_sort_particles after a mesh move) */
} pf2_params;

const char* pf2_last_error(void);

/* pfem2Solver constructor + build_grid/setup_system hand-over
 * (cylinder2D/pfem2particle.cpp:467-480).  Also builds the vertex-to-cell
 * star and the cell-centre directions, i.e. pfem2ParticleHandler::initialize_maps
 * (TurekBenchmark/pfem2particle.cpp:215-219).  `capacity` = maximum number of
 * particles; 0 chooses the bound that reseeding guarantees. */
int pf2_create(const pf2_mesh* mesh, const pf2_params* params, long long capacity, pf2_handle* out);
/* pf2_create inside a caller-owned device memory block of block_bytes
 * (at least pf2_estimate_bytes) and on a caller-owned HIP stream (a
 * hipStream_t; NULL creates one).  No counterpart in the reference: it lets a
 * server place many simulations in one allocation without hipMalloc/hipFree
 * while others run.  pf2_destroy then frees neither the block nor the stream. */
int pf2_create_in(const pf2_mesh* mesh, const pf2_params* params, long long capacity, void* block,
                  unsigned long long block_bytes, void* stream, pf2_handle* out);
/* Device bytes pf2_create takes for this mesh / params / capacity (upper
 * bound), and the bytes a handle holds. */
int pf2_estimate_bytes(const pf2_mesh* mesh, const pf2_params* params, long long capacity,
                       unsigned long long* bytes);
int pf2_memory_bytes(pf2_handle h, unsigned long long* bytes);
int pf2_destroy(pf2_handle h);
int pf2_set_params(pf2_handle h, const pf2_params* params);

/* Nodal fields.  V replaces solutionVx/Vy[/Vz]; V_old replaces
 * old_solutionVx/... (cylinder2D/pfem2particle.h:163-164).  Either pointer
 * may be NULL to keep the resident copy.  After a projection both resident
 * copies hold the projected field (the reference sets old_solution =
 * solution before the FEM step, cylinder2D/cylinder2D.cc:389-391), so a
 * driver uploads only the new V each step. */
int pf2_upload_fields(pf2_handle h, const double* V, const double* V_old);
/* Same with device pointers (copied on the handle's stream). */
int pf2_upload_fields_device(pf2_handle h, const double* V, const double* V_old);
int pf2_download_field(pf2_handle h, double* V);

/* pfem2Solver::seed_particles (cylinder2D/pfem2particle.cpp:563-580). */
int pf2_seed_particles(pf2_handle h);
/* pfem2ParticleHandler::insert_particle (cylinder2D/pfem2particle.cpp:178-184)
 * for n particles given by cell and reference coordinates; the location is
 * mapped from the reference point, the velocity is interpolated from V when
 * v is NULL, and ids continue the particle counter in input order. */
int pf2_insert_particles(pf2_handle h, int n, const int* cell, const double* xi, const double* v);

/* pfem2Solver::correct_particles_velocities (cylinder2D/pfem2particle.cpp:582-603). */
int pf2_correct_particles_velocities(pf2_handle h);
/* pfem2Solver::move_particles (cylinder2D/pfem2particle.cpp:605-647):
 * substeps of advection + relocation, then reseed/delete in every cell. */
int pf2_move_particles(pf2_handle h);
/* pfem2Solver::distribute_particle_velocities_to_grid
 * (TurekBenchmark/pfem2particle.cpp:858-918) including the wall fix-up of
 * cylinderDyM/pfem2particle.cpp:960-992.  V_out (host, component-major) may
 * be NULL. */
int pf2_distribute_particle_velocities_to_grid(pf2_handle h, double* V_out);
/* No counterpart in the reference.  pf2_distribute_particle_velocities_to_grid
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the
 * per-cell projection sums go to a static device array of
 * PF2_STATIC_SCRATCH_BYTES per GPU instead of the handle's buffer, and the
 * projected field is written to V_out_device (device memory, dim x n_dofs,
 * component-major; NULL: kept resident only), with no host synchronisation.
 * Results equal those of pf2_distribute_particle_velocities_to_grid.
 * Needs n_cells * 2^dim * (dim + 1) * 8 <= PF2_STATIC_SCRATCH_BYTES
 * (PF2_E_CAPACITY otherwise) and the handle's device current.  All calls on
 * one GPU share the array: calls that overlap on one GPU (other streams,
 * other threads) race. */
#define PF2_STATIC_SCRATCH_BYTES (64ull << 20)
int pf2_distribute_particle_velocities_to_grid_static(pf2_handle h, double* V_out_device);
/* correct -> move -> distribute in one call (the driver sequence of
 * cylinder2D/cylinder2D.cc:865-867); V_out may be NULL. */
int pf2_step(pf2_handle h, double* V_out);
/* One step of a driver whose FEM update leaves V in device memory
 * (component-major): V is refreshed from V_device, then correct -> move ->
 * distribute.  No host synchronisation. */
int pf2_step_with_field(pf2_handle h, const double* V_device);
/* Batched steps (no counterpart in the reference): pf2_batch_step does
 * pf2_step_with_field(hs[i], fields[i]) for n simulations with one launch per
 * kernel of the step for all of them, on the batch's stream (ordered after
 * each simulation's earlier work, and each simulation's later work after
 * it).  Each simulation's results equal those of its own step, bit for bit.
 * The simulations share one dimension and precision, have degree-1 fields
 * and no walls, and appear once each.  A batch holds up to max_sims
 * simulations per step; stream NULL creates one.  Calls on one batch must
 * not overlap. */
typedef struct pf2_batch_s* pf2_batch;
int pf2_batch_create(int max_sims, void* stream, pf2_batch* out);
int pf2_batch_step(pf2_batch b, int n, const pf2_handle* hs, const double* const* fields);
int pf2_batch_destroy(pf2_batch b);
/* Capture pf2_step / pf2_step_with_field (when no V_out is requested and
 * profiling is off) as HIP graphs, one per buffer parity, replayed on later
 * steps; results are identical to the eager launches.  pf2_set_params,
 * pf2_set_walls, pf2_seed_particles and pf2_insert_particles drop the
 * graphs (they change launch arguments or sizes).  Off by default. */
int pf2_set_graphs(pf2_handle h, int on);

/* Wall / body nodes of the projection fix-up (wallsAndBodyDoFs,
 * cylinderDyM/pfem2particle.h:253): vertex index, DoF index, boundary values
 * (n x dim) and a mask of enforced components per node.  n = 0 clears. */
int pf2_set_walls(pf2_handle h, int n, const int* vertex, const int* dof, const double* bc, const int* mask);

/* pfem2ParticleHandler::sort_particles_into_subdomains_and_cells
 * (cylinder2D/pfem2particle.cpp:214-445), as called after a mesh move
 * (cylinderDyM/cylinderDyM.cc:1159). */
int pf2_sort_particles_into_subdomains_and_cells(pf2_handle h);
/* pfem2Solver::check_empty_cells (cylinderDyM/pfem2particle.cpp:759-763). */
int pf2_check_empty_cells(pf2_handle h);
/* pfem2Solver::transform_grid (cylinderDyM/pfem2particle.cpp:1252-1273):
 * new absolute vertex coordinates (n_vertices x dim).  The cached centre
 * directions are kept unless recompute_maps != 0, as in the reference where
 * initialize_maps runs once. */
int pf2_update_vertices(pf2_handle h, const double* vertices, int recompute_maps);
/* The moving-mesh sequence of cylinderDyM/cylinderDyM.cc:1157-1160 in one
 * call: transform_grid (new vertex positions), then
 * sort_particles_into_subdomains_and_cells and check_empty_cells.  The
 * relocation and the reseed share one census; survivors, ids, counts and
 * statistics equal those of the three calls.  The FP32-mixed path keeps its
 * FP32 state (locations from the mesh before the move, relocation in FP64).
 * Without step graphs the census's results are applied in place and the
 * particles are regrouped by the next move (or by the first call that
 * reads them, such as pf2_download_particles or a projection); callers see
 * the same particles, ids and order either way. */
int pf2_move_mesh(pf2_handle h, const double* vertices, int recompute_maps);
/* A prescribed mesh motion X(t) = X0 + a(t) W, for the moving-mesh case
 * without a host elasticity solve: pf2_set_motion stores the current
 * vertices as X0 and the displacement shape W (n_vertices x dim, host);
 * pf2_move_mesh_scaled(h, a) is pf2_move_mesh with the vertices X0 + a W
 * formed on the device.  Needs keep_velocity_ext (which reserves X0, W). */
int pf2_set_motion(pf2_handle h, const double* W);
int pf2_move_mesh_scaled(pf2_handle h, double a);
/* pfem2ParticleHandler::initialize_maps (TurekBenchmark/pfem2particle.cpp:215-219). */
int pf2_initialize_maps(pf2_handle h);

/* pfem2Solver::reinterpolate_fields (cylinderDyM/pfem2particle.cpp:1144-1250),
 * called before the mesh moves (cylinderDyM/cylinderDyM.cc:1157): the
 * resident V and n_extra further nodal scalar fields on the same DoFs
 * (extra: n_extra x n_dofs, host, updated in place, e.g. the pressure) take
 * the values of the current mesh's fields at the vertices displaced by U
 * (n_vertices x dim).  Vertices on a boundary face keep their values.  The
 * search starts in the vertex's first cell and tries the rest of its star
 * (ascending), accepting reference coordinates within 1e-5 of the unit
 * cell.  V_old is not changed.  Q1 fields. */
int pf2_reinterpolate_fields(pf2_handle h, const double* U, int n_extra, double* extra);

/* pfem2Solver::calculate_loads (TurekBenchmark/pfem2particle.cpp:920-970;
 * pressure part: cylinder2D/pfem2particle.cpp:691-735), 2-D Q1: force
 * coefficients over the boundary faces of a patch, given as (cell, face
 * number) pairs in the reference's loop order (cells ascending, faces
 * ascending).  V: velocity (2 x n_dofs, host, component-major; NULL: the
 * resident V), P: pressure (n_dofs).  out5 = {Cx_nu, Cx_p, Cy_nu, Cy_p,
 * P[dof_a] - P[dof_e]} with C = 2 F / (rho u_mean^2 diam); a negative DoF
 * drops its term of the pressure difference. */
typedef struct pf2_loads_params {
    double mu, rho, u_mean, diam;
    int dof_a, dof_e;
} pf2_loads_params;
int pf2_calculate_loads(pf2_handle h, const double* V, const double* P, int n_faces, const int* faces,
                        const pf2_loads_params* lp, double* out5);

/* pfem2ParticleHandler::n_global_particles (cylinder2D/pfem2particle.cpp:186-189). */
int pf2_n_global_particles(pf2_handle h, long long* n);
/* iteration over particle_handler.begin()/end() reading location, reference
 * location, velocity and id (cylinder2D/cylinder2D.cc:816-835).  Any output
 * pointer may be NULL; arrays must hold pf2_n_global_particles entries. */
int pf2_download_particles(pf2_handle h, int* cell, uint32_t* id, double* x, double* xi, double* v);
/* pfem2ParticleHandler::n_particles_in_cell for every cell (n_cells ints). */
int pf2_cell_counts(pf2_handle h, int* counts);
/* the particleCount member of pfem2Solver (last id handed out) */
int pf2_particle_counter(pf2_handle h, uint32_t* count);

/* Counters since creation: [0] particles lost in relocation, [1] deleted by
 * the per-part limit, [2] spawned by reseeding, [3] relocation searches,
 * [4] zero-weight projection nodes. */
int pf2_stats(pf2_handle h, unsigned long long* out5);
/* Work counts of the last pf2_move_particles / pf2_step (no reference
 * counterpart; diagnostics): [0] particles before the step, [1] particles
 * after it, [2..3] particles queued for the vertex-star search in each
 * relocation round of the last move (the FP64 path searches in one round,
 * the FP32-mixed path in two; an unused round reads 0), [4] particles that
 * the last re-sort after a mesh move relocated by a star search, [5]
 * particles that changed cell in the last regroup.
 * Returns the number of entries written (at most cap). */
int pf2_step_info(pf2_handle h, long long* out, int cap);
int pf2_synchronize(pf2_handle h);

/* Per-kernel timing with HIP events on the handle's stream (off by
 * default).  pf2_kernel_times writes up to `cap` entries of total
 * milliseconds and launch counts in the order of pf2_kernel_name(i), and
 * returns the number of kernels. */
int pf2_set_profiling(pf2_handle h, int on);
int pf2_kernel_times(pf2_handle h, double* total_ms, long long* launches, int cap);
const char* pf2_kernel_name(int i);

/* ------------------------------------------------------------------------
 * Serving loop (no counterpart in the reference, which runs one simulation
 * per process).  Independent simulations are submitted with their mesh,
 * parameters, a fixed nodal field V and a step count; each is seeded with
 * the lattice and then advanced n_steps times by pf2_step_with_field.
 * Admission is first come, first served while the simulation's device
 * footprint fits the HBM budget, which the server reserves once at creation
 * and hands out in ranges.  Single-threaded: calls on one server must not
 * overlap. */
typedef struct pf2_server_s* pf2_server;
typedef struct pf2_server_config {
    unsigned long long hbm_budget;  /* bytes reserved for all simulations (0: 80% of free HBM) */
    int max_active;                 /* simulations in flight at once (0: limited by HBM only) */
    int steps_per_chunk;            /* steps issued per simulation between completion events (0: 8) */
    int chunks_ahead;               /* chunks kept queued per simulation (0: 2) */
    int use_graphs;                 /* replay each simulation's step as a HIP graph */
    int batched;                    /* step all active simulations of one dimension and precision
                                       together with pf2_batch_step (degree-1 fields) */
    int batch_max_cells;            /* batched mode: only simulations of at most this many cells are
                                       batched; larger ones replay their own step graphs on their own
                                       streams (0: every simulation is batched) */
} pf2_server_config;

const char* pf2_server_last_error(void);
int pf2_server_create(const pf2_server_config* cfg, pf2_server* out);
int pf2_server_destroy(pf2_server s);
/* Copies every input; V is dim x n_dofs, component-major.  Fails with
 * PF2_E_CAPACITY when the simulation alone exceeds the budget. */
int pf2_server_submit(pf2_server s, const pf2_mesh* mesh, const pf2_params* params, const double* V, int n_steps,
                      long long* request_id);
/* A moving-mesh request: after each step k (1-based) the mesh moves to
 * X0 + amplitude * sin(omega * k * dt + phase) * W (pf2_move_mesh_scaled);
 * W is n_vertices x dim.  keep_velocity_ext is set for it. */
int pf2_server_submit_moving(pf2_server s, const pf2_mesh* mesh, const pf2_params* params, const double* V,
                             int n_steps, const double* W, double amplitude, double omega, double phase,
                             long long* request_id);
/* Retires finished simulations, admits pending ones, issues steps; never
 * waits for the GPU.  n_pending / n_active may be NULL. */
int pf2_server_poll(pf2_server s, int* n_pending, int* n_active);
/* Polls until every submitted simulation has finished. */
int pf2_server_drain(pf2_server s);
/* Result of a finished simulation: the last projected field (dim x n_dofs),
 * the final particle count and the latency from submission to completion.
 * PF2_E_STATE while it runs; the simulation's error code if it failed. */
int pf2_server_result(pf2_server s, long long request_id, double* V_out, long long* n_particles, double* latency_ms);
/* [0] finished, [1] failed, [2] p50 latency ms, [3] p99 latency ms,
 * [4] particle-updates, [5] seconds from first submission to last
 * completion, [6] particle-updates/s over that span, [7] peak bytes in use,
 * [8] most simulations active at once, [9] mean latency ms.  Returns the
 * number of entries written (at most cap). */
int pf2_server_stats(pf2_server s, double* out, int cap);

#ifdef __cplusplus
}
#endif

#endif /* PFEM2_AMD_H */
