/* drape_hip.h — C-ABI of the MI355X implicit-Euler Newton–PCG cloth step.
 *
 * This is the drop-in boundary for the reference's `drape::sim::Solver` surface
 * (libs/sim/include/drape/sim/solver.hpp:85-98) and its backend switch
 * (libs/sim/include/drape/sim/solver_factory.hpp:12-20). Every entry point is plain C:
 * opaque handles, plain pointers and sizes, an int status code, and a thread-local last-error
 * string. A C++ caller uses the header-only `drape::sim::hip::HipSolver` wrapper
 * (include/drape/sim/hip_solver.hpp), which maps status codes back to the reference's
 * exception types.
 *
 * Status codes mirror the exceptions the reference throws:
 *   DRAPE_OK                      0
 *   DRAPE_ERR_INVALID_ARGUMENT    std::invalid_argument  (sim_model.cpp:81-86, :201, :257-258; cpu_solver.cpp:341)
 *   DRAPE_ERR_OUT_OF_RANGE        std::out_of_range      (sim_model.cpp:15, :184, :190)
 *   DRAPE_ERR_RUNTIME             std::runtime_error     (device failures, metal_device.cpp:58-88, :136)
 *   DRAPE_ERR_LOGIC               std::logic_error       (API misuse, metal_device.cpp:108, :115, :126)
 */
#ifndef DRAPE_HIP_H
#define DRAPE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRAPE_OK 0
#define DRAPE_ERR_INVALID_ARGUMENT 1
#define DRAPE_ERR_OUT_OF_RANGE 2
#define DRAPE_ERR_RUNTIME 3
#define DRAPE_ERR_LOGIC 4

#define DRAPE_PHASE_ASSEMBLING 0 /* SimPhase::Assembling (solver.hpp:17) */
#define DRAPE_PHASE_SETTLING 1   /* SimPhase::Settling */

/* FabricPhysical (libs/core/include/drape/core/fabric.hpp:10-18), SI units. */
typedef struct drape_fabric {
  double stretch_warp, stretch_weft, stretch_bias; /* N/m */
  double strain_limit_warp, strain_limit_weft;     /* fraction */
  double bend_warp, bend_weft;                     /* N m */
  double weight;                                   /* kg/m^2 */
  double thickness;                                /* m */
  double friction, damping;                        /* 0..1 */
} drape_fabric;

/* SimSettings (solver.hpp:19-30). drape_settings_default() fills the reference defaults. */
typedef struct drape_settings {
  double gravity[3];
  double collision_offset;
  double assembly_closing_speed;
  double assembly_extra_damping;
  double damping_scale;
  double strain_limit_stiffness_scale;
  double seam_stiffness;
  double friction_slip_speed;
  int32_t max_cg_iterations;
  int32_t pad0;
  double cg_tolerance;
} drape_settings;

/* SimMesh (libs/core/include/drape/core/sim_mesh.hpp:11-27), one garment, flat arrays. */
typedef struct drape_mesh_desc {
  uint32_t n_vertices;
  const double* rest;            /* 2 per vertex, pattern space (m) */
  const uint32_t* vertex_piece;  /* 1 per vertex */
  uint32_t n_triangles;
  const uint32_t* triangles;     /* 3 per triangle, CCW in pattern space */
  const uint32_t* triangle_piece;
  const double* vertex_area;     /* 1/3 of incident rest areas (m^2) */
  const uint8_t* boundary;       /* may be NULL (unused by the step) */
  uint32_t n_pieces;
  const double* piece_grain;     /* 2 per piece, unit grainline */
  uint32_t n_seam_pairs;
  const uint32_t* seam_pairs;    /* 4 per pair: a, b, seam, index (SeamPair, sim_mesh.hpp:8-13) */
} drape_mesh_desc;

/* SimGarment (solver.hpp:32-37). */
typedef struct drape_garment_desc {
  const char* id;
  drape_mesh_desc mesh;
  const double* initial_positions; /* 3 per mesh vertex */
  uint32_t n_piece_fabrics;
  const drape_fabric* piece_fabric; /* one per piece */
} drape_garment_desc;

/* CollisionField (libs/core/include/drape/core/collision_field.hpp:11-20): signed distance on a regular grid,
 * samples at origin + (i,j,k) * cell_size, x fastest. The library copies the values at build time. */
typedef struct drape_field_desc {
  double origin[3];
  double cell_size;
  int32_t dims[3];
  int32_t pad0;
  double band;
  const float* values; /* dims[0]*dims[1]*dims[2] */
} drape_field_desc;

/* SimScene (solver.hpp:49-55). body may be NULL. Pin/force particle indices are global. */
typedef struct drape_scene_desc {
  uint32_t n_garments;
  const drape_garment_desc* garments;
  const drape_field_desc* body;
  uint32_t n_pins;
  const uint32_t* pin_particles;
  const double* pin_positions; /* 3 per pin */
  uint32_t n_forces;
  const uint32_t* force_particles;
  const double* force_vectors; /* 3 per force, newtons */
  drape_settings settings;
} drape_scene_desc;

/* SimStats (solver.hpp:63-69). */
typedef struct drape_stats {
  int32_t has_non_finite;
  int32_t pad0;
  double kinetic_energy;
  double rms_speed;
  double max_seam_gap;
  double max_body_penetration;
} drape_stats;

/* Work counters of the step (not exposed by the reference's Solver; phase-b spec :196 notes their absence).
 * Cumulative since build or the last drape_solver_reset_counters(). Summed over segments. */
typedef struct drape_counters {
  uint64_t substeps;
  uint64_t newton_iterations;
  uint64_t pcg_iterations;
  uint64_t line_search_trials;
  uint64_t warm_start_fallbacks; /* learned guesses rejected because they raised the energy */
  uint64_t warm_start_accepted;
} drape_counters;

typedef struct drape_solver drape_solver; /* opaque */

/* Last error message of the calling thread ("" if none). */
const char* drape_last_error(void);

/* Reference defaults of SimSettings (solver.hpp:19-30). */
int drape_settings_default(drape_settings* out);

/* Backend probe (solver_factory.hpp:17-18, isAvailable): 1 when a gfx950 device is usable, else 0 (reason in
 * drape_last_error). */
int drape_hip_available(void);

/* makeSolver(SolverBackend::Hip) (solver_factory.hpp:20): runtime_error naming "hip" when no device. */
int drape_solver_create(drape_solver** out);
void drape_solver_destroy(drape_solver* s);

/* Solver::build (solver.hpp:88). The scene is copied except body values, which are copied to the device. */
int drape_solver_build(drape_solver* s, const drape_scene_desc* scene);

/* Batched build: n independent scenes stepped together, one segment each (own particles, own body field, own
 * line search and convergence). Particles are concatenated in scene order. drape_solver_build(s, sc) is
 * drape_solver_build_batch(s, 1, sc). */
int drape_solver_build_batch(drape_solver* s, uint32_t n_scenes, const drape_scene_desc* scenes);

/* Solver::updateFabric / setPhase / range / particleCount (solver.hpp:89-90, :96-97). For batched builds the
 * garment id is looked up in scene order; the first match wins. */
int drape_solver_update_fabric(drape_solver* s, const char* garment, uint32_t n, const drape_fabric* piece_fabric);
int drape_solver_set_phase(drape_solver* s, const char* garment, int32_t phase);
int drape_solver_range(const drape_solver* s, const char* garment, uint32_t* offset, uint32_t* count);
int drape_solver_particle_count(const drape_solver* s, size_t* out);
int drape_solver_segment_count(const drape_solver* s, uint32_t* out);
int drape_solver_segment_range(const drape_solver* s, uint32_t segment, uint32_t* offset, uint32_t* count);

/* Solver::step (solver.hpp:91): substeps = max(1, substeps), h = dt / substeps, iterations = max(1, iterations)
 * (cpu_solver.cpp:333-338). */
int drape_solver_step(drape_solver* s, double dt, int32_t substeps, int32_t iterations);

/* Solver::readPositions (solver.hpp:92): out holds 3 floats per particle; n must equal particleCount()
 * (invalid_argument otherwise, cpu_solver.cpp:340-343). */
int drape_solver_read_positions(const drape_solver* s, float* out, size_t n);
int drape_solver_read_velocities(const drape_solver* s, float* out, size_t n);

/* Solver::stats (solver.hpp:93; SimModel::stats, sim_model.cpp:222-241), reduced on the device in FP64 (per-tile
 * partials summed per segment in a fixed order; the body is sampled in FP64): no positions or velocities cross
 * PCIe, 48 bytes per segment do. Over all segments of a batched build. */
int drape_solver_stats(const drape_solver* s, drape_stats* out);

/* Stats of segments [first, first + count) of a batched build, one drape_stats per segment (each as
 * drape_solver_stats of that scene built alone), from one device reduction: the per-frame check of a settle
 * loop (drape_session.cpp:32-95 calls Solver::stats every frame) for many scenes at once. out_of_range when the
 * range exceeds drape_solver_segment_count. No counterpart in the reference's Solver. */
int drape_solver_segment_stats(const drape_solver* s, uint32_t first, uint32_t count, drape_stats* out);

/* Solver::snapshot / restore (solver.hpp:94-95) as one flat byte blob: query the size, save into a caller
 * buffer, restore from one (sim_model.cpp:243-274). Little-endian layout, no padding:
 *   uint64 n_particles, uint64 n_garments, uint64 n_seams
 *   float  positions[n_particles][3], velocities[n_particles][3]
 *   int32  phase[n_garments]            (DRAPE_PHASE_*)
 *   double assembly_time[n_garments]
 *   float  seam_start_gap[n_seams]      (seam springs in build order)
 *   double contact_force[n_particles]
 * restore rejects a blob whose counts differ from the built scene (std::invalid_argument). */
int drape_solver_snapshot_size(const drape_solver* s, size_t* bytes);
int drape_solver_snapshot_save(const drape_solver* s, void* out, size_t bytes);
int drape_solver_snapshot_restore(drape_solver* s, const void* in, size_t bytes);

int drape_solver_counters(const drape_solver* s, drape_counters* out);
/* Linear-solve report (no counterpart in the reference's Solver; it makes the PCG's stopping visible): per Newton
 * iteration one solve, counted by why it ended (cpu_solver.cpp:188-204: |r| <= tol |r0|, the iteration cap,
 * p.Ap <= 0, r0 = 0), and its FP64 true relative residual |S(-grad - A step)| / |S(-grad)| at the end (S the
 * contact filter; the recursive FP32 residual the stopping test uses can drift from it). Summed over segments;
 * cumulative since build or drape_solver_reset_counters(). */
typedef struct drape_pcg_report {
  uint64_t solves;
  uint64_t exit_tolerance;
  uint64_t exit_cap;
  uint64_t exit_curvature;
  uint64_t exit_zero;
  uint64_t restarts;    /* solves continued from their FP64 true residual (drape_solver_set_true_residual) */
  double true_rel_max;  /* largest true relative residual at a solve's end */
  double true_rel_mean;
  double true_rel_last; /* of the last solve (largest over segments) */
  double true_rel_max_converged; /* largest over the solves that ended on the tolerance test */
  uint64_t persistent_aborts;    /* persistent solves whose barrier wait gave up and were re-run (halo = 2 below) */
  /* Over the solves that ended on the tolerance test: the largest true relative residual divided by max(tol, its
   * FP32 rounding floor). The floor is what storing step in FP32 leaves in |S(-grad - A step)| however exact the
   * solve: each component's rounding error, uniform within half a unit in the last place, gives an expected
   * 2^-24 sqrt(sum (A_ij step_j)^2 / 3) over every product term, relative to |r0|. */
  double true_rel_excess_max;
} drape_pcg_report;
int drape_solver_pcg_report(const drape_solver* s, drape_pcg_report* out);
int drape_solver_reset_counters(drape_solver* s);

/* Execution options. graph: capture each substep's launch sequence in a hipGraph and replay it (1, default) or
 * launch eagerly (0). Results are bit-identical either way. */
int drape_solver_set_graph(drape_solver* s, int32_t enabled);

/* PCG form: 1 (default) one kernel and one global reduction per iteration (the standard preconditioned CG with
 * p.Ap reduced directly and r.z, r.r expanded exactly over the previous iteration's vectors: the same iterates and
 * stopping rules); 0 the two-kernel standard form (cpu_solver.cpp:169-213 step by step). */
int drape_solver_set_pcg_variant(drape_solver* s, int32_t variant);
/* FP64 true-residual refinement of the one-kernel PCG (no counterpart in the reference, whose CG runs in FP64):
 * after a solve stops on the FP32 recursion's test |r| <= tol |r0|, its true residual S(-grad - A step) is formed
 * with FP64 accumulation, and a solve whose true residual is still above 4 max(tol |r0|, its FP32 rounding floor;
 * drape_pcg_report.true_rel_excess_max) continues from it (same r0, same iteration cap), at most `restarts` times
 * per solve. 0 turns it off; default 4. */
int drape_solver_set_refinement(drape_solver* s, int32_t restarts);
/* Launch shape of the one-kernel PCG iteration: `waves` wavefronts per 64-row slice and `slots` lower and upper
 * matrix slots per batch; 1x8, 1x4 (smaller register footprint: two workgroups per CU, for launches with many more
 * tiles than CUs), 2x8 and 4x8, or 0x0 (default) to time every shape during the first solves and keep the fastest.
 * halo = 1 (with 1x8 or 1x4) stages each 256-row tile's column window in LDS and gathers from there; it runs only
 * on matrices whose every tile reaches at most 1,536 consecutive rows (a step with it fails otherwise), and the
 * tuner tries it only there. halo = 2 (with 1x8) is the persistent solve: one kernel launch runs a whole PCG
 * solve, each 256-row tile's workgroup keeping its slices' matrix and private vectors in registers, the tiles of a
 * scene meeting at a barrier of their own each iteration (scenes of a batch never wait for each other); it needs
 * at most 16 + 16 matrix slots per 64-row slice and every scene's tiles resident together: one tile per CU, each
 * 256-row tile's columns within at most four row ranges of 1,536 rows in all, or, where the tiles outnumber the CUs,
 * a 512-thread workgroup per pair of consecutive tiles, each pair's columns within eight row ranges of 3,072 rows
 * (at most 512 tiles, 131,072 particle rows, per scene on MI355X; a batch's scenes pass through the CUs one after
 * another when their tiles outnumber that). If a barrier wait gives up (drape_solver_pcg_report counts
 * it), the solve is re-run with one launch per iteration and the solver stops using the persistent form. Results
 * are bit-identical for every shape. drape_solver_pcg_shape reports the shape in use (0x0 while still tuning). No
 * reference counterpart: the reference's GPU path has one thread per row (kernels.h:272-289). */
int drape_solver_set_pcg_shape(drape_solver* s, int32_t waves, int32_t slots, int32_t halo);
int drape_solver_pcg_shape(const drape_solver* s, int32_t* waves, int32_t* slots, int32_t* halo);
/* Persistent solve: polls of a barrier's generation counter before a workgroup gives up waiting (default 2^20, about
 * a second); every wait of the kernel is bounded by it. *available (optional) reports whether the persistent solve
 * fits this build and has not given up. No reference counterpart. */
int drape_solver_set_solve_spin_cap(drape_solver* s, uint32_t polls);
int drape_solver_persistent_available(const drape_solver* s, int32_t* available);

/* Kernel timing inside the step: every `every`-th PCG SpMV launch (0 = off) is bracketed by HIP events on the
 * solver's stream. drape_solver_spmv_timing reports how many launches were timed since sampling was set, their
 * mean duration in microseconds and their mean algorithmic bytes: each timed launch counts the scenes still solving
 * at the host's last look at the convergence flags (all of them when nothing was timed). The matrix is stored
 * symmetric: 40 B per block on or above the diagonal (values and column index) and 8 B per slot below it (column
 * and the index of the transposed block); plus per particle 192 B for the one-kernel form (zr, t, p gathered; r,
 * ap, step, preconditioner (6 floats), contact and row flags read; r, ap, p, zr, t, step written) or 56 B for the
 * two-kernel SpMV (row and contact flags, z and p_old gathered, p_new and ap written). */
int drape_solver_set_spmv_sampling(drape_solver* s, int32_t every);
int drape_solver_spmv_timing(const drape_solver* s, uint64_t* launches, double* mean_us, double* bytes_per_launch);

/* ---- Structure and kernel access (parity tests and benchmarks; same kernels as the step) ---- */

/* Block-CSR pattern and term slot tables built on the device (sim_model.cpp:41-74). Sizes first (any pointer may
 * be NULL to skip), then copies. */
typedef struct drape_structure_sizes {
  uint32_t n_particles, n_stretch, n_hinges, n_seams, n_blocks;
} drape_structure_sizes;
int drape_solver_structure_sizes(const drape_solver* s, drape_structure_sizes* out);
int drape_solver_read_structure(const drape_solver* s, uint32_t* row_start, uint32_t* col_index, uint32_t* diag_slot,
                                uint32_t* stretch_slots, uint32_t* hinge_slots, uint32_t* seam_slots);
/* Terms as built (sim_model.cpp:107-175): stretch vertices (3 each), hinge vertices (4 each), hinge K, seam
 * a/b (2 each). */
int drape_solver_read_terms(const drape_solver* s, uint32_t* stretch_v, uint32_t* hinge_v, double* hinge_k,
                            uint32_t* seam_ab);

/* The assembly's summation order (replaces GpuSolverCore's gather lists, gpu_solver_core.cpp:184-224, in the
 * contribution numbering of kernels.h:6-9): per particle the contributions summed into its gradient [own i,
 * stretch n+3t+a, hinge n+3S+4h+a, seam n+3S+4H+2k+a], per BSR slot those summed into its block [diag i, stretch
 * n+9t+3a+c, hinge n+9S+16h+4a+c, seam n+9S+16H+4k+2a+c], each in term order — derived from the incidence lists and
 * slot tables built on the device, which fix the order the row-owner assembly adds its terms in. Sizes first; any
 * output may be NULL. */
int drape_solver_gather_sizes(const drape_solver* s, uint32_t* n_grad, uint32_t* n_hess);
int drape_solver_read_gather(const drape_solver* s, uint32_t* grad_start, uint32_t* grad_idx, uint32_t* block_start,
                             uint32_t* block_idx);

/* Assembles at substep start with displacement u (3 floats per particle, relative to the current positions)
 * and no contact: total energy (FP64 sum), gradient (3 per particle) and the blocks in BSR slot order (9 per
 * block, row-major). Uses h for the inertia term. Any output may be NULL. */
int drape_solver_debug_assemble(drape_solver* s, double h, const float* u, double* energy, float* grad,
                                float* blocks);
/* y = A x over the last assembled matrix, pinned rows/columns excluded (cpu_solver.cpp:147-159). */
int drape_solver_debug_spmv(drape_solver* s, const float* x, float* y);
/* The linear system of the last Newton iteration as the device holds it (cpu_solver.cpp:163-213): blocks in BSR
 * slot order (9 per block), gradient (3 per particle), block-Jacobi inverse (9 per particle, zero when pinned),
 * contact flags (bit 0 near, bit 1 active) and contact normals (4 per particle: normal, offset). Any output may be
 * NULL. */
int drape_solver_debug_system(drape_solver* s, float* blocks, float* grad, float* precond, uint32_t* contact_flags,
                              float* contact_normals);

/* BakeOptions (libs/avatar/include/drape/avatar/collision_bake.hpp:10-14). */
typedef struct drape_bake_options {
  double cell_size; /* 0.004 m */
  double padding;   /* 0.05 m */
  double band;      /* 0.02 m */
} drape_bake_options;

/* TriangleMesh (libs/core/include/drape/core/triangle_mesh.hpp:11-14): a closed surface, triangles
 * counter-clockwise seen from outside. */
typedef struct drape_body_mesh {
  const double* vertices; /* 3 per vertex */
  uint32_t n_vertices;
  const uint32_t* triangles; /* 3 per triangle */
  uint32_t n_triangles;
} drape_body_mesh;

/* bakeCollisionField (collision_bake.hpp:16-18, collision_bake.cpp:98-208): the grid of the field of the union
 * of `n` closed components — origin, cell_size, dims and band of `out` (values left NULL). options NULL = the
 * defaults. invalid_argument when a component is not a closed manifold (every undirected edge used by exactly two
 * triangles, in opposite directions). */
int drape_bake_field_grid(const drape_body_mesh* components, uint32_t n, const drape_bake_options* options,
                          drape_field_desc* out);
/* The field itself, baked on the device into `values` (dims[0] * dims[1] * dims[2] floats, x fastest): per
 * component the signed distance within the band (closest triangle, sign from the angle-weighted pseudonormal of
 * the closest feature), +-band beyond it (outside = connected to the grid boundary), clamped to +-band; the
 * minimum over components. FP64 arithmetic, stored as float. */
int drape_bake_collision_field(const drape_body_mesh* components, uint32_t n, const drape_bake_options* options,
                               float* values);

/* vertexNormals (libs/render/include/drape/render/normals.hpp:13-18; normals.cpp:5-24) on the device, on caller
 * arrays (host memory): positions 3 floats per vertex, triangles 3 indices each, weld_pairs 2 indices each (a
 * seam's two coincident particles). Each vertex sums the cross products (x1 - x0) x (x2 - x0) of its triangles
 * in triangle order, each weld pair in order sets both of its vertices to their sum, and every normal is
 * normalised; a vertex without area gets (0, 1, 0). FP32, products and sums not fused. out: 3 floats per vertex.
 * std::out_of_range for an index >= n. */
int drape_vertex_normals(const float* positions, uint32_t n, const uint32_t* triangles, uint32_t n_triangles,
                         const uint32_t* weld_pairs, uint32_t n_welds, float* out);

/* drape_vertex_normals without any allocation. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). positions (3
 * floats per vertex), triangles, weld_pairs and out are DEVICE pointers the caller owns; stream is a hipStream_t
 * (NULL: the null stream), synchronised before the call returns. The working memory (incident-triangle lists and
 * per-vertex sums) is a fixed device array of the library, about 11.5 MB per GPU, held until the process ends:
 * at most 262,144 vertices and 524,288 triangles (std::invalid_argument beyond). Every call on a GPU shares that
 * array: calls that overlap on one GPU race. Same results as drape_vertex_normals, bit for bit; std::out_of_range
 * for an index >= n (out is then undefined). */
int drape_vertex_normals_noalloc(const float* positions, uint32_t n, const uint32_t* triangles, uint32_t n_triangles,
                                 const uint32_t* weld_pairs, uint32_t n_welds, float* out, void* stream);

/* The solver's current positions and their vertex normals (as drape_vertex_normals over every garment's mesh
 * triangles, every seam pair welded), in particle order, in one call: the position exchange of a live viewer
 * (the reference's viewport reads positions, then Renderer::updateGarment computes normals on the CPU,
 * renderer.cpp:285-289). positions may be NULL; n must equal particleCount() (invalid_argument). */
int drape_solver_read_positions_normals(const drape_solver* s, float* positions, float* normals, size_t n);

/* Standalone block-CSR SpMV on caller arrays (host memory): copies in, runs the solver's SpMV kernel `reps`
 * times, copies y out, and reports the mean kernel time in microseconds (HIP events). pinned may be NULL. */
int drape_bsr_spmv(uint32_t n, uint32_t n_blocks, const uint32_t* row_start, const uint32_t* col_index,
                   const float* blocks /* 9 per block, row-major */, const uint8_t* pinned, const float* x,
                   float* y, int32_t reps, double* mean_us);

/* Batch term evaluation on the device (the FP32 term math of the step): n samples of one kind.
 * kind 0 stretch: x 9 floats, params [dmInv00, dmInv10, dmInv01, dmInv11, area], fabric per sample;
 * kind 1 hinge:   x 12 floats, params [K, restAngle];
 * kind 2 seam:    x 6 floats,  params [target, K];
 * kind 3 friction: x 3 floats (x - x0), params [nx, ny, nz, muLambda, eps].
 * Outputs: energy (n), grad (n * 3N), hess (n * 9 N^2, block (a,b) row-major at (a*N+b)*9). */
int drape_eval_terms(int32_t kind, uint32_t n, const float* x, const float* params, const drape_fabric* fabrics,
                     float limit_scale, float* energy, float* grad, float* hess);

/* Body field queries on the device: phi (n) and unit gradient (3n) at points p (3n), FP32
 * (CollisionField::sample/gradient, collision_field.cpp:8-41). */
int drape_field_query(const drape_field_desc* field, uint32_t n, const float* p, float* phi, float* normal);

/* ---- Learned warm start. No counterpart in the reference: these entry points replace its previous-velocity
 * Newton guess u = h v (cpu_solver.cpp:244) with u = y + h d, where y = h v + h^2 (g + f_ext / m) is the inertial
 * target of the step and d, the velocity residual of the step beyond free flight (its dv minus h (g + f_ext / m)),
 * is predicted by a per-vertex message-passing network over the mesh edge graph; the plain guess is kept whenever
 * the prediction's energy is higher. Network: 3 rounds, hidden width 64, BF16 MFMA layers with FP32
 * accumulation, trained on the GPU. */
typedef struct drape_net drape_net;     /* weights, Adam state and working buffers on the device */
typedef struct drape_graph drape_graph; /* a mesh edge graph on the device */

int drape_net_create(drape_net** out, uint64_t seed);
void drape_net_destroy(drape_net* net);
int drape_net_param_count(const drape_net* net, size_t* out);
int drape_net_get_params(const drape_net* net, float* out, size_t n);
int drape_net_set_params(drape_net* net, const float* in, size_t n);
int drape_net_reset_optimizer(drape_net* net);

/* Graph over n vertices: CSR neighbour lists (row_start n + 1 entries, sorted neighbours) and pattern-space
 * rest positions (2 per vertex). */
int drape_graph_create(drape_graph** out, uint32_t n, const uint32_t* row_start, const uint32_t* nbr,
                       const double* rest2);
void drape_graph_destroy(drape_graph* g);

/* Host arrays of n vertices: x, v, dv_prev, target, dv_out 3 floats each; phi_normal 4 (normal, SDF distance).
 * predict: the velocity residual d (m/s) per vertex, into dv_out. train_step: the mean squared error of
 * (prediction - target) / 0.5 m/s into *loss and, when grad
 * is not NULL, the parameter gradient; lr > 0 then takes one Adam step. */
int drape_net_predict(drape_net* net, const drape_graph* g, const float* x, const float* v, const float* dv_prev,
                      const float* phi_normal, float* dv_out);
int drape_net_train_step(drape_net* net, const drape_graph* g, const float* x, const float* v,
                         const float* dv_prev, const float* phi_normal, const float* target, float lr, double* loss,
                         float* grad);

/* Attach a network to a solver (NULL detaches; the caller keeps the network alive while attached). Each substep
 * then predicts d, evaluates the energy at both guesses and keeps the lower (drape_counters.warm_start_*). */
int drape_solver_set_warm_start(drape_solver* s, drape_net* net);
/* The network inputs of the solver's current state, particle order (any output may be NULL). After a step,
 * dv_prev holds the dv that step realised: the training target for the inputs read before it. */
int drape_solver_frame_inputs(drape_solver* s, float* x, float* v, float* dv_prev, float* phi_normal);
/* The solver's mesh edge graph (triangle edges) in particle order: size, then CSR arrays. */
int drape_solver_mesh_graph_size(const drape_solver* s, size_t* n_entries);
int drape_solver_mesh_graph(const drape_solver* s, uint32_t* row_start, uint32_t* nbr);

#ifdef __cplusplus
}
#endif

#endif /* DRAPE_HIP_H */
