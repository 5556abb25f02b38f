/* oracle.h — C entry points of the FP64 CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY. The oracle is this project's own double-precision restatement of the reference's
 * CpuSolver path (libs/sim/src/cpu_solver.cpp, sim_model.cpp, energies.cpp, stretch.hpp, bending.hpp/.cpp,
 * collision.cpp, libs/core/src/collision_field.cpp). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker or the timed CPU baseline — never as a product path.
 * The reference's own build cannot run here (Eigen 3.4 and FetchContent dependencies are absent), so no
 * oracle/_ref build exists.
 */
#ifndef DRAPE_ORACLE_H
#define DRAPE_ORACLE_H

#include "drape_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_solver oracle_solver;

const char* oracle_last_error(void);
int oracle_solver_create(oracle_solver** out);
void oracle_solver_destroy(oracle_solver* s);
int oracle_solver_build(oracle_solver* s, const drape_scene_desc* scene);
int oracle_solver_update_fabric(oracle_solver* s, const char* garment, uint32_t n, const drape_fabric* fabrics);
int oracle_solver_set_phase(oracle_solver* s, const char* garment, int32_t phase);
int oracle_solver_step(oracle_solver* s, double dt, int32_t substeps, int32_t iterations);
int oracle_solver_particle_count(const oracle_solver* s, size_t* out);
int oracle_solver_range(const oracle_solver* s, const char* garment, uint32_t* offset, uint32_t* count);
int oracle_solver_read_positions(const oracle_solver* s, float* out, size_t n);
int oracle_solver_read_velocities(const oracle_solver* s, float* out, size_t n);
int oracle_solver_write_state(oracle_solver* s, const float* x, const float* v, size_t n);
int oracle_solver_stats(const oracle_solver* s, drape_stats* out);
int oracle_solver_snapshot_size(const oracle_solver* s, size_t* bytes);
int oracle_solver_snapshot_save(const oracle_solver* s, void* out, size_t bytes);
int oracle_solver_snapshot_restore(oracle_solver* s, const void* in, size_t bytes);
int oracle_solver_counters(const oracle_solver* s, drape_counters* out);
/* Threads for the row-parallel parts (term evaluation, SpMV, vector updates). Results are bit-identical for any
 * thread count: every sum keeps the serial order. */
int oracle_set_threads(int32_t n);

int oracle_solver_structure_sizes(const oracle_solver* s, drape_structure_sizes* out);
int oracle_solver_read_structure(const oracle_solver* s, uint32_t* row_start, uint32_t* col_index,
                                 uint32_t* diag_slot, uint32_t* stretch_slots, uint32_t* hinge_slots,
                                 uint32_t* seam_slots);
int oracle_solver_read_terms(const oracle_solver* s, uint32_t* stretch_v, uint32_t* hinge_v, double* hinge_k,
                             uint32_t* seam_ab);
/* Element -> gradient/block gather lists of the reference GPU path (gpu_solver_core.cpp:184-224). */
int oracle_solver_gather_sizes(const oracle_solver* s, uint32_t* n_grad, uint32_t* n_hess);
int oracle_solver_read_gather(const oracle_solver* s, uint32_t* grad_start, uint32_t* grad_idx,
                              uint32_t* block_start, uint32_t* block_idx);
/* Assembly at xd = x + u with no contact (evaluate(..., true), cpu_solver.cpp:92-144): energy, gradient and
 * blocks (9 per slot, row-major). */
int oracle_solver_debug_assemble(oracle_solver* s, double h, const double* u, double* energy, double* grad,
                                 double* blocks);
/* out = A in over the last assembled blocks (multiply, cpu_solver.cpp:147-159). */
int oracle_solver_debug_spmv(oracle_solver* s, const double* in, double* out);

/* Term math (energies.cpp, stretch.hpp, bending.hpp/.cpp). Outputs as in drape_eval_terms, in double. */
int oracle_eval_stretch(const double* x9, const double* dm_inv4 /* (0,0),(1,0),(0,1),(1,1) */, double area,
                        const drape_fabric* f, double limit_scale, double* energy, double* grad9, double* hess81);
int oracle_eval_hinge(const double* x12, double K, double rest_angle, double* energy, double* grad12,
                      double* hess144);
int oracle_eval_seam(const double* x6, double target, double K, double* energy, double* grad6, double* hess36);
int oracle_eval_friction(const double* x3, const double* x0, const double* n3, double mu_lambda, double eps,
                         double* energy, double* grad3, double* hess9);
int oracle_strain(const double* x9, const double* dm_inv4, int32_t which, double* C, double* grad9);
int oracle_dihedral(const double* x12, double* theta, double* grad12);
int oracle_bending_coefficients(double bend_warp, double bend_weft, double* p, double* q);
double oracle_edge_bend_stiffness(double p, double q, const double* edge2, const double* grain2);

/* CollisionField::sample / gradient (collision_field.cpp:8-41) and the post-solve helpers (collision.cpp). */
int oracle_field_query(const drape_field_desc* field, uint32_t n, const double* p, double* phi, double* normal);
int oracle_collide_body(const drape_field_desc* field, float radius, float friction, const float* x_prev, float* x);
int oracle_land_on_body(const drape_field_desc* field, float radius, const float* from, const float* to,
                        float* out);

/* 3x3 symmetric inverse used for the block-Jacobi preconditioner: ok = 0 when the block is not positive
 * definite (the reference's Eigen LDLT with identity fallback, cpu_solver.cpp:171-178). */
int oracle_invert_spd3(const double* a9, double* inv9);

/* Standalone PCG on a caller BSR system (solveLinear, cpu_solver.cpp:169-213) with no contact. */
int oracle_bsr_pcg(uint32_t n, const uint32_t* row_start, const uint32_t* col_index, const double* blocks,
                   const uint8_t* pinned, const double* rhs, double tol, int32_t max_iter, double* x,
                   int32_t* iterations);
/* The same with held contacts: rows with active[i] != 0 solve in the plane normal to normals[3i..3i+2]
 * (filterContact, cpu_solver.cpp:163-167). contact_force (may be NULL) receives each held row's normal force
 * from the unfiltered residual (cpu_solver.cpp:210-212). */
int oracle_bsr_pcg_contact(uint32_t n, const uint32_t* row_start, const uint32_t* col_index, const double* blocks,
                           const uint8_t* pinned, const uint8_t* active, const double* normals, const double* rhs,
                           double tol, int32_t max_iter, double* x, int32_t* iterations, double* contact_force);

/* bakeCollisionField (libs/avatar/src/collision_bake.cpp:98-208), sequentially in FP64 as the reference: triangles
 * in order with the bounding-box reject and strict-less updates, pseudonormal signs, breadth-first flood fill from
 * the grid boundary, minimum over components. options NULL = defaults. oracle_bake_field_grid fills the grid of
 * `out` (values NULL); oracle_bake_collision_field fills dims product floats. */
int oracle_bake_field_grid(const drape_body_mesh* components, uint32_t n, const drape_bake_options* options,
                           drape_field_desc* out);
int oracle_bake_collision_field(const drape_body_mesh* components, uint32_t n, const drape_bake_options* options,
                                float* values);

/* vertexNormals (libs/render/src/normals.cpp:5-24), sequentially in FP32: per triangle in order add
 * c = (x1 - x0) x (x2 - x0) to its three vertices; per weld pair in order set both vertices to their sum; then
 * normalise each, (0, 1, 0) for a zero vector. */
int oracle_vertex_normals(const float* positions, uint32_t n, const uint32_t* triangles, uint32_t n_triangles,
                          const uint32_t* weld_pairs, uint32_t n_welds, float* out);

#ifdef __cplusplus
}
#endif

#endif /* DRAPE_ORACLE_H */
