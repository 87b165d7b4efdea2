/* CPU oracle of the PlanetSim implicit heat-transport path — C ABI.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg load liboracle.so, and only as the checker
 * (or the timed CPU baseline). The product path never calls it.
 *
 * The oracle restates the reference algorithm in plain serial C++ with the
 * reference's block structure and operation order (no contraction, no
 * fast-math), so that its outputs are bit-identical to the reference's:
 *   mesh           sim/planet/mesh/icosphere.cpp:200-419
 *   basis          sim/planet/coordinates/local_tangent_basis.cpp:8-29
 *   FV operators   sim/planet/operators/finite_volume.cpp:23-111
 *   graph          sim/planet/surface/heat_transport.cpp:433-452
 *   multigrid      sim/planet/surface/heat_transport.cpp:54-326,545-577
 *   BiCGSTAB       sim/planet/surface/heat_transport.cpp:820-897
 *   diffusion      sim/planet/surface/heat_transport.cpp:28-36,579-593
 *   agglomeration  sim/planet/surface/heat_transport.cpp:467-512
 *   group maps     sim/planet/surface/surface_energy.cpp:1285-1300,1306-1324
 *   block ILU(0)   sim/core/math/incomplete_lu.cpp:40-146
 *   tracer step    sim/planet/surface/heat_transport.cpp:901-1081
 *   advection-diffusion  sim/planet/surface/heat_transport.cpp:774-816,1083-1318
 *   RNG            sim/core/random/counter_rng.hpp:19-45
 * and the synthetic heat-step workload (this project's input generator).
 *
 * Every function returns 0 on success, non-zero on error (message from
 * orc_last_error()), unless it returns a handle (NULL on error).
 */
#ifndef PSIM_ORACLE_H
#define PSIM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

const char* orc_last_error(void);

uint64_t orc_mix_random_key(uint64_t value);
uint64_t orc_keyed_random_u64(uint64_t seed, uint32_t stream, int64_t tick, uint32_t cell,
                              uint32_t sample);
double orc_keyed_random_unit_double(uint64_t seed, uint32_t stream, int64_t tick, uint32_t cell,
                                    uint32_t sample);

/* Mesh + its transport graph. */
void* orc_mesh_build(uint32_t level, double radius);
void orc_mesh_free(void* mesh);
void orc_mesh_counts(const void* mesh, uint64_t* cells, uint64_t* edges, uint64_t* directed);
void orc_mesh_cells(const void* mesh, double* center, double* east, double* north, double* area,
                    uint32_t* edge_offset, uint8_t* edge_count);
void orc_mesh_cell_edges(const void* mesh, uint32_t* neighbor, uint32_t* edge,
                         float* normal_east, float* normal_north);
void orc_mesh_edges(const void* mesh, uint32_t* first, uint32_t* second, double* length,
                    double* distance);
void orc_graph(const void* mesh, double* area, uint64_t* offset, uint64_t* neighbour,
               double* weight, double* weight_sum);

/* FV operators (reference templates for float and double). */
int orc_gradient_f64(const void* mesh, const double* scalar, double* east, double* north,
                     uint64_t workers);
int orc_gradient_f32(const void* mesh, const float* scalar, float* east, float* north,
                     uint64_t workers);
int orc_divergence_f64(const void* mesh, const double* flux, double* out, uint64_t workers);
int orc_divergence_f32(const void* mesh, const float* flux, float* out, uint64_t workers);
int orc_laplacian_f64(const void* mesh, const double* scalar, double* out, uint64_t workers);
int orc_laplacian_f32(const void* mesh, const float* scalar, float* out, uint64_t workers);
int orc_edge_orientation(const void* mesh, uint32_t cell, uint32_t edge, double* out);
int orc_diffusion_source(const void* mesh, double conductance, const double* temperature,
                         double* out, uint64_t workers);

/* Multigrid hierarchy structure on the unit-conductance graph. */
void* orc_hierarchy_build(const void* mesh);
void orc_hierarchy_free(void* hierarchy);
uint64_t orc_hierarchy_depth(const void* hierarchy);
void orc_hierarchy_level_counts(const void* hierarchy, uint64_t j, uint64_t* size, uint64_t* nnz);
void orc_hierarchy_level(const void* hierarchy, uint64_t j, uint64_t* offset, uint64_t* column,
                         uint64_t* aggregate, uint64_t* target, double* diagonal, double* value);

/* GraphMultigrid: set_matrix + one symmetric Gauss–Seidel V-cycle. */
void* orc_mg_create(const void* mesh);
void orc_mg_free(void* mg);
int orc_mg_set_matrix(void* mg, const double* diagonal, uint64_t n, const double* off,
                      uint64_t nnz);
int orc_mg_precondition(const void* mg, const double* rhs, double* out, uint64_t n,
                        uint64_t workers);
/* Coarsest level of the last set_matrix: size and its dense Cholesky factor
 * (row-major lower triangle, as the reference stores it). */
uint64_t orc_mg_coarsest_size(const void* mg);
void orc_mg_coarsest_factor(const void* mg, double* factor);
/* Galerkin values of level j after set_matrix. */
void orc_mg_level_values(const void* mg, uint64_t j, double* diagonal, double* value);

/* Right-preconditioned BiCGSTAB (reference form) on diag + off, from x = 0. */
int orc_bicgstab(const void* mesh, const void* mg, const double* diagonal, const double* off,
                 const double* b, double* x, double tolerance, int max_iterations,
                 uint64_t workers, int* iterations);

/* solve_newton_system (heat_transport.cpp:328-429): PCG with the V-cycle on
 * the Newton matrix of slope (held cells: identity rows), on a fresh
 * hierarchy of the mesh. */
int orc_newton_system(const void* mesh, double conductance, const double* slope, const double* rhs,
                      double cg_tolerance, int max_cg_iterations, uint64_t workers, double* solution,
                      int* iterations);

/* solve_implicit_transport (heat_transport.cpp:595-772) with a C response
 * callback: response(user, source[n], mean[n] out, slope[n] out, n). */
typedef void (*orc_response_fn)(void* user, const double* source, double* mean, double* slope, uint64_t n);
typedef struct {
    int max_newton_iterations, max_line_search_halvings;
    double rounding_floor;
    int max_non_monotone_steps, pad;
    double max_non_monotone_growth, newton_tolerance, cg_relative_tolerance;
    int max_cg_iterations, pad2;
} orc_transport_settings;
typedef struct {
    int newton_iterations, cg_iterations;
    double consistency_residual, sum_W, absolute_sum_W, dissipation_W_K;
} orc_transport_result;
int orc_solve_implicit_transport(const void* mesh, double conductance, const double* floor, orc_response_fn response,
                                 void* user, const orc_transport_settings* settings, uint64_t workers,
                                 const double* latent, double* source, double* mean, orc_transport_result* result);

/* build_agglomeration (heat_transport.cpp:467-512): group_of_cell[cells]
 * and a new mesh handle (free with orc_mesh_free) for the agglomerated
 * transport graph: the mesh one level coarser (the mesh itself at level 0)
 * whose cell areas are the group areas. NULL on error. */
void* orc_agglomerate(const void* mesh, uint32_t* group_of_cell);
/* The cell-to-group restriction around the agglomerated transport
 * (surface_energy.cpp:1285-1300): mean[g] = sum_c A_c v_c / A_g over the
 * group's cells in cell order; and the groups' floor (:1306,1323-1324). */
int orc_group_mean(uint64_t cells, uint64_t groups, const uint32_t* group_of_cell, const double* cell_area,
                   const double* group_area, const double* cell_values, double* group_mean);
int orc_group_max(uint64_t cells, uint64_t groups, const uint32_t* group_of_cell, const double* cell_values,
                  double* group_max);

/* Block ILU(0) (incomplete_lu.cpp:40-146): rows [block_start[b],
 * block_start[b+1]) factored on their own, diagonal scaled by `scale`;
 * orc_ilu_solve: x <- (LU)^-1 x. NULL / non-zero on error (zero pivot). */
void* orc_ilu_create(uint64_t n, const uint64_t* offset, const uint64_t* neighbour, const double* diagonal,
                     const double* off, const uint64_t* block_start, uint64_t blocks, double scale);
void orc_ilu_free(void* ilu);
int orc_ilu_solve(const void* ilu, double* x);

/* solve_implicit_tracer (heat_transport.cpp:901-1081) on the mesh's graph:
 * conductance[E] (W/K), outflow[layers x E] (kg/s), layer_mass[n] (kg),
 * tracer[layers x n] (layer-major in and out). */
typedef struct {
    int iterations, pad;
    double relative_residual, clipped_kg, column_divergence_kg_s;
} orc_tracer_result;
int orc_tracer_step(const void* mesh, uint32_t layers, const double* conductance, const double* outflow,
                    double heat_capacity, const double* layer_mass, const double* tracer, double dt, double tol,
                    int max_iterations, uint64_t workers, double* out, orc_tracer_result* result);

/* advection_diffusion_source (heat_transport.cpp:774-816): eddy_out[n] and
 * advective_out[n] (W/m^2) from mean[n] and energy[layers x n];
 * eddy_latent NULL: none. */
int orc_advection_source(const void* mesh, uint32_t layers, const double* conductance, const double* outflow,
                         const double* eddy_latent, const double* mean, const double* energy, double* eddy_out,
                         double* advective_out);
/* solve_implicit_advection_diffusion (heat_transport.cpp:1083-1318) with a
 * C response: (user, source[n], mean[n], slope[n], energy[L n],
 * energy_slope[L n], n, L). initial NULL: start at max(0, floor). */
typedef void (*orc_advection_response_fn)(void* user, const double* source, double* mean, double* slope,
                                          double* energy, double* energy_slope, uint64_t n, uint32_t layers);
typedef struct {
    int newton_iterations, linear_iterations;
    double consistency_residual, sum_W, absolute_sum_W, eddy_absolute_sum_W, advective_absolute_sum_W;
} orc_advection_result;
int orc_solve_advection_diffusion(const void* mesh, uint32_t layers, const double* conductance, const double* outflow,
                                  const double* eddy_latent, const double* floor, orc_advection_response_fn response,
                                  void* user, const orc_transport_settings* settings, const double* initial,
                                  uint64_t workers, double* source_out, double* mean_out, double* energy_out,
                                  orc_advection_result* result);

/* Synthetic heat-step workload (the project's seeded input generator).
 *   instance_seed = orc_mix_random_key(base_seed ^ instance)
 *   conductance_m2_s[E], capacity_area[n] = C_i A_i, land[n] in {0,1},
 *   temperature[n] = T_old for step 0. k_base: 1e6 for the literal
 *   configurations; sigma: log-normal spread of K_e. */
int orc_heat_instance(const void* mesh, uint64_t instance_seed, double k_base, double sigma,
                      double* conductance_m2_s, double* capacity_area, uint8_t* land,
                      double* temperature);
/* S_i for a step: 100 cos(lat) + 10 z. */
int orc_heat_forcing(const void* mesh, uint64_t instance_seed, uint32_t step, double* forcing);
/* diag = C A + dt Σ G, off = −dt G, G_k = K_e(k) w_k. */
int orc_heat_matrix(const void* mesh, const double* conductance_m2_s, const double* capacity_area,
                    double dt, double* diagonal, double* off);
/* b = C A T + dt A S. */
int orc_heat_rhs(const void* mesh, const double* capacity_area, const double* temperature,
                 const double* forcing, double dt, double* rhs);
/* True relative residual ||b − A x|| / ||b|| (block-ordered dots). */
double orc_relative_residual(const void* mesh, const double* diagonal, const double* off,
                             const double* b, const double* x);

#ifdef __cplusplus
}
#endif

#endif
