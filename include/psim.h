/* libpsim — MI355X (gfx950) implementation of PlanetSim's implicit
 * diffusive heat-transport path, C ABI.
 *
 * This is the boundary the C++20 drop-in headers under include/planetsim/
 * call, and the one a ctypes/cffi/FFI binding would bind. Plain pointers
 * and sizes only. Every function returns PSIM_OK (0) or an error code, and
 * leaves a message for psim_last_error() (thread-local) on failure, unless
 * it is documented otherwise.
 *
 * Memory: functions that take a `where` argument read and write host
 * memory (PSIM_HOST: synchronous, the call returns when outputs are
 * written) or device memory (PSIM_DEVICE: pointers from psim_device_alloc,
 * work is queued on the context's stream; psim_ctx_synchronize waits).
 * Device state (meshes, hierarchies, matrices, vectors) is owned by the
 * library and lives in HBM until its handle is destroyed.
 *
 * Reference interfaces each entry point replaces are cited as
 * file:line of fgianni/planet_game.
 */
#ifndef PSIM_H
#define PSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum psim_status {
    PSIM_OK = 0,
    PSIM_INVALID_ARGUMENT = 1, /* std::invalid_argument in the reference */
    PSIM_LOGIC_ERROR = 2,      /* std::logic_error (internal invariant) */
    PSIM_DEVICE_ERROR = 3,     /* HIP runtime error or no device */
    PSIM_OUT_OF_MEMORY = 4,    /* device allocation / admission refused */
    PSIM_NOT_CONVERGED = 5,    /* solver stopped above its tolerance */
    PSIM_RUNTIME_ERROR = 6     /* std::runtime_error (numerical breakdown) */
};

enum psim_where { PSIM_HOST = 0, PSIM_DEVICE = 1 };
enum psim_dtype { PSIM_F64 = 0, PSIM_F32 = 1 };
/* Solver precision: PSIM_PLAIN_FP64 keeps everything in FP64;
 * PSIM_MIXED keeps the BiCGSTAB vectors, inner products, true residual and
 * coarsest solve in FP64 and runs the V-cycle after its first level-0
 * Jacobi sweep in FP32 (every level's operator is also stored in FP32): the
 * FP64 first sweep carries the part of the preconditioned residual that
 * FP32 could not represent to 1e-10. */
enum psim_precision { PSIM_PLAIN_FP64 = 0, PSIM_MIXED = 1 };

const char* psim_last_error(void);
int psim_version(void);

/* ------------------------------------------------------------- host mesh
 * make_icosphere (sim/planet/mesh/planet_mesh.hpp:15, icosphere.cpp:200-419),
 * bit-identical; PlanetMesh accessors (planet_mesh.hpp:84-125). Host only. */
typedef struct psim_mesh psim_mesh;
int psim_mesh_create(uint32_t subdivision, double radius_m, psim_mesh** out);
void psim_mesh_destroy(psim_mesh* mesh);
int psim_mesh_counts(const psim_mesh* mesh, uint64_t* cells, uint64_t* edges, uint64_t* directed_edges);
/* center/east/north: 3 doubles (x, y, z) per cell. */
int psim_mesh_cells(const psim_mesh* mesh, double* center, double* east, double* north, double* area_m2,
                    uint32_t* edge_offset, uint8_t* edge_count);
int psim_mesh_cell_edges(const psim_mesh* mesh, uint32_t* neighbor, uint32_t* edge, float* outward_normal_east,
                         float* outward_normal_north);
int psim_mesh_edges(const psim_mesh* mesh, uint32_t* first_cell, uint32_t* second_cell, double* length_m,
                    double* centroid_distance_m);
/* mesh_transport_graph (heat_transport.cpp:433-452): area, offset[n+1],
 * neighbour[nnz], weight = l/d, weight_sum. */
int psim_mesh_graph(const psim_mesh* mesh, double* area_m2, uint64_t* offset, uint64_t* neighbour, double* weight,
                    double* weight_sum);

/* ------------------------------------------------------ agglomeration
 * build_agglomeration (sim/planet/surface/heat_transport.cpp:467-512),
 * bit-identical: each cell joins the nearest cell of the mesh one level
 * coarser (greedy walk from the previous cell's group); a group's area is
 * its cells' areas summed in cell order. The handle replaces the process
 * cache behind agglomerated_transport_graph and agglomerated_mesh
 * (heat_transport.hpp:45-50, heat_transport.cpp:520-543):
 *   psim_agglomeration_mesh      the mesh one level coarser (at subdivision
 *                                0 the mesh itself), owned by the handle;
 *   the agglomerated graph       psim_mesh_graph of that mesh with its
 *                                area replaced by group_area_m2.
 * PSIM_LOGIC_ERROR when a coarse cell receives no fine cell (:505-509). */
typedef struct psim_agglomeration psim_agglomeration;
int psim_agglomeration_create(const psim_mesh* mesh, psim_agglomeration** out);
void psim_agglomeration_destroy(psim_agglomeration* agglomeration);
int psim_agglomeration_counts(const psim_agglomeration* agglomeration, uint64_t* cells, uint64_t* groups);
/* group_of_cell[cells], group_area_m2[groups]; either may be NULL. */
int psim_agglomeration_groups(const psim_agglomeration* agglomeration, uint32_t* group_of_cell,
                              double* group_area_m2);
const psim_mesh* psim_agglomeration_mesh(const psim_agglomeration* agglomeration);

/* --------------------------------------------------- hierarchy (host view)
 * build_hierarchy (heat_transport.cpp:175-200): the aggregation hierarchy's
 * structure and unit-conductance Galerkin values, bit-identical. Exposed for
 * parity checks; the device copy is built by psim_graph_create. */
typedef struct psim_hierarchy psim_hierarchy;
int psim_hierarchy_create(uint64_t n, const uint64_t* offset, const uint64_t* neighbour, const double* weight,
                          const double* weight_sum, psim_hierarchy** out);
void psim_hierarchy_destroy(psim_hierarchy* h);
uint64_t psim_hierarchy_depth(const psim_hierarchy* h);
int psim_hierarchy_level_counts(const psim_hierarchy* h, uint64_t level, uint64_t* size, uint64_t* nnz);
/* aggregate[size] and target[nnz] (all-ones = coarse diagonal) are left
 * untouched on the coarsest level. */
int psim_hierarchy_level(const psim_hierarchy* h, uint64_t level, uint64_t* offset, uint64_t* column,
                         uint64_t* aggregate, uint64_t* target, double* diagonal, double* value);

/* ------------------------------------------------------------- workload
 * The seeded heat-step inputs of the benchmark configurations (no
 * reference counterpart; draws use counter_rng.hpp:19-45). */
uint64_t psim_instance_seed(uint64_t base_seed, uint64_t index);
int psim_heat_instance(const psim_mesh* mesh, uint64_t seed, double k_base, double sigma,
                       double* conductance_per_edge, double* capacity_area, uint8_t* land, double* temperature);
int psim_heat_forcing(const psim_mesh* mesh, uint64_t seed, uint32_t step, double* forcing);

/* ------------------------------------------------------------- device */
typedef struct psim_ctx psim_ctx;
/* One context per GPU: a HIP stream (default priority) and a workspace.
 * Fails with PSIM_DEVICE_ERROR where no GPU is visible. */
int psim_ctx_create(int device, psim_ctx** out);
void psim_ctx_destroy(psim_ctx* ctx);
int psim_ctx_synchronize(psim_ctx* ctx);
typedef struct psim_device_info {
    int32_t compute_units;
    int32_t program_blocks_per_cu;   /* co-resident solver workgroups per CU */
    int32_t cooperative;             /* device-wide groups supported */
    int32_t reserved;
    uint64_t lds_per_block;
    uint64_t lds_per_cu;
    uint64_t hbm_bytes;
    double wall_clock_hz;
} psim_device_info;
int psim_ctx_info(const psim_ctx* ctx, psim_device_info* info);
int psim_device_alloc(psim_ctx* ctx, uint64_t bytes, void** out);
int psim_device_free(psim_ctx* ctx, void* ptr);
int psim_copy_to_device(psim_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int psim_copy_to_host(psim_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* Device-to-device copy through a plain streaming kernel, queued on the
 * context's stream (no counterpart in the reference: the kernel of known
 * traffic that HBM byte counters are calibrated against). 16-byte aligned
 * pointers and size. */
int psim_device_copy(psim_ctx* ctx, void* dst, const void* src, uint64_t bytes);
/* HIP events on the context's stream, for timing device work in place:
 * psim_ctx_events makes `count` events, psim_ctx_record records event
 * `index` on the stream, psim_ctx_elapsed waits for event `to` and returns
 * the milliseconds between `from` and `to`. */
int psim_ctx_events(psim_ctx* ctx, uint32_t count);
int psim_ctx_record(psim_ctx* ctx, uint32_t index);
int psim_ctx_elapsed(psim_ctx* ctx, uint32_t from, uint32_t to, float* ms);


/* ------------------------------------------------- FV operators on device
 * finite_volume.hpp:21-44 / finite_volume.cpp:23-123. The mesh is uploaded
 * once; each call sweeps `fields` fields stored one after another (cell
 * fields: fields x cells, edge fields: fields x edges). Results are
 * bit-identical to the reference for both dtypes. */
typedef struct psim_fv psim_fv;
int psim_fv_create(psim_ctx* ctx, const psim_mesh* mesh, psim_fv** out);
/* The same from a PlanetMesh's arrays (the drop-in wrappers' route). */
int psim_fv_create_from_arrays(psim_ctx* ctx, uint64_t cells, uint64_t edges, const double* area_m2,
                               const uint32_t* edge_offset, const uint8_t* edge_count, const uint32_t* neighbor,
                               const uint32_t* edge, const float* outward_normal_east,
                               const float* outward_normal_north, const uint32_t* first_cell,
                               const uint32_t* second_cell, const double* length_m,
                               const double* centroid_distance_m, psim_fv** out);
void psim_fv_destroy(psim_fv* fv);
/* edge_orientation (finite_volume.cpp:23-32): +1, -1, or
 * PSIM_INVALID_ARGUMENT "cell is not incident to edge". */
int psim_edge_orientation(const psim_fv* fv, uint32_t cell, uint32_t edge, double* out);
int psim_gradient(psim_fv* fv, int dtype, uint64_t fields, const void* scalar, void* east, void* north, int where);
int psim_divergence(psim_fv* fv, int dtype, uint64_t fields, const void* edge_flux, void* result, int where);
int psim_laplacian(psim_fv* fv, int dtype, uint64_t fields, const void* scalar, void* result, int where);

/* ------------------------------------------- transport graph on device
 * TransportGraph (heat_transport.hpp:27-36) plus the multigrid hierarchy
 * structure built from it (GraphMultigrid's constructor,
 * heat_transport.cpp:550-553). Shared by every system on the graph.
 * offset[0..n]: CSR row starts, offset[0] == 0, never decreasing, fewer
 * than 2^31 entries; any row width (empty rows included). neighbour and
 * weight may be NULL when the graph has no entries. Otherwise
 * PSIM_INVALID_ARGUMENT. (psim_hierarchy_create checks its graph alike.) */
typedef struct psim_graph psim_graph;
int psim_graph_create(psim_ctx* ctx, uint64_t n, const double* area_m2, const uint64_t* offset,
                      const uint64_t* neighbour, const double* weight, const double* weight_sum, psim_graph** out);
void psim_graph_destroy(psim_graph* graph);
uint64_t psim_graph_size(const psim_graph* graph);
uint64_t psim_graph_entries(const psim_graph* graph);
uint64_t psim_graph_depth(const psim_graph* graph);
/* diffusion_source (heat_transport.cpp:579-593), bit-identical. */
int psim_diffusion_source(psim_graph* graph, double conductance_W_K, const double* temperature_K,
                          double* source_W_m2, int where);

/* ------------------------------------------------- groups on the device
 * The cell-to-group maps around a transport solve on the agglomerated graph
 * (sim/planet/surface/surface_energy.cpp:1226-1247,1285-1300,1306-1338; the
 * reference writes them as loops inline around its response). A group's
 * cells are visited in cell order. Fields are stored one after another
 * (fields x cells in, fields x groups out, or the reverse), 1..65535. */
typedef struct psim_groups psim_groups;
/* Host arrays: group_of_cell[cells] (each < groups, every group non-empty),
 * cell_area_m2[cells], group_area_m2[groups] (the agglomeration's). */
int psim_groups_create(psim_ctx* ctx, uint64_t cells, uint64_t groups, const uint32_t* group_of_cell,
                       const double* cell_area_m2, const double* group_area_m2, psim_groups** out);
void psim_groups_destroy(psim_groups* groups);
/* Area-weighted mean, bit-identical: group[g] = (sum over the group's cells
 * of A_c v_c, from 0) / A_g (surface_energy.cpp:1285-1300). */
int psim_groups_mean(psim_groups* groups, uint32_t fields, const double* cell_values, double* group_values, int where);
/* Highest value, from -inf with std::max in cell order (NaN cells leave the
 * running value as std::max does) (surface_energy.cpp:1306,1323-1324). */
int psim_groups_max(psim_groups* groups, uint32_t fields, const double* cell_values, double* group_values, int where);
/* cell[c] = group[group_of_cell[c]] (surface_energy.cpp:1247,1369). */
int psim_groups_broadcast(psim_groups* groups, uint32_t fields, const double* group_values, double* cell_values,
                          int where);

/* ---------------------------------------------- linear system on a graph
 * GraphMultigrid::set_matrix / precondition (heat_transport.hpp:60-76) and
 * the right-preconditioned BiCGSTAB of heat_transport.cpp:820-897, run on
 * the device. The V-cycle smooths with Chebyshev-weighted Jacobi (not the
 * reference's Gauss–Seidel), so preconditioner outputs differ from the
 * reference's while solutions agree to the solver tolerance. */
typedef struct psim_system psim_system;
typedef struct psim_solve_info {
    int32_t iterations;         /* BiCGSTAB iterations, summed over restarts */
    int32_t restarts;           /* true-residual restarts taken */
    int32_t status;             /* PSIM_OK or PSIM_NOT_CONVERGED */
    int32_t v_cycles;           /* preconditioner applications (half-iterations) */
    double relative_residual;   /* true ||b - A x|| / ||b|| in FP64 */
    double device_ms;           /* the solve's duration on the device clock */
} psim_solve_info;

int psim_system_create(psim_graph* graph, int precision, psim_system** out);
void psim_system_destroy(psim_system* system);
/* Workgroups that run one solve (the group walks the tiles and meets at
 * device-wide barriers); default min(tiles, co-resident capacity). Results
 * are bit-identical for any value. No reference counterpart (the
 * reference's worker_count plays this role, deterministic_executor.hpp:14). */
int psim_system_set_workgroups(psim_system* system, int32_t workgroups);
int32_t psim_system_workgroups(const psim_system* system);
/* Pre-smoothing of the V-cycle inside BiCGSTAB solves: bit j of
 * level_mask set = level j takes two weighted-Jacobi sweeps before its
 * restriction (each sweep a barrier-separated phase), clear = one; -1 (the
 * default) = automatic: every level whose Jacobi weight is damped below 1
 * (Gershgorin ratio of D^-1 A above 0.6), plus level 0 in the FP64 form.
 * precondition and the PCG always use the symmetric one-sweep V(1,1). No
 * reference counterpart (the reference smooths with one symmetric
 * Gauss-Seidel sweep, heat_transport.cpp:252-326; the second Jacobi sweep
 * brings the BiCGSTAB iteration counts to the reference's). */
int psim_system_set_pre_sweeps(psim_system* system, int32_t level_mask);
int32_t psim_system_pre_sweeps(const psim_system* system);
/* Phase trace (diagnostics): with capacity > 0, every later launch on the
 * system stamps the device wall clock (100 MHz) when it starts and after
 * each device-wide barrier; psim_system_trace_read returns the stamps of
 * the last launch (synchronises). capacity 0 turns tracing off. */
int psim_system_trace(psim_system* system, int32_t capacity);
int psim_system_trace_read(psim_system* system, uint64_t* ticks, int32_t capacity, int32_t* count);
/* FP64 Galerkin values of a multigrid level after set_matrix (the
 * reference's Level::diagonal / Level::value, heat_transport.cpp:54-67). */
int psim_system_level_values(psim_system* system, uint64_t level, double* diagonal, double* off_diagonal);
/* diagonal[n], off_diagonal[nnz] in the graph's CSR order; throws the
 * reference's "multigrid matrix does not match the graph" on size
 * mismatch (heat_transport.cpp:562-564). */
int psim_system_set_matrix(psim_system* system, const double* diagonal, uint64_t n, const double* off_diagonal,
                           uint64_t nnz, int where);
/* One V-cycle: out ~= A^-1 rhs. */
int psim_system_precondition(psim_system* system, const double* rhs, double* out, int where);
/* out = A x (diag * x, then entries in CSR order: heat_transport.cpp:1204-1214). */
int psim_system_apply(psim_system* system, const double* x, double* out, int where);
/* x = A^-1 b from x = 0 to true relative residual <= tolerance. */
int psim_system_solve(psim_system* system, const double* b, double* x, double tolerance, int32_t max_iterations,
                      psim_solve_info* info, int where);

/* Preconditioned conjugate gradients with the V-cycle on the system's matrix
 * (plain FP64 systems), the reference's Newton-system solve
 * (heat_transport.cpp:328-429, solve_newton_system): from x = 0, stops when
 * ||b - A x|| <= tolerance ||b|| (recursive residual), after max_iterations,
 * or when (r, M r) <= 0. hold (n bytes, or NULL) marks rows held at 0: the
 * caller makes them decoupled identity rows of the matrix with b = 0 there,
 * and x stays exactly 0 on them. info->iterations counts CG iterations. */
int psim_system_pcg(psim_system* system, const double* b, const uint8_t* hold, double* x, double tolerance,
                    int32_t max_iterations, psim_solve_info* info, int where);
/* The same PCG where the operator's diagonal differs from the V-cycle's
 * matrix (off-diagonals shared): pressure_redistribution.cpp:48-58,150-178
 * solves the two-point Laplacian with the V-cycle of its regularised form
 * (diagonal x (1 + 1e-10)). operator_diagonal NULL: the system's own. */
int psim_system_pcg_operator(psim_system* system, const double* operator_diagonal, const double* b,
                             const uint8_t* hold, double* x, double tolerance, int32_t max_iterations,
                             psim_solve_info* info, int where);

/* ------------------------------------------------ implicit heat step
 * The benchmark's unit of work: (diag(C_i A_i) + dt L_K) T = C_i A_i T_old
 * + dt A_i S_i, with L_K the two-point Laplacian with entry weights
 * K_e l_e / d_e. In the reference's terms this is the linear-response form
 * of ADR-0009's Newton system (heat_transport.cpp:328-336) solved with the
 * BiCGSTAB + V-cycle of heat_transport.cpp:820-897.
 *
 * psim_heat_set_matrix assembles diag = C A + dt sum_k G_k, off_k = -dt G_k,
 * G_k = K_k w_k on the device and builds the multigrid values;
 * conductance_per_entry is K_e of each CSR entry's edge (nnz).
 * psim_heat_step assembles b = C A T_old + dt A S on the device, solves, and
 * writes T. */
/* Result record of the last solve on the system, once its work has
 * completed (synchronises the stream). */
int psim_system_last_info(psim_system* system, psim_solve_info* info);

int psim_heat_set_matrix(psim_system* system, const double* conductance_per_entry, const double* capacity_area,
                         double dt_s, int where);
int psim_heat_step(psim_system* system, const double* temperature_old, const double* forcing_W_m2, double dt_s,
                   double tolerance, int32_t max_iterations, double* temperature_new, psim_solve_info* info,
                   int where);

/* A chain of `steps` heat steps in one device launch (e.g. a year of
 * monthly steps): step k solves with forcings[k * n ..] and the previous step's T,
 * and the last step's T is written to temperature_new. When
 * conductance_per_entry and capacity_area are both given, the same launch
 * first assembles the matrix as psim_heat_set_matrix does. step_info, if
 * not NULL, receives one record per step (synchronises). */
int psim_heat_chain(psim_system* system, const double* conductance_per_entry, const double* capacity_area,
                    const double* temperature_old, const double* forcings, int32_t steps, double dt_s,
                    double tolerance, int32_t max_iterations, double* temperature_new, psim_solve_info* step_info,
                    int where);

/* psim_heat_step on device pointers with no allocation in the call: the
 * solver's workspace (every BiCGSTAB and V-cycle vector, the reduction
 * partials, the group barrier words, the result record, the instance
 * descriptor) lives in one static device array of 8 MiB per GPU, held until
 * the process ends, instead of the system's own buffers; the matrix and
 * hierarchy come from `system` (psim_heat_set_matrix first). Results are
 * bit-identical to psim_heat_step. Systems whose workspace exceeds 8 MiB are
 * rejected with PSIM_INVALID_ARGUMENT (FP64 systems up to 40,962 cells, L6,
 * fit). The array keeps the last call's contents and is shared by every call
 * on the GPU: calls that overlap on one GPU (other streams, other threads)
 * race. info, if not NULL, receives the result (synchronises).
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under stream capture, tight serving loops). */
int psim_heat_step_noalloc(psim_system* system, const double* temperature_old, const double* forcing_W_m2,
                           double dt_s, double tolerance, int32_t max_iterations, double* temperature_new,
                           psim_solve_info* info);

/* ---- The implicit transport (heat_transport.hpp:78-138,
 * solve_implicit_transport; heat_transport.cpp:595-772). Newton on
 * F(h) = h - K lap T(h) (+ latent) with the reference's backtracking line
 * search, kink handling and result sums; the response is the caller's
 * (host callback, called at every iterate and trial); the diffusion and the
 * Newton system's PCG run on the device. */
typedef void (*psim_response_fn)(void* user, const double* source_W_m2, double* mean_K, double* slope_K_m2_W,
                                 uint64_t n);

typedef struct {
    int32_t max_newton_iterations, max_line_search_halvings;
    double rounding_floor_W_m2;
    int32_t max_non_monotone_steps, pad;
    double max_non_monotone_growth, newton_tolerance_W_m2, cg_relative_tolerance;
    int32_t max_cg_iterations, pad2;
} psim_transport_settings;

typedef struct {
    int32_t newton_iterations, cg_iterations;
    double consistency_residual_W_m2, sum_W, absolute_sum_W, dissipation_W_K;
} psim_transport_result;

/* The reference's ImplicitTransportSettings defaults. */
void psim_transport_settings_default(psim_transport_settings* settings);
/* source_floor_W_m2, latent_W_m2 (or NULL), source_W_m2 and mean_K: host,
 * n = psim_graph_size. settings NULL: the defaults. */
int psim_implicit_transport(psim_graph* graph, double conductance_W_K, const double* source_floor_W_m2,
                            psim_response_fn response, void* user, const psim_transport_settings* settings,
                            const double* latent_W_m2, double* source_W_m2, double* mean_K,
                            psim_transport_result* result);

/* The same driver with a device-resident response, so that no field crosses
 * PCIe inside the Newton loop (the reference's response is always a
 * callback; this form has no counterpart there). Per node, a piecewise-
 * linear table of `points` (source, temperature) pairs, stored point-major
 * (pair k of node i at k n + i), sources increasing: T(h) follows the
 * segment whose left end is the last one <= h (the first segment below the
 * table, the last above it), and the slope is that segment's
 * (T1 - T0) / (h1 - h0); a flat segment holds the node (slope 0). A linear
 * response T0 + s h is a 2-point table. where = PSIM_DEVICE: the table,
 * floor, latent and outputs are device arrays; PSIM_HOST: host arrays. */
typedef struct {
    uint32_t points, reserved;
    const double* source_W_m2;   /* points x n */
    const double* mean_K;        /* points x n */
} psim_response_table;
int psim_implicit_transport_table(psim_graph* graph, double conductance_W_K, const double* source_floor_W_m2,
                                  const psim_response_table* table, const psim_transport_settings* settings,
                                  const double* latent_W_m2, double* source_W_m2, double* mean_K,
                                  psim_transport_result* result, int where);

/* ------------------------------------------------------- block ILU(0)
 * IncompleteLU (sim/core/math/incomplete_lu.hpp:14-43, .cpp:40-146), block
 * form, on the device: rows [block_start[b], block_start[b+1]) keep their
 * diagonal and their in-block off-diagonals of the CSR pattern
 * (offset[n+1], neighbour[offset[n]]; offsets checked as for
 * psim_graph_create) and are factored on their own.
 * Factors and solves are bit-identical to the reference's. One workgroup
 * per block walks its rows' level schedule. */
typedef struct psim_ilu psim_ilu;
int psim_ilu_create(psim_ctx* ctx, uint64_t n, const uint64_t* offset, const uint64_t* neighbour,
                    const uint64_t* block_start, uint64_t blocks, psim_ilu** out);
void psim_ilu_destroy(psim_ilu* ilu);
/* Factor diag x diagonal_scale + off (CSR order). PSIM_RUNTIME_ERROR
 * "incomplete LU: zero pivot" as the reference throws. */
int psim_ilu_factor(psim_ilu* ilu, const double* diagonal, const double* off, double diagonal_scale, int where);
/* x <- (LU)^-1 x. */
int psim_ilu_solve(psim_ilu* ilu, double* x, int where);

/* solve_implicit_tracer (heat_transport.hpp:210-239, heat_transport.cpp:
 * 901-1081): a tracer's backward-Euler step under layer mass fluxes and
 * eddy exchange on the graph, the stacked system solved on the device by
 * BiCGSTAB with block ILU(0) (blocks of 256 whole columns). Host arrays:
 * conductance_W_K[E], outflow_kg_s[layers x E] (per CSR entry, layer-
 * major), layer_mass_kg[n], tracer[layers x n] in, tracer_out[layers x n]. */
typedef struct {
    int32_t iterations, reserved;
    double relative_residual;        /* of the linear solve */
    double clipped_kg;               /* mass added where rounding left c' < 0 */
    double column_divergence_kg_s;   /* max over nodes of |sum_l sum_k F| */
} psim_tracer_result;
int psim_tracer_step(psim_graph* graph, uint32_t layers, const double* conductance_W_K, const double* outflow_kg_s,
                     double heat_capacity_J_kg_K, const double* layer_mass_kg, const double* tracer, double dt_s,
                     double relative_tolerance, int32_t max_iterations, double* tracer_out,
                     psim_tracer_result* result);

/* ----------------------------------------- operator solves (other consumers)
 * BiCGSTAB on an operator that differs from the multigrid's matrix: a
 * diagonal and off-diagonals in the graph's CSR order (host or device
 * arrays by `where`). Preconditioner (right): the V-cycle of `multigrid`
 * (a plain FP64 system of the same graph whose matrix was set, typically
 * the operator's symmetric part), its input multiplied by pre_scale and
 * its output divided by post_divide when given, then the block ILU(0) of
 * the operator (`ilu`, factored) applied to what the V-cycle leaves:
 *   out = M in;  out += ILU^-1 (in - A out).
 * Without a multigrid the preconditioner is the ILU(0) alone.
 *   PSIM_FORM_BALANCE    balanced_circulation.cpp:316-444: iteration cap
 *                        -> status PSIM_NOT_CONVERGED, breakdown ->
 *                        PSIM_RUNTIME_ERROR (the reference throws both);
 *   PSIM_FORM_TRANSPORT  heat_transport.cpp:820-897 (the advection-
 *                        diffusion Newton system, :1176-1226): stops at the
 *                        cap or a breakdown.
 * info->relative_residual is the true ||b - A x|| / ||b||. */
enum psim_solve_form { PSIM_FORM_TRANSPORT = 0, PSIM_FORM_BALANCE = 1 };
typedef struct {
    psim_system* multigrid;      /* or NULL */
    const double* pre_scale;     /* n, or NULL */
    const double* post_divide;   /* n, or NULL */
    psim_ilu* ilu;               /* or NULL */
} psim_preconditioner;
typedef struct {
    int32_t iterations, status;
    double relative_residual;
} psim_operator_info;
int psim_operator_solve(psim_graph* graph, const double* diagonal, const double* off,
                        const psim_preconditioner* preconditioner, int form, const double* b, double* x,
                        double relative_tolerance, int32_t max_iterations, psim_operator_info* info, int where);

/* ------------------------------------------- coupled advection-diffusion
 * AdvectionDiffusion (heat_transport.hpp:140-170): conductance_W_K[E]
 * (symmetric, >= 0), outflow_kg_s[layers x E] (antisymmetric, per CSR
 * entry, layer-major), eddy_latent_W_m2[n] or NULL.
 * advection_diffusion_source (:172-178, heat_transport.cpp:774-816),
 * bit-identical: eddy and advective sources (W/m^2) of mean_K[n] and
 * energy_J_kg[layers x n]. */
int psim_advection_source(psim_graph* graph, uint32_t layers, const double* conductance_W_K,
                          const double* outflow_kg_s, const double* eddy_latent_W_m2, const double* mean_K,
                          const double* energy_J_kg, double* eddy_W_m2, double* advective_W_m2, int where);
/* solve_implicit_advection_diffusion (heat_transport.hpp:241-250,
 * heat_transport.cpp:1083-1318) with a host callback for the columns'
 * response (AdvectionResponse, hpp:162-166): source[n] in; mean_K[n],
 * slope[n], energy[layers x n], energy_slope[layers x n] out. Host arrays;
 * initial_source_W_m2 NULL: start at max(0, floor). The Jacobian's block
 * ILU(0), the V-cycle of A / theta' + K and the BiCGSTAB run on the device. */
typedef void (*psim_advection_response_fn)(void* user, const double* source_W_m2, double* mean_K,
                                           double* slope_K_m2_W, double* energy_J_kg, double* energy_slope_J_kg_m2_W,
                                           uint64_t n, uint32_t layers);
typedef struct {
    int32_t newton_iterations, linear_iterations;
    double consistency_residual_W_m2, sum_W, absolute_sum_W, eddy_absolute_sum_W, advective_absolute_sum_W;
} psim_advection_result;
int psim_implicit_advection_diffusion(psim_graph* graph, uint32_t layers, const double* conductance_W_K,
                                      const double* outflow_kg_s, const double* eddy_latent_W_m2,
                                      const double* source_floor_W_m2, psim_advection_response_fn response,
                                      void* user, const psim_transport_settings* settings,
                                      const double* initial_source_W_m2, double* source_W_m2, double* mean_K,
                                      double* energy_J_kg, psim_advection_result* result);

/* ---- Batches (no counterpart in the reference, whose solves run one at a
 * time): many systems, of any graphs and precisions on one context, solved
 * in ONE launch of the persistent solver. Each system owns a group of
 * workgroups for the whole launch; its results are bit-identical to its own
 * psim_heat_step (they do not depend on the group size or the other
 * members). A batch holds at most as many systems as the device has
 * co-resident solver workgroups (psim_batch_size). Systems stay owned by
 * the caller and must outlive the batch; one system may belong to several
 * batches, whose calls are ordered on the context's stream. */
typedef struct psim_batch psim_batch;

int psim_batch_create(psim_ctx* ctx, psim_system* const* systems, int32_t count, psim_batch** out);
void psim_batch_destroy(psim_batch* batch);
/* Number of systems and co-resident workgroups (either output may be NULL). */
int psim_batch_size(const psim_batch* batch, int32_t* count, int32_t* capacity);
/* Workgroups per system (count entries). */
int psim_batch_groups(const psim_batch* batch, int32_t* groups);
/* Set workgroups per system (each >= 1, sum <= capacity); NULL restores the
 * default allocation (tiles per workgroup levelled greedily). */
int psim_batch_set_groups(psim_batch* batch, const int32_t* groups);

/* One heat step of every system in one launch. All arrays hold `count`
 * DEVICE pointers. conductance_per_entry / capacity_area may be NULL, or
 * hold NULL for a system: that system keeps its matrix (it must have one);
 * otherwise its matrix is assembled first in the same launch. infos (host,
 * count records) synchronises; NULL queues the work and returns (up to four
 * calls may be in flight; psim_batch_synchronize waits for them). */
/* Program instance of the batch's launches: 3 (default) or 2 workgroups per
 * CU (register budgets of 168 or 256 VGPRs per thread); setting it goes back
 * to the default allocation of that occupancy's capacity. Results are
 * bit-identical at either. */
int psim_batch_set_occupancy(psim_batch* batch, int32_t workgroups_per_cu);
int32_t psim_batch_occupancy(const psim_batch* batch);
int psim_batch_heat_step(psim_batch* batch, const double* const* conductance_per_entry,
                         const double* const* capacity_area, const double* const* temperature_old,
                         const double* const* forcing_W_m2, double dt_s, double tolerance, int32_t max_iterations,
                         double* const* temperature_new, psim_solve_info* infos);
int psim_batch_synchronize(psim_batch* batch);

/* Autotuner: times candidate launch shapes on this exact step, `repeats`
 * launches each, and keeps the fastest (its mean launch time in best_ms):
 * the occupancy (2 or 3 workgroups per CU) and, for each, workgroup
 * allocations (shares of the device's co-resident workgroups and tile-count
 * weightings). Outputs are written as by psim_batch_heat_step. */
int psim_batch_autotune(psim_batch* batch, const double* const* conductance_per_entry,
                        const double* const* capacity_area, const double* const* temperature_old,
                        const double* const* forcing_W_m2, double dt_s, double tolerance, int32_t max_iterations,
                        double* const* temperature_new, int32_t repeats, double* best_ms);

/* ---- Serving (no counterpart in the reference): continuous batching of
 * heat-step requests on one persistent launch. The server splits the
 * device's solver workgroups into slots per graph (slot size from the
 * graph's tiles, slot count from the expected share of requests x tiles)
 * and keeps, per graph, a pool of systems sized by an HBM budget: the
 * graphs' hierarchies and every live session's matrix and solver state stay
 * resident. A session's first request assembles its matrix on a free
 * system of its graph (admission waits for one) unless a free system
 * already holds the matrix its key names; its later steps reuse it; its
 * last request frees it. Requests that arrive while others run start as
 * soon as a slot of their graph is free, and run exactly as they would
 * alone (the same result bits). Graphs must have a coarsest level of at
 * most 168 rows (all icosphere levels do). */
typedef struct psim_server psim_server;

typedef struct {
    double arrival_s;          /* arrival time, seconds after the run starts (non-decreasing) */
    int32_t graph;             /* index into the server's graphs */
    int32_t session;           /* requests of one session share a resident system, in order */
    int32_t first, last;       /* first: assemble from conductance/capacity; last: free the system */
    const double* conductance_per_entry;   /* device, first request only */
    const double* capacity_area;           /* device, first request only */
    const double* temperature_old;         /* device (or device-visible pinned host) memory */
    const double* forcing_W_m2;
    double* temperature_new;
    /* first request only: 0, or a key naming the matrix (the contents of
     * conductance_per_entry and capacity_area) that the caller guarantees
     * for every session it is given to. A session whose key equals the one
     * a free pooled system of the graph last assembled with the same dt
     * takes that system and skips the assembly (kept across runs). */
    uint64_t matrix_key;
} psim_serve_request;

typedef struct {
    double arrival_s, submitted_s, completed_s;   /* host clock, seconds after the run starts */
    double device_ms;                             /* the solve on the device clock */
    int32_t status, iterations, restarts, v_cycles;
    double relative_residual;
} psim_serve_record;

/* request_share: expected fraction of requests per graph (NULL: equal).
 * slot_sizes: workgroups per slot per graph (NULL or 0: about 12 tiles of
 * the graph per workgroup).
 * memory_budget_bytes: HBM for pooled systems, split by slots x bytes per
 * system (at least one system per slot, at most 4096 per graph; 0 = one
 * per slot). */
int psim_server_create(psim_ctx* ctx, psim_graph* const* graphs, const double* request_share,
                       const int32_t* slot_sizes, int32_t graph_count, int precision, uint64_t memory_budget_bytes,
                       psim_server** out);
/* The same with the server program's occupancy: 2 or 3 workgroups per CU
 * (register budgets of 256 or 168 VGPRs; 0 = 3, the default of
 * psim_server_create). */
int psim_server_create_occupancy(psim_ctx* ctx, psim_graph* const* graphs, const double* request_share,
                                 const int32_t* slot_sizes, int32_t graph_count, int precision,
                                 uint64_t memory_budget_bytes, int32_t workgroups_per_cu, psim_server** out);
void psim_server_destroy(psim_server* server);
/* Per graph: slots, workgroups per slot, pool systems, bytes per system
 * (any output may be NULL). */
int psim_server_plan(const psim_server* server, int32_t* slots, int32_t* slot_size, int32_t* pool,
                     uint64_t* system_bytes);
/* Serve `count` requests (sorted by arrival) in real time on one persistent
 * launch and fill one record per request; returns when all are done. */
int psim_server_run(psim_server* server, const psim_serve_request* requests, int32_t count, double dt_s,
                    double tolerance, int32_t max_iterations, psim_serve_record* records);

#ifdef __cplusplus
}
#endif

#endif /* PSIM_H */
