// C-ABI shim over the PlanetSim reference sources, compiled from
// /root/reference by oracle/Makefile into oracle/_ref/libref.so.
//
// Test infrastructure only: the golden-fixture generator and the tests load
// it (through ctypes) to pin the oracle's restatement against the real
// reference code. Nothing in the product links or loads it.
//
// heat_transport.cpp is included as a translation unit of this file (not
// compiled separately) so that its internal hierarchy builder and its
// BiCGSTAB template (anonymous namespace, heat_transport.cpp:54-326,820-897)
// can be driven directly. The reference source is not modified.

#include "sim/planet/surface/heat_transport.cpp"  // NOLINT: see header comment

#include "sim/core/math/incomplete_lu.hpp"
#include "sim/core/random/counter_rng.hpp"
#include "sim/planet/coordinates/local_tangent_basis.hpp"
#include "sim/planet/mesh/icosphere.hpp"
#include "sim/planet/operators/finite_volume.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <string>

namespace {

thread_local std::string g_error;

template <class F> int guarded(F&& body) {
    try {
        body();
        return 0;
    } catch (const std::exception& error) {
        g_error = error.what();
        return 1;
    } catch (...) {
        g_error = "unknown exception";
        return 1;
    }
}

struct RefMesh {
    planetsim::PlanetMesh mesh;
    planetsim::TransportGraph graph;
};

template <class T>
int divergence_impl(const void* handle, const T* flux, T* out, std::uint64_t workers) {
    return guarded([&] {
        const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
        planetsim::EdgeField<T> in(mesh.edge_count());
        for (std::size_t e = 0; e < mesh.edge_count(); ++e) {
            in[planetsim::EdgeId{static_cast<std::uint32_t>(e)}] = flux[e];
        }
        planetsim::Field2D<T> result(mesh.cell_count());
        planetsim::divergence(mesh, in, result, workers);
        std::memcpy(out, result.values().data(), mesh.cell_count() * sizeof(T));
    });
}

template <class T>
int laplacian_impl(const void* handle, const T* scalar, T* out, std::uint64_t workers) {
    return guarded([&] {
        const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
        const std::size_t n = mesh.cell_count();
        planetsim::Field2D<T> in(std::vector<T>(scalar, scalar + n));
        planetsim::Field2D<T> result(n);
        planetsim::laplacian(mesh, in, result, workers);
        std::memcpy(out, result.values().data(), n * sizeof(T));
    });
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_error.c_str(); }

std::uint64_t ref_mix_random_key(std::uint64_t value) { return planetsim::mix_random_key(value); }

std::uint64_t ref_keyed_random_u64(std::uint64_t seed, std::uint32_t stream, std::int64_t tick,
                                   std::uint32_t cell, std::uint32_t sample) {
    return planetsim::keyed_random_u64(seed, static_cast<planetsim::RandomStreamId>(stream), tick,
                                       cell, sample);
}

double ref_keyed_random_unit_double(std::uint64_t seed, std::uint32_t stream, std::int64_t tick,
                                    std::uint32_t cell, std::uint32_t sample) {
    return planetsim::keyed_random_unit_double(
        seed, static_cast<planetsim::RandomStreamId>(stream), tick, cell, sample);
}

void* ref_mesh_build(std::uint32_t level, double radius) {
    RefMesh* result = nullptr;
    const int status = guarded([&] {
        auto mesh = planetsim::make_icosphere(level, radius);
        auto graph = planetsim::mesh_transport_graph(mesh);
        result = new RefMesh{std::move(mesh), std::move(graph)};
    });
    return status == 0 ? result : nullptr;
}

void ref_mesh_free(void* handle) { delete static_cast<RefMesh*>(handle); }

void ref_mesh_counts(const void* handle, std::uint64_t* cells, std::uint64_t* edges,
                     std::uint64_t* directed) {
    const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
    *cells = mesh.cell_count();
    *edges = mesh.edge_count();
    *directed = mesh.directed_edge_count();
}

// center/east/north: 3 doubles per cell (x, y, z).
void ref_mesh_cells(const void* handle, double* center, double* east, double* north,
                    double* area, std::uint32_t* edge_offset, std::uint8_t* edge_count) {
    const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
    std::size_t index = 0;
    for (const auto& cell : mesh.cells()) {
        const planetsim::Vec3d* vectors[3] = {&cell.center_unit, &cell.east_unit,
                                              &cell.north_unit};
        double* outputs[3] = {center, east, north};
        for (int v = 0; v < 3; ++v) {
            outputs[v][3 * index + 0] = vectors[v]->x;
            outputs[v][3 * index + 1] = vectors[v]->y;
            outputs[v][3 * index + 2] = vectors[v]->z;
        }
        area[index] = cell.area_m2;
        edge_offset[index] = cell.edge_offset;
        edge_count[index] = cell.edge_count;
        ++index;
    }
}

void ref_mesh_cell_edges(const void* handle, std::uint32_t* neighbor, std::uint32_t* edge,
                         float* normal_east, float* normal_north) {
    const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
    std::size_t k = 0;
    for (const auto& cell : mesh.cells()) {
        for (const auto& cell_edge : mesh.cell_edges(cell.id)) {
            neighbor[k] = cell_edge.neighbor.value();
            edge[k] = cell_edge.edge.value();
            normal_east[k] = cell_edge.outward_normal_east;
            normal_north[k] = cell_edge.outward_normal_north;
            ++k;
        }
    }
}

void ref_mesh_edges(const void* handle, std::uint32_t* first, std::uint32_t* second,
                    double* length, double* distance) {
    const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
    std::size_t e = 0;
    for (const auto& edge : mesh.edges()) {
        first[e] = edge.first_cell.value();
        second[e] = edge.second_cell.value();
        length[e] = edge.length_m;
        distance[e] = edge.centroid_distance_m;
        ++e;
    }
}

void ref_graph(const void* handle, double* area, std::uint64_t* offset, std::uint64_t* neighbour,
               double* weight, double* weight_sum) {
    const auto& graph = static_cast<const RefMesh*>(handle)->graph;
    for (std::size_t i = 0; i < graph.size(); ++i) {
        area[i] = graph.area_m2[i];
        weight_sum[i] = graph.weight_sum[i];
    }
    for (std::size_t i = 0; i < graph.offset.size(); ++i) offset[i] = graph.offset[i];
    for (std::size_t k = 0; k < graph.neighbour.size(); ++k) {
        neighbour[k] = graph.neighbour[k];
        weight[k] = graph.weight[k];
    }
}

int ref_gradient_f64(const void* handle, const double* scalar, double* east, double* north,
                     std::uint64_t workers) {
    return guarded([&] {
        const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
        const std::size_t n = mesh.cell_count();
        planetsim::Field2D<double> in(std::vector<double>(scalar, scalar + n));
        planetsim::Field2D<double> e(n), no(n);
        planetsim::gradient(mesh, in, e, no, workers);
        std::memcpy(east, e.values().data(), n * sizeof(double));
        std::memcpy(north, no.values().data(), n * sizeof(double));
    });
}

int ref_gradient_f32(const void* handle, const float* scalar, float* east, float* north,
                     std::uint64_t workers) {
    return guarded([&] {
        const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
        const std::size_t n = mesh.cell_count();
        planetsim::Field2D<float> in(std::vector<float>(scalar, scalar + n));
        planetsim::Field2D<float> e(n), no(n);
        planetsim::gradient(mesh, in, e, no, workers);
        std::memcpy(east, e.values().data(), n * sizeof(float));
        std::memcpy(north, no.values().data(), n * sizeof(float));
    });
}

int ref_divergence_f64(const void* h, const double* flux, double* out, std::uint64_t w) {
    return divergence_impl(h, flux, out, w);
}
int ref_divergence_f32(const void* h, const float* flux, float* out, std::uint64_t w) {
    return divergence_impl(h, flux, out, w);
}

int ref_laplacian_f64(const void* h, const double* s, double* out, std::uint64_t w) {
    return laplacian_impl(h, s, out, w);
}
int ref_laplacian_f32(const void* h, const float* s, float* out, std::uint64_t w) {
    return laplacian_impl(h, s, out, w);
}

int ref_edge_orientation(const void* handle, std::uint32_t cell, std::uint32_t edge,
                         double* out) {
    return guarded([&] {
        const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
        *out = planetsim::edge_orientation(mesh, planetsim::CellId{cell}, planetsim::EdgeId{edge});
    });
}

int ref_diffusion_source(const void* handle, double conductance, const double* temperature,
                         double* out, std::uint64_t workers) {
    return guarded([&] {
        const auto& graph = static_cast<const RefMesh*>(handle)->graph;
        const std::size_t n = graph.size();
        planetsim::Field2D<double> t(std::vector<double>(temperature, temperature + n));
        planetsim::Field2D<double> source;
        planetsim::diffusion_source(graph, conductance, t, source, workers);
        std::memcpy(out, source.values().data(), n * sizeof(double));
    });
}

// ---- multigrid hierarchy structure (heat_transport.cpp:175-200) ----

struct RefHierarchy {
    std::vector<planetsim::Level> levels;
};

void* ref_hierarchy_build(const void* handle) {
    const auto& graph = static_cast<const RefMesh*>(handle)->graph;
    return new RefHierarchy{planetsim::build_hierarchy(graph, graph.size())};
}

void ref_hierarchy_free(void* h) { delete static_cast<RefHierarchy*>(h); }

std::uint64_t ref_hierarchy_depth(const void* h) {
    return static_cast<const RefHierarchy*>(h)->levels.size();
}

// size, off-diagonal count of level j.
void ref_hierarchy_level_counts(const void* h, std::uint64_t j, std::uint64_t* size,
                                std::uint64_t* nnz) {
    const auto& level = static_cast<const RefHierarchy*>(h)->levels[j];
    *size = level.size;
    *nnz = level.column.size();
}

// aggregate/target are size/nnz long on every level but the coarsest
// (zero there); diagonal/value are the unit-conductance Galerkin values.
void ref_hierarchy_level(const void* h, std::uint64_t j, std::uint64_t* offset,
                         std::uint64_t* column, std::uint64_t* aggregate, std::uint64_t* target,
                         double* diagonal, double* value) {
    const auto& level = static_cast<const RefHierarchy*>(h)->levels[j];
    for (std::size_t i = 0; i < level.offset.size(); ++i) offset[i] = level.offset[i];
    for (std::size_t k = 0; k < level.column.size(); ++k) {
        column[k] = level.column[k];
        value[k] = level.value[k];
    }
    for (std::size_t i = 0; i < level.size; ++i) diagonal[i] = level.diagonal[i];
    for (std::size_t i = 0; i < level.aggregate.size(); ++i) aggregate[i] = level.aggregate[i];
    for (std::size_t k = 0; k < level.target.size(); ++k) target[k] = level.target[k];
}

// ---- GraphMultigrid (heat_transport.hpp:60-76) ----

void* ref_mg_create(const void* handle) {
    planetsim::GraphMultigrid* result = nullptr;
    const int status = guarded([&] {
        result = new planetsim::GraphMultigrid(static_cast<const RefMesh*>(handle)->graph);
    });
    return status == 0 ? result : nullptr;
}

void ref_mg_free(void* mg) { delete static_cast<planetsim::GraphMultigrid*>(mg); }

int ref_mg_set_matrix(void* mg, const double* diagonal, std::uint64_t n, const double* off,
                      std::uint64_t nnz) {
    return guarded([&] {
        static_cast<planetsim::GraphMultigrid*>(mg)->set_matrix({diagonal, n}, {off, nnz});
    });
}

int ref_mg_precondition(const void* mg, const double* rhs, double* out, std::uint64_t n,
                        std::uint64_t workers) {
    return guarded([&] {
        std::vector<double> in(rhs, rhs + n);
        std::vector<double> solution;
        static_cast<const planetsim::GraphMultigrid*>(mg)->precondition(in, solution, workers);
        std::memcpy(out, solution.data(), n * sizeof(double));
    });
}

// The reference BiCGSTAB template (heat_transport.cpp:820-897) on the CSR
// operator diag + off (the `apply` form of heat_transport.cpp:1204-1214),
// preconditioned by GraphMultigrid::precondition (set_matrix must have run).
int ref_bicgstab(const void* handle, const void* mg, const double* diagonal, const double* off,
                 const double* b, double* x, double tolerance, int max_iterations,
                 std::uint64_t workers, int* iterations) {
    return guarded([&] {
        const auto& graph = static_cast<const RefMesh*>(handle)->graph;
        const auto& multigrid = *static_cast<const planetsim::GraphMultigrid*>(mg);
        const std::size_t n = graph.size();
        const auto apply = [&](const std::vector<double>& in, std::vector<double>& out) {
            planetsim::for_each_deterministic_block(
                std::span<const planetsim::CellBlock>(graph.blocks), workers,
                [&](std::size_t, const planetsim::CellBlock& block) {
                    for (std::size_t a = block.begin; a < block.end; ++a) {
                        double sum = diagonal[a] * in[a];
                        for (std::size_t k = graph.offset[a]; k < graph.offset[a + 1U]; ++k) {
                            sum += off[k] * in[graph.neighbour[k]];
                        }
                        out[a] = sum;
                    }
                });
        };
        const auto precondition = [&](const std::vector<double>& in, std::vector<double>& out) {
            multigrid.precondition(in, out, workers);
        };
        std::vector<double> rhs(b, b + n);
        std::vector<double> solution;
        *iterations = planetsim::bicgstab(graph, apply, precondition, rhs, solution, tolerance,
                                          max_iterations, workers);
        std::memcpy(x, solution.data(), n * sizeof(double));
    });
}

// The reference's public coupled solve (heat_transport.hpp:246-249) with no
// layers, the heat step's conductances G (per CSR entry) and a linear column
// response theta(h) = t0 + slope * h: an independent route to the solution
// of (diag(A / slope) + L_G) theta = A t0 / slope.
int ref_coupled_linear_solve(const void* handle, const double* conductance, const double* t0,
                             const double* slope, double newton_tolerance,
                             double linear_tolerance, double* theta, int* newton_iterations,
                             int* linear_iterations) {
    return guarded([&] {
        const auto& graph = static_cast<const RefMesh*>(handle)->graph;
        const std::size_t n = graph.size();
        planetsim::AdvectionDiffusion transport;
        transport.layers = 0;
        transport.conductance_W_K.assign(conductance, conductance + graph.neighbour.size());
        const planetsim::Field2D<double> floor(n, -1.0e30);
        const auto response = [&](const planetsim::Field2D<double>& h,
                                  planetsim::Field2D<double>& mean,
                                  planetsim::Field2D<double>& d_mean, std::vector<double>& energy,
                                  std::vector<double>& energy_slope) {
            mean = planetsim::Field2D<double>(n, 0.0);
            d_mean = planetsim::Field2D<double>(n, 0.0);
            energy.clear();
            energy_slope.clear();
            for (std::size_t a = 0; a < n; ++a) {
                mean[a] = t0[a] + slope[a] * h[a];
                d_mean[a] = slope[a];
            }
        };
        planetsim::ImplicitTransportSettings settings;
        settings.newton_tolerance_W_m2 = newton_tolerance;
        settings.cg_relative_tolerance = linear_tolerance;
        settings.max_cg_iterations = 10'000;
        const auto result = planetsim::solve_implicit_advection_diffusion(graph, transport, floor,
                                                                          response, settings, 1U);
        std::memcpy(theta, result.mean_K.values().data(), n * sizeof(double));
        *newton_iterations = result.newton_iterations;
        *linear_iterations = result.linear_iterations;
    });
}

// solve_newton_system (heat_transport.cpp:328-429, anonymous namespace) on
// a fresh hierarchy of the graph: PCG with the Gauss-Seidel V-cycle.
int ref_newton_system(const void* handle, double conductance, const double* slope, const double* rhs,
                      double cg_tolerance, int max_cg_iterations, std::uint64_t workers, double* solution,
                      int* iterations) {
    return guarded([&] {
        const auto& graph = static_cast<const RefMesh*>(handle)->graph;
        const std::size_t n = graph.size();
        auto levels = planetsim::build_hierarchy(graph, n);
        planetsim::ImplicitTransportSettings settings;
        settings.cg_relative_tolerance = cg_tolerance;
        settings.max_cg_iterations = max_cg_iterations;
        const planetsim::Field2D<double> s(std::vector<double>(slope, slope + n));
        const planetsim::Field2D<double> b(std::vector<double>(rhs, rhs + n));
        planetsim::Field2D<double> x;
        *iterations = planetsim::solve_newton_system(graph, conductance, s, b, levels, x, settings, workers);
        std::memcpy(solution, x.values().data(), n * sizeof(double));
    });
}

// The reference's public Newton transport solve (heat_transport.hpp:135)
// with a C response callback.
typedef void (*ref_response_fn)(void* user, const double* source, double* mean, double* slope, std::uint64_t n);
struct RefTransportSettings {
    int max_newton_iterations, max_line_search_halvings;
    double rounding_floor;
    int max_non_monotone_steps, pad;
    double max_non_monotone_growth, newton_tolerance, cg_relative_tolerance;
    int max_cg_iterations, pad2;
};
struct RefTransportResult {
    int newton_iterations, cg_iterations;
    double consistency_residual, sum_W, absolute_sum_W, dissipation_W_K;
};

int ref_solve_implicit_transport(const void* handle, double conductance, const double* floor,
                                 ref_response_fn response, void* user, const RefTransportSettings* st,
                                 std::uint64_t workers, const double* latent, double* source, double* mean,
                                 RefTransportResult* result) {
    return guarded([&] {
        const auto& graph = static_cast<const RefMesh*>(handle)->graph;
        const std::size_t n = graph.size();
        const planetsim::TransportResponse bridge = [&](const planetsim::Field2D<double>& h,
                                                        planetsim::Field2D<double>& m,
                                                        planetsim::Field2D<double>& d) {
            response(user, h.values().data(), m.values().data(), d.values().data(), n);
        };
        planetsim::ImplicitTransportSettings settings;
        settings.max_newton_iterations = st->max_newton_iterations;
        settings.max_line_search_halvings = st->max_line_search_halvings;
        settings.rounding_floor_W_m2 = st->rounding_floor;
        settings.max_non_monotone_steps = st->max_non_monotone_steps;
        settings.max_non_monotone_growth = st->max_non_monotone_growth;
        settings.newton_tolerance_W_m2 = st->newton_tolerance;
        settings.cg_relative_tolerance = st->cg_relative_tolerance;
        settings.max_cg_iterations = st->max_cg_iterations;
        const planetsim::Field2D<double> fl(std::vector<double>(floor, floor + n));
        std::vector<double> lat;
        if (latent != nullptr) lat.assign(latent, latent + n);
        const auto r = planetsim::solve_implicit_transport(graph, conductance, fl, bridge, settings, workers,
                                                           latent != nullptr ? &lat : nullptr);
        std::memcpy(source, r.source_W_m2.values().data(), n * sizeof(double));
        std::memcpy(mean, r.mean_K.values().data(), n * sizeof(double));
        result->newton_iterations = r.newton_iterations;
        result->cg_iterations = r.cg_iterations;
        result->consistency_residual = r.consistency_residual_W_m2;
        result->sum_W = r.sum_W;
        result->absolute_sum_W = r.absolute_sum_W;
        result->dissipation_W_K = r.dissipation_W_K;
    });
}

// agglomerated_transport_graph / agglomerated_mesh (heat_transport.cpp:
// 520-543) of the handle's mesh: the group count, then the group map, the
// graph's arrays and the coarse mesh's subdivision.
std::uint64_t ref_agglomeration_groups(const void* handle) {
    const std::vector<std::size_t>* map = nullptr;
    return planetsim::agglomerated_transport_graph(static_cast<const RefMesh*>(handle)->mesh, map).size();
}

int ref_agglomeration(const void* handle, std::uint64_t* group_of_cell, double* area, std::uint64_t* offset,
                      std::uint64_t* neighbour, double* weight, double* weight_sum, std::uint32_t* subdivision) {
    return guarded([&] {
        const auto& mesh = static_cast<const RefMesh*>(handle)->mesh;
        const std::vector<std::size_t>* map = nullptr;
        const auto& graph = planetsim::agglomerated_transport_graph(mesh, map);
        for (std::size_t c = 0; c < map->size(); ++c) group_of_cell[c] = (*map)[c];
        for (std::size_t i = 0; i < graph.size(); ++i) {
            area[i] = graph.area_m2[i];
            weight_sum[i] = graph.weight_sum[i];
        }
        for (std::size_t i = 0; i < graph.offset.size(); ++i) offset[i] = graph.offset[i];
        for (std::size_t k = 0; k < graph.neighbour.size(); ++k) {
            neighbour[k] = graph.neighbour[k];
            weight[k] = graph.weight[k];
        }
        *subdivision = planetsim::agglomerated_mesh(mesh).subdivision();
    });
}

// IncompleteLU (incomplete_lu.hpp:14-43), block form.
void* ref_ilu_create(std::uint64_t n, const std::uint64_t* offset, const std::uint64_t* neighbour,
                     const double* diagonal, const double* off, const std::uint64_t* block_start,
                     std::uint64_t blocks, double scale) {
    planetsim::IncompleteLU* out = nullptr;
    const int status = guarded([&] {
        const std::vector<std::size_t> o(offset, offset + n + 1), nb(neighbour, neighbour + offset[n]),
            bs(block_start, block_start + blocks + 1);
        out = new planetsim::IncompleteLU(o, nb, std::span<const double>(diagonal, n),
                                          std::span<const double>(off, offset[n]), bs, scale, 1U);
    });
    return status == 0 ? out : nullptr;
}

void ref_ilu_free(void* f) { delete static_cast<planetsim::IncompleteLU*>(f); }

int ref_ilu_solve(const void* f, double* x, std::uint64_t n) {
    return guarded([&] {
        std::vector<double> v(x, x + n);
        static_cast<const planetsim::IncompleteLU*>(f)->solve(v);
        std::copy(v.begin(), v.end(), x);
    });
}

// solve_implicit_tracer (heat_transport.hpp:227-239) on the mesh's graph.
int ref_tracer_step(const void* handle, std::uint32_t layers, const double* conductance, const double* outflow,
                    double heat_capacity, const double* layer_mass, const double* tracer, double dt, double tol,
                    int max_iterations, std::uint64_t workers, double* out, int* iterations,
                    double* relative_residual, double* clipped_kg, double* column_divergence_kg_s) {
    return guarded([&] {
        const auto& graph = static_cast<const RefMesh*>(handle)->graph;
        const std::size_t n = graph.size(), e = graph.neighbour.size();
        planetsim::AdvectionDiffusion transport;
        transport.layers = layers;
        transport.conductance_W_K.assign(conductance, conductance + e);
        transport.outflow_kg_s.assign(outflow, outflow + layers * e);
        transport.heat_capacity_J_kg_K = heat_capacity;
        const std::vector<double> c(tracer, tracer + layers * n);
        const auto r = planetsim::solve_implicit_tracer(graph, transport, std::span<const double>(layer_mass, n), c,
                                                        dt, tol, max_iterations, workers);
        std::copy(r.tracer.begin(), r.tracer.end(), out);
        *iterations = r.iterations;
        *relative_residual = r.relative_residual;
        *clipped_kg = r.clipped_kg;
        *column_divergence_kg_s = r.column_divergence_kg_s;
    });
}

// advection_diffusion_source and solve_implicit_advection_diffusion
// (heat_transport.hpp:140-200) on the mesh's graph.
int ref_advection_source(const void* handle, std::uint32_t layers, const double* conductance, const double* outflow,
                         const double* eddy_latent, const double* mean, const double* energy, double* eddy_out,
                         double* advective_out) {
    return guarded([&] {
        const auto& graph = static_cast<const RefMesh*>(handle)->graph;
        const std::size_t n = graph.size(), e = graph.neighbour.size();
        planetsim::AdvectionDiffusion transport;
        transport.layers = layers;
        transport.conductance_W_K.assign(conductance, conductance + e);
        transport.outflow_kg_s.assign(outflow, outflow + layers * e);
        if (eddy_latent != nullptr) transport.eddy_latent_W_m2.assign(eddy_latent, eddy_latent + n);
        const planetsim::Field2D<double> m(std::vector<double>(mean, mean + n));
        const std::vector<double> s(energy, energy + layers * n);
        planetsim::Field2D<double> eddy, advective;
        planetsim::advection_diffusion_source(graph, transport, m, s, eddy, advective, 1U);
        std::memcpy(eddy_out, eddy.values().data(), n * sizeof(double));
        std::memcpy(advective_out, advective.values().data(), n * sizeof(double));
    });
}

typedef void (*ref_advection_response_fn)(void* user, const double* source, double* mean, double* slope,
                                          double* energy, double* energy_slope, std::uint64_t n,
                                          std::uint32_t layers);
struct RefAdvectionResult {
    int newton_iterations, linear_iterations;
    double consistency_residual, sum_W, absolute_sum_W, eddy_absolute_sum_W, advective_absolute_sum_W;
};

int ref_solve_advection_diffusion(const void* handle, std::uint32_t layers, const double* conductance,
                                  const double* outflow, const double* eddy_latent, const double* floor,
                                  ref_advection_response_fn response, void* user, const RefTransportSettings* st,
                                  const double* initial, std::uint64_t workers, double* source_out, double* mean_out,
                                  double* energy_out, RefAdvectionResult* result) {
    return guarded([&] {
        const auto& graph = static_cast<const RefMesh*>(handle)->graph;
        const std::size_t n = graph.size(), e = graph.neighbour.size();
        planetsim::AdvectionDiffusion transport;
        transport.layers = layers;
        transport.conductance_W_K.assign(conductance, conductance + e);
        transport.outflow_kg_s.assign(outflow, outflow + layers * e);
        if (eddy_latent != nullptr) transport.eddy_latent_W_m2.assign(eddy_latent, eddy_latent + n);
        const planetsim::AdvectionResponse bridge =
            [&](const planetsim::Field2D<double>& h, planetsim::Field2D<double>& m, planetsim::Field2D<double>& d,
                std::vector<double>& s, std::vector<double>& ds) {
                m = planetsim::Field2D<double>(n, 0.0);
                d = planetsim::Field2D<double>(n, 0.0);
                s.assign(layers * n, 0.0);
                ds.assign(layers * n, 0.0);
                response(user, h.values().data(), m.values().data(), d.values().data(), s.data(), ds.data(), n,
                         layers);
            };
        planetsim::ImplicitTransportSettings settings;
        settings.max_newton_iterations = st->max_newton_iterations;
        settings.max_line_search_halvings = st->max_line_search_halvings;
        settings.rounding_floor_W_m2 = st->rounding_floor;
        settings.max_non_monotone_steps = st->max_non_monotone_steps;
        settings.max_non_monotone_growth = st->max_non_monotone_growth;
        settings.newton_tolerance_W_m2 = st->newton_tolerance;
        settings.cg_relative_tolerance = st->cg_relative_tolerance;
        settings.max_cg_iterations = st->max_cg_iterations;
        planetsim::Field2D<double> start;
        if (initial != nullptr) {
            start = planetsim::Field2D<double>(std::vector<double>(initial, initial + n));
            settings.initial_source_W_m2 = &start;
        }
        const planetsim::Field2D<double> fl(std::vector<double>(floor, floor + n));
        const auto r = planetsim::solve_implicit_advection_diffusion(graph, transport, fl, bridge, settings, workers);
        std::memcpy(source_out, r.source_W_m2.values().data(), n * sizeof(double));
        std::memcpy(mean_out, r.mean_K.values().data(), n * sizeof(double));
        std::memcpy(energy_out, r.energy_J_kg.data(), layers * n * sizeof(double));
        result->newton_iterations = r.newton_iterations;
        result->linear_iterations = r.linear_iterations;
        result->consistency_residual = r.consistency_residual_W_m2;
        result->sum_W = r.sum_W;
        result->absolute_sum_W = r.absolute_sum_W;
        result->eddy_absolute_sum_W = r.eddy_absolute_sum_W;
        result->advective_absolute_sum_W = r.advective_absolute_sum_W;
    });
}

}  // extern "C"
