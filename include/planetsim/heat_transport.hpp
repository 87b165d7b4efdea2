// The transport graph, GraphMultigrid and diffusion source with the
// reference's signatures (sim/planet/surface/heat_transport.hpp:27-76,
// 136-138), and the implicit heat step as a new entry point in the same
// style (the reference's BiCGSTAB is internal, heat_transport.cpp:820-897).
// All numerical work runs on the device through libpsim.
#pragma once

#include "field.hpp"
#include "planet_mesh.hpp"

#include <cstddef>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <utility>
#include <span>
#include <stdexcept>
#include <exception>
#include <vector>

namespace planetsim {

struct TransportGraph {
    std::vector<double> area_m2;
    std::vector<std::size_t> offset;   // size() + 1
    std::vector<std::size_t> neighbour;
    std::vector<double> weight;        // l / d
    std::vector<double> weight_sum;    // sum of weight per node
    std::vector<CellBlock> blocks;

    [[nodiscard]] std::size_t size() const noexcept { return area_m2.size(); }

    // Device copy with its multigrid hierarchy structure, made on first
    // device use and remade if the arrays are reallocated or resized. The
    // copy is keyed on the arrays' addresses and sizes only: values changed
    // in place (element writes, assign() with the same size) are not seen
    // until invalidate() is called. (The reference's TransportGraph is a plain
    // struct with no device copy; this cache is the drop-in's own.)
    [[nodiscard]] psim_graph* device() const {
        const Key key{area_m2.data(), offset.data(), neighbour.data(), weight.data(), weight_sum.data(),
                      area_m2.size(), neighbour.size()};
        if (!cache_ || !(cache_->key == key)) {
            static_assert(sizeof(std::size_t) == sizeof(std::uint64_t));
            psim_graph* g = nullptr;
            gpu::check(psim_graph_create(gpu::context(), size(), area_m2.data(),
                                         reinterpret_cast<const std::uint64_t*>(offset.data()),
                                         reinterpret_cast<const std::uint64_t*>(neighbour.data()), weight.data(),
                                         weight_sum.data(), &g));
            cache_ = std::make_shared<Cache>(key, g);
        }
        return cache_->graph;
    }

    // Drop the device copy: the next device() call uploads the arrays again.
    void invalidate() const noexcept { cache_.reset(); }

  private:
    struct Key {
        const void* p[5];
        std::size_t n, nnz;
        bool operator==(const Key& o) const {
            for (int i = 0; i < 5; ++i)
                if (p[i] != o.p[i]) return false;
            return n == o.n && nnz == o.nnz;
        }
    };
    struct Cache {
        Cache(const Key& k, psim_graph* g) : key(k), graph(g) {}
        ~Cache() { psim_graph_destroy(graph); }
        Key key;
        psim_graph* graph;
    };
    mutable std::shared_ptr<Cache> cache_;
};

// heat_transport.cpp:433-452: CSR rows in cell_edges order, w = l / d,
// row sums in entry order.
[[nodiscard]] inline TransportGraph mesh_transport_graph(const PlanetMesh& mesh) {
    TransportGraph graph;
    const std::size_t cells = mesh.cell_count();
    graph.offset.assign(cells + 1U, 0U);
    for (std::size_t c = 0; c < cells; ++c) {
        const CellId id{static_cast<CellId::value_type>(c)};
        graph.area_m2.push_back(mesh.cells()[c].area_m2);
        double sum = 0.0;
        for (const auto& ce : mesh.cell_edges(id)) {
            const EdgeGeometry& e = mesh.edge(ce.edge);
            graph.neighbour.push_back(ce.neighbor.to_index());
            graph.weight.push_back(e.length_m / e.centroid_distance_m);
            sum += graph.weight.back();
        }
        graph.offset[c + 1U] = graph.neighbour.size();
        graph.weight_sum.push_back(sum);
    }
    graph.blocks.assign(mesh.blocks().begin(), mesh.blocks().end());
    return graph;
}

namespace detail {
// One agglomeration per subdivision and radius for the process, as the
// reference caches it (heat_transport.cpp:517-530); built by libpsim's
// psim_agglomeration (build_agglomeration, :467-512, bit-identical).
struct Agglomeration {
    TransportGraph graph;
    std::vector<std::size_t> group_of_cell;
    std::unique_ptr<PlanetMesh> coarse_mesh;   // null at subdivision 0: the mesh itself
};

inline const Agglomeration& cached_agglomeration(const PlanetMesh& mesh) {
    static std::mutex mutex;
    static std::map<std::pair<std::uint32_t, double>, std::unique_ptr<Agglomeration>> cache;
    const std::lock_guard lock(mutex);
    auto& entry = cache[{mesh.subdivision(), mesh.radius_m()}];
    if (!entry || entry->group_of_cell.size() != mesh.cell_count()) {
        psim_mesh* raw = nullptr;
        gpu::check(psim_mesh_create(mesh.subdivision(), mesh.radius_m(), &raw));
        const std::unique_ptr<psim_mesh, void (*)(psim_mesh*)> fine(raw, psim_mesh_destroy);
        psim_agglomeration* handle = nullptr;
        gpu::check(psim_agglomeration_create(raw, &handle));
        const std::unique_ptr<psim_agglomeration, void (*)(psim_agglomeration*)> owned(handle,
                                                                                      psim_agglomeration_destroy);
        std::uint64_t cells = 0, groups = 0;
        gpu::check(psim_agglomeration_counts(handle, &cells, &groups));
        std::vector<std::uint32_t> group(cells);
        std::vector<double> area(groups);
        gpu::check(psim_agglomeration_groups(handle, group.data(), area.data()));
        auto next = std::make_unique<Agglomeration>();
        if (mesh.subdivision() == 0U) {
            next->graph = mesh_transport_graph(mesh);
        } else {
            next->coarse_mesh = std::make_unique<PlanetMesh>(
                mesh_from_handle(psim_agglomeration_mesh(handle), mesh.subdivision() - 1U, mesh.radius_m()));
            next->graph = mesh_transport_graph(*next->coarse_mesh);
        }
        next->graph.area_m2 = std::move(area);
        next->group_of_cell.assign(group.begin(), group.end());
        entry = std::move(next);
    }
    return *entry;
}
}  // namespace detail

// heat_transport.hpp:45-50: the transport graph of the mesh's groups (each
// cell joins the nearest cell of the mesh one level coarser; a group's area
// is its cells'), and that coarser mesh (the mesh itself at subdivision 0).
// References stay valid for the process, as in the reference.
[[nodiscard]] inline const TransportGraph& agglomerated_transport_graph(
    const PlanetMesh& mesh, const std::vector<std::size_t>*& group_of_cell) {
    const detail::Agglomeration& entry = detail::cached_agglomeration(mesh);
    group_of_cell = &entry.group_of_cell;
    return entry.graph;
}

[[nodiscard]] inline const PlanetMesh& agglomerated_mesh(const PlanetMesh& mesh) {
    const detail::Agglomeration& entry = detail::cached_agglomeration(mesh);
    return entry.coarse_mesh ? *entry.coarse_mesh : mesh;
}

// The cell-to-group maps of a transport on the agglomerated graph, on the
// device. The reference writes them as loops around its response
// (surface_energy.cpp:1226-1247,1285-1300,1306-1324); this class has no
// counterpart there. Several fields may be given one after another.
class GroupMaps {
  public:
    explicit GroupMaps(const PlanetMesh& mesh) {
        const std::vector<std::size_t>* map = nullptr;
        const TransportGraph& graph = agglomerated_transport_graph(mesh, map);
        const std::vector<std::uint32_t> group(map->begin(), map->end());
        std::vector<double> area(mesh.cell_count());
        for (std::size_t c = 0; c < area.size(); ++c) area[c] = mesh.cells()[c].area_m2;
        psim_groups* handle = nullptr;
        gpu::check(psim_groups_create(gpu::context(), area.size(), graph.size(), group.data(), area.data(),
                                      graph.area_m2.data(), &handle));
        handle_.reset(handle);
        cells_ = area.size();
        groups_ = graph.size();
    }

    [[nodiscard]] std::size_t cells() const noexcept { return cells_; }
    [[nodiscard]] std::size_t groups() const noexcept { return groups_; }

    // A group's area-weighted mean of its cells (surface_energy.cpp:1285-1300).
    void mean(std::span<const double> cell_values, std::span<double> group_values) const {
        gpu::check(psim_groups_mean(handle_.get(), fields(cell_values.size(), cells_, group_values.size(), groups_),
                                    cell_values.data(), group_values.data(), PSIM_HOST));
    }
    // A group's highest cell value, from -inf (surface_energy.cpp:1306,1323-1324).
    void max(std::span<const double> cell_values, std::span<double> group_values) const {
        gpu::check(psim_groups_max(handle_.get(), fields(cell_values.size(), cells_, group_values.size(), groups_),
                                   cell_values.data(), group_values.data(), PSIM_HOST));
    }
    // Each cell takes its group's value (surface_energy.cpp:1247).
    void broadcast(std::span<const double> group_values, std::span<double> cell_values) const {
        gpu::check(psim_groups_broadcast(handle_.get(),
                                         fields(group_values.size(), groups_, cell_values.size(), cells_),
                                         group_values.data(), cell_values.data(), PSIM_HOST));
    }

  private:
    static std::uint32_t fields(std::size_t in, std::size_t per_in, std::size_t out, std::size_t per_out) {
        const std::size_t count = in / per_in;
        if (count == 0U || in != count * per_in || out != count * per_out || count > 65535U)
            throw std::invalid_argument("field size does not match the groups");
        return static_cast<std::uint32_t>(count);
    }
    struct Free {
        void operator()(psim_groups* g) const noexcept { psim_groups_destroy(g); }
    };
    std::unique_ptr<psim_groups, Free> handle_;
    std::size_t cells_ = 0, groups_ = 0;
};

namespace detail {
struct SystemFree {
    void operator()(psim_system* s) const noexcept { psim_system_destroy(s); }
};
struct BatchFree {
    void operator()(psim_batch* b) const noexcept { psim_batch_destroy(b); }
};
}  // namespace detail

// The aggregation-multigrid V-cycle on a transport graph. The device form
// smooths with Chebyshev-weighted Jacobi (the reference uses block
// Gauss-Seidel), so precondition() outputs differ from the reference's;
// aggregates, coarse patterns and Galerkin values are bit-identical.
class GraphMultigrid {
  public:
    explicit GraphMultigrid(const TransportGraph& graph) : n_(graph.size()), nnz_(graph.neighbour.size()) {
        graph_hold_ = std::make_shared<TransportGraph>(graph);
        psim_system* s = nullptr;
        gpu::check(psim_system_create(graph_hold_->device(), PSIM_PLAIN_FP64, &s));
        system_.reset(s);
    }
    GraphMultigrid(const GraphMultigrid&) = delete;
    GraphMultigrid& operator=(const GraphMultigrid&) = delete;
    GraphMultigrid(GraphMultigrid&&) noexcept = default;
    GraphMultigrid& operator=(GraphMultigrid&&) noexcept = default;

    void set_matrix(std::span<const double> diagonal, std::span<const double> off_diagonal) {
        if (diagonal.size() != n_ || off_diagonal.size() != nnz_)
            throw std::invalid_argument("multigrid matrix does not match the graph");
        gpu::check(psim_system_set_matrix(system_.get(), diagonal.data(), diagonal.size(), off_diagonal.data(),
                                          off_diagonal.size(), PSIM_HOST));
    }

    void precondition(const std::vector<double>& rhs, std::vector<double>& solution,
                      std::size_t worker_count = 1U) const {
        gpu::check_workers(worker_count);
        if (rhs.size() != n_) throw std::invalid_argument("multigrid right-hand side does not match the graph");
        solution.assign(n_, 0.0);
        gpu::check(psim_system_precondition(system_.get(), rhs.data(), solution.data(), PSIM_HOST));
    }

  private:
    std::size_t n_ = 0, nnz_ = 0;
    std::shared_ptr<TransportGraph> graph_hold_;
    std::unique_ptr<psim_system, detail::SystemFree> system_;
};

// heat_transport.cpp:579-593: K sum_e w_e (T_n - T_c) / A_c in W/m^2;
// the output is assigned.
inline void diffusion_source(const TransportGraph& graph, double conductance_W_K, const Field2D<double>& temperature_K,
                             Field2D<double>& source_W_m2, std::size_t worker_count = 1U) {
    if (temperature_K.size() != graph.size())
        throw std::invalid_argument("transport temperature does not match the mesh");
    gpu::check_workers(worker_count);
    source_W_m2 = Field2D<double>(graph.size(), 0.0);
    gpu::check(psim_diffusion_source(graph.device(), conductance_W_K, temperature_K.values().data(),
                                     source_W_m2.values().data(), PSIM_HOST));
}

// ---------------------------------------------------------------------------
// The implicit transport (heat_transport.hpp:78-138): Newton on
// F(h) = h - K lap T(h) with the caller's column response, the reference's
// line search, kink handling and result sums. The response runs on the host
// (called at every iterate and line-search trial); the diffusion and the
// Newton system's PCG run on the device.

using TransportResponse = std::function<void(const Field2D<double>& source_W_m2, Field2D<double>& mean_K,
                                             Field2D<double>& slope_K_m2_W)>;

struct ImplicitTransportSettings {
    int max_newton_iterations = 30;
    int max_line_search_halvings = 12;
    double rounding_floor_W_m2 = 1e-3;
    int max_non_monotone_steps = 4;
    double max_non_monotone_growth = 100.0;
    double newton_tolerance_W_m2 = 1e-6;
    double cg_relative_tolerance = 1e-6;
    int max_cg_iterations = 2'000;
    const Field2D<double>* initial_source_W_m2 = nullptr;   // accepted; the first iterate is the floor
};

inline constexpr double coupled_newton_tolerance_W_m2 = 1e-4;

struct ImplicitTransportResult {
    Field2D<double> source_W_m2;
    Field2D<double> mean_K;
    double consistency_residual_W_m2 = 0.0;
    int newton_iterations = 0;
    int cg_iterations = 0;
    double sum_W = 0.0;
    double absolute_sum_W = 0.0;
    double dissipation_W_K = 0.0;
};

namespace detail {
struct ResponseCall {
    const TransportResponse* response;
    std::size_t n;
    std::exception_ptr error;
};
// C callback into the std::function; an exception is held and rethrown
// once the library call returns (it cannot cross the C boundary).
inline void response_trampoline(void* user, const double* source, double* mean, double* slope, std::uint64_t n) {
    auto* call = static_cast<ResponseCall*>(user);
    if (call->error) return;
    try {
        const Field2D<double> h(std::vector<double>(source, source + n));
        Field2D<double> m(n, 0.0), s(n, 0.0);
        (*call->response)(h, m, s);
        std::copy(m.values().begin(), m.values().end(), mean);
        std::copy(s.values().begin(), s.values().end(), slope);
    } catch (...) {
        call->error = std::current_exception();
        std::fill(mean, mean + n, 0.0);
        std::fill(slope, slope + n, 0.0);
    }
}

inline psim_transport_settings transport_settings(const ImplicitTransportSettings& settings) {
    psim_transport_settings st{};
    st.max_newton_iterations = settings.max_newton_iterations;
    st.max_line_search_halvings = settings.max_line_search_halvings;
    st.rounding_floor_W_m2 = settings.rounding_floor_W_m2;
    st.max_non_monotone_steps = settings.max_non_monotone_steps;
    st.max_non_monotone_growth = settings.max_non_monotone_growth;
    st.newton_tolerance_W_m2 = settings.newton_tolerance_W_m2;
    st.cg_relative_tolerance = settings.cg_relative_tolerance;
    st.max_cg_iterations = settings.max_cg_iterations;
    return st;
}

inline void fill_result(const psim_transport_result& r, ImplicitTransportResult& result) {
    result.consistency_residual_W_m2 = r.consistency_residual_W_m2;
    result.newton_iterations = r.newton_iterations;
    result.cg_iterations = r.cg_iterations;
    result.sum_W = r.sum_W;
    result.absolute_sum_W = r.absolute_sum_W;
    result.dissipation_W_K = r.dissipation_W_K;
}
}  // namespace detail

[[nodiscard]] inline ImplicitTransportResult solve_implicit_transport(
    const TransportGraph& graph, double conductance_W_K, const Field2D<double>& source_floor_W_m2,
    const TransportResponse& response, const ImplicitTransportSettings& settings = {},
    std::size_t worker_count = 1U, const std::vector<double>* latent_W_m2 = nullptr) {
    const std::size_t n = graph.size();
    if (source_floor_W_m2.size() != n) throw std::invalid_argument("transport floor does not match the mesh");
    if (latent_W_m2 != nullptr && latent_W_m2->size() != n)
        throw std::invalid_argument("transport latent heat does not match the mesh");
    gpu::check_workers(worker_count);
    const psim_transport_settings st = detail::transport_settings(settings);
    detail::ResponseCall call{&response, n, nullptr};
    ImplicitTransportResult result;
    result.source_W_m2 = Field2D<double>(n, 0.0);
    result.mean_K = Field2D<double>(n, 0.0);
    psim_transport_result r{};
    const int status = psim_implicit_transport(graph.device(), conductance_W_K, source_floor_W_m2.values().data(),
                                               &detail::response_trampoline, &call, &st,
                                               latent_W_m2 != nullptr ? latent_W_m2->data() : nullptr,
                                               result.source_W_m2.values().data(), result.mean_K.values().data(), &r);
    if (call.error) std::rethrow_exception(call.error);
    gpu::check(status);
    detail::fill_result(r, result);
    return result;
}

// A response held on the device, so that no field crosses PCIe inside the
// Newton loop (no reference counterpart: the reference's response is a
// callback). Per node a piecewise-linear table of `points` (source,
// temperature) pairs, point-major (pair k of node i at k size() + i),
// sources increasing: T(h) follows the segment whose left end is the last
// one <= h (the end segments extend past the table), the slope is that
// segment's, and a flat segment holds the node.
struct TransportResponseTable {
    std::size_t points = 0;
    std::vector<double> source_W_m2;
    std::vector<double> mean_K;
};

[[nodiscard]] inline ImplicitTransportResult solve_implicit_transport(
    const TransportGraph& graph, double conductance_W_K, const Field2D<double>& source_floor_W_m2,
    const TransportResponseTable& table, const ImplicitTransportSettings& settings = {},
    std::size_t worker_count = 1U, const std::vector<double>* latent_W_m2 = nullptr) {
    const std::size_t n = graph.size();
    if (source_floor_W_m2.size() != n) throw std::invalid_argument("transport floor does not match the mesh");
    if (latent_W_m2 != nullptr && latent_W_m2->size() != n)
        throw std::invalid_argument("transport latent heat does not match the mesh");
    if (table.source_W_m2.size() != table.points * n || table.mean_K.size() != table.points * n)
        throw std::invalid_argument("transport response table does not match the mesh");
    gpu::check_workers(worker_count);
    const psim_transport_settings st = detail::transport_settings(settings);
    const psim_response_table t{static_cast<std::uint32_t>(table.points), 0U, table.source_W_m2.data(),
                                table.mean_K.data()};
    ImplicitTransportResult result;
    result.source_W_m2 = Field2D<double>(n, 0.0);
    result.mean_K = Field2D<double>(n, 0.0);
    psim_transport_result r{};
    gpu::check(psim_implicit_transport_table(graph.device(), conductance_W_K, source_floor_W_m2.values().data(), &t,
                                             &st, latent_W_m2 != nullptr ? latent_W_m2->data() : nullptr,
                                             result.source_W_m2.values().data(), result.mean_K.values().data(), &r,
                                             PSIM_HOST));
    detail::fill_result(r, result);
    return result;
}

// ---------------------------------------------------------------------------
// The coupled climate transport (heat_transport.hpp:140-250 of the
// reference): eddy diffusion of the columns' theta plus the upwind advection
// of each layer's dry static energy, its source, its implicit Newton solve
// (block ILU(0), V-cycle and BiCGSTAB on the device) and a tracer's implicit
// step under the same fluxes.
struct AdvectionDiffusion {
    std::size_t layers = 0;
    std::vector<double> conductance_W_K;
    std::vector<double> outflow_kg_s;
    double heat_capacity_J_kg_K = 1004.64;
    std::vector<double> eddy_latent_W_m2;
};

using AdvectionResponse =
    std::function<void(const Field2D<double>& source_W_m2, Field2D<double>& mean_K, Field2D<double>& slope_K_m2_W,
                       std::vector<double>& energy_J_kg, std::vector<double>& energy_slope_J_kg_m2_W)>;

struct AdvectionDiffusionResult {
    Field2D<double> source_W_m2;
    Field2D<double> mean_K;
    std::vector<double> energy_J_kg;
    double consistency_residual_W_m2 = 0.0;
    int newton_iterations = 0;
    int linear_iterations = 0;
    double sum_W = 0.0;
    double absolute_sum_W = 0.0;
    double eddy_absolute_sum_W = 0.0;
    double advective_absolute_sum_W = 0.0;
};

namespace detail {
inline void check_advection(const TransportGraph& graph, const AdvectionDiffusion& transport) {
    const std::size_t e = graph.neighbour.size();
    if (transport.layers == 0U || transport.conductance_W_K.size() != e ||
        transport.outflow_kg_s.size() != transport.layers * e ||
        (!transport.eddy_latent_W_m2.empty() && transport.eddy_latent_W_m2.size() != graph.size()))
        throw std::invalid_argument("advection-diffusion inputs do not match the graph");
}

struct AdvectionCall {
    const AdvectionResponse* response;
    std::exception_ptr error;
};
// C callback into the std::function (sizes checked as the reference
// checks them); an exception is held and rethrown after the library call.
inline void advection_trampoline(void* user, const double* source, double* mean, double* slope, double* energy,
                                 double* energy_slope, std::uint64_t n, std::uint32_t layers) {
    auto* call = static_cast<AdvectionCall*>(user);
    const std::size_t cells = static_cast<std::size_t>(n), values = cells * layers;
    if (!call->error) {
        try {
            const Field2D<double> h(std::vector<double>(source, source + cells));
            Field2D<double> m(cells, 0.0), s(cells, 0.0);
            std::vector<double> e(values, 0.0), es(values, 0.0);
            (*call->response)(h, m, s, e, es);
            if (m.size() != cells || s.size() != cells || e.size() != values || es.size() != values)
                throw std::logic_error("advection-diffusion response has the wrong size");
            std::copy(m.values().begin(), m.values().end(), mean);
            std::copy(s.values().begin(), s.values().end(), slope);
            std::copy(e.begin(), e.end(), energy);
            std::copy(es.begin(), es.end(), energy_slope);
            return;
        } catch (...) {
            call->error = std::current_exception();
        }
    }
    std::fill(mean, mean + cells, 0.0);
    std::fill(slope, slope + cells, 0.0);
    std::fill(energy, energy + values, 0.0);
    std::fill(energy_slope, energy_slope + values, 0.0);
}
}  // namespace detail

// heat_transport.cpp:774-816, bit-identical.
inline void advection_diffusion_source(const TransportGraph& graph, const AdvectionDiffusion& transport,
                                       const Field2D<double>& mean_K, const std::vector<double>& energy_J_kg,
                                       Field2D<double>& eddy_W_m2, Field2D<double>& advective_W_m2,
                                       std::size_t worker_count = 1U) {
    detail::check_advection(graph, transport);
    const std::size_t n = graph.size();
    if (mean_K.size() != n || energy_J_kg.size() != transport.layers * n)
        throw std::invalid_argument("advection-diffusion inputs do not match the graph");
    gpu::check_workers(worker_count);
    if (eddy_W_m2.size() != n) eddy_W_m2 = Field2D<double>(n, 0.0);
    if (advective_W_m2.size() != n) advective_W_m2 = Field2D<double>(n, 0.0);
    gpu::check(psim_advection_source(graph.device(), static_cast<std::uint32_t>(transport.layers),
                                     transport.conductance_W_K.data(), transport.outflow_kg_s.data(),
                                     transport.eddy_latent_W_m2.empty() ? nullptr : transport.eddy_latent_W_m2.data(),
                                     mean_K.values().data(), energy_J_kg.data(), eddy_W_m2.values().data(),
                                     advective_W_m2.values().data(), PSIM_HOST));
}

// heat_transport.cpp:1083-1318.
[[nodiscard]] inline AdvectionDiffusionResult solve_implicit_advection_diffusion(
    const TransportGraph& graph, const AdvectionDiffusion& transport, const Field2D<double>& source_floor_W_m2,
    const AdvectionResponse& response, const ImplicitTransportSettings& settings = {},
    std::size_t worker_count = 1U) {
    detail::check_advection(graph, transport);
    const std::size_t n = graph.size();
    if (source_floor_W_m2.size() != n) throw std::invalid_argument("advection-diffusion inputs do not match the graph");
    if (settings.initial_source_W_m2 != nullptr && settings.initial_source_W_m2->size() != n)
        throw std::invalid_argument("the initial sources do not match the graph");
    gpu::check_workers(worker_count);
    const psim_transport_settings st = detail::transport_settings(settings);
    detail::AdvectionCall call{&response, nullptr};
    AdvectionDiffusionResult result;
    result.source_W_m2 = Field2D<double>(n, 0.0);
    result.mean_K = Field2D<double>(n, 0.0);
    result.energy_J_kg.assign(transport.layers * n, 0.0);
    psim_advection_result r{};
    const int status = psim_implicit_advection_diffusion(
        graph.device(), static_cast<std::uint32_t>(transport.layers), transport.conductance_W_K.data(),
        transport.outflow_kg_s.data(), transport.eddy_latent_W_m2.empty() ? nullptr : transport.eddy_latent_W_m2.data(),
        source_floor_W_m2.values().data(), &detail::advection_trampoline, &call, &st,
        settings.initial_source_W_m2 != nullptr ? settings.initial_source_W_m2->values().data() : nullptr,
        result.source_W_m2.values().data(), result.mean_K.values().data(), result.energy_J_kg.data(), &r);
    if (call.error) std::rethrow_exception(call.error);
    gpu::check(status);
    result.consistency_residual_W_m2 = r.consistency_residual_W_m2;
    result.newton_iterations = r.newton_iterations;
    result.linear_iterations = r.linear_iterations;
    result.sum_W = r.sum_W;
    result.absolute_sum_W = r.absolute_sum_W;
    result.eddy_absolute_sum_W = r.eddy_absolute_sum_W;
    result.advective_absolute_sum_W = r.advective_absolute_sum_W;
    return result;
}

struct TracerTransportResult {
    std::vector<double> tracer;
    int iterations = 0;
    double relative_residual = 0.0;
    double clipped_kg = 0.0;
    double column_divergence_kg_s = 0.0;
};

// heat_transport.cpp:901-1081.
[[nodiscard]] inline TracerTransportResult solve_implicit_tracer(const TransportGraph& graph,
                                                                 const AdvectionDiffusion& transport,
                                                                 std::span<const double> layer_mass_kg,
                                                                 const std::vector<double>& tracer, double dt_s,
                                                                 double relative_tolerance = 1e-12,
                                                                 int max_iterations = 2'000,
                                                                 std::size_t worker_count = 1U) {
    const std::size_t n = graph.size(), e = graph.neighbour.size(), layers = transport.layers;
    if (layers == 0U || layer_mass_kg.size() != n || tracer.size() != layers * n ||
        transport.conductance_W_K.size() != e || transport.outflow_kg_s.size() != layers * e || !(dt_s > 0.0))
        throw std::invalid_argument("tracer transport inputs do not match the graph");
    gpu::check_workers(worker_count);
    TracerTransportResult result;
    result.tracer.assign(layers * n, 0.0);
    psim_tracer_result r{};
    gpu::check(psim_tracer_step(graph.device(), static_cast<std::uint32_t>(layers), transport.conductance_W_K.data(),
                                transport.outflow_kg_s.data(), transport.heat_capacity_J_kg_K, layer_mass_kg.data(),
                                tracer.data(), dt_s, relative_tolerance, max_iterations, result.tracer.data(), &r));
    result.iterations = r.iterations;
    result.relative_residual = r.relative_residual;
    result.clipped_kg = r.clipped_kg;
    result.column_divergence_kg_s = r.column_divergence_kg_s;
    return result;
}

// ---------------------------------------------------------------------------
// The implicit heat step (new; no reference counterpart beyond its parts):
//   (diag(C_i A_i) + dt L_K) T = C_i A_i T_old + dt A_i S_i,
// L_K the two-point Laplacian with entry weights K_e l_e / d_e, solved by
// right-preconditioned BiCGSTAB with the V-cycle to a true relative residual.

struct HeatStepSettings {
    double dt_s = 2.63e6;
    double relative_tolerance = 1.0e-10;
    int max_iterations = 1000;
    bool mixed_precision = false;   // PSIM_MIXED: FP32 V-cycle after an FP64 first sweep, FP64 BiCGSTAB
};

struct HeatStepResult {
    int iterations = 0;
    int restarts = 0;
    int v_cycles = 0;
    double relative_residual = 0.0;
    bool converged = false;
};

// Keeps the device matrix and hierarchy of one graph between steps.
class HeatTransportSolver {
  public:
    explicit HeatTransportSolver(const TransportGraph& graph, bool mixed_precision = false)
        : n_(graph.size()), nnz_(graph.neighbour.size()) {
        graph_hold_ = std::make_shared<TransportGraph>(graph);
        psim_system* s = nullptr;
        gpu::check(psim_system_create(graph_hold_->device(), mixed_precision ? PSIM_MIXED : PSIM_PLAIN_FP64, &s));
        system_.reset(s);
    }

    // conductance_per_entry: K_e (m^2/s) of each CSR entry's edge;
    // capacity_area: C_i A_i.
    void set_conductance(std::span<const double> conductance_per_entry, std::span<const double> capacity_area,
                         double dt_s) {
        if (conductance_per_entry.size() != nnz_ || capacity_area.size() != n_)
            throw std::invalid_argument("heat-step inputs do not match the graph");
        gpu::check(psim_heat_set_matrix(system_.get(), conductance_per_entry.data(), capacity_area.data(), dt_s,
                                        PSIM_HOST));
        dt_s_ = dt_s;
    }

    HeatStepResult step(const Field2D<double>& temperature_K, const Field2D<double>& forcing_W_m2,
                        Field2D<double>& next_K, double relative_tolerance = 1.0e-10, int max_iterations = 1000) {
        if (temperature_K.size() != n_ || forcing_W_m2.size() != n_)
            throw std::invalid_argument("transport temperature does not match the mesh");
        next_K = Field2D<double>(n_, 0.0);
        psim_solve_info info{};
        gpu::check(psim_heat_step(system_.get(), temperature_K.values().data(), forcing_W_m2.values().data(), dt_s_,
                                  relative_tolerance, max_iterations, next_K.values().data(), &info, PSIM_HOST));
        return {info.iterations, info.restarts, info.v_cycles, info.relative_residual, info.status == PSIM_OK};
    }

    std::size_t size() const noexcept { return n_; }
    double time_step() const noexcept { return dt_s_; }
    psim_system* device_system() const noexcept { return system_.get(); }

  private:
    std::size_t n_ = 0, nnz_ = 0;
    double dt_s_ = 2.63e6;
    std::shared_ptr<TransportGraph> graph_hold_;
    std::unique_ptr<psim_system, detail::SystemFree> system_;
};

// Heat steps of many solvers (any graphs, any precisions, each with its own
// matrix set by set_conductance) in ONE device launch (new: the reference
// solves one instance at a time). Each result equals that solver's own
// step() bit for bit. The solvers must outlive the batch and share dt.
class HeatStepBatch {
  public:
    explicit HeatStepBatch(std::vector<HeatTransportSolver*> solvers) : solvers_(std::move(solvers)) {
        if (solvers_.empty()) throw std::invalid_argument("a batch needs at least one solver");
        std::vector<psim_system*> systems;
        for (HeatTransportSolver* s : solvers_) {
            if (s == nullptr) throw std::invalid_argument("null solver");
            systems.push_back(s->device_system());
        }
        psim_batch* b = nullptr;
        gpu::check(psim_batch_create(gpu::context(), systems.data(), static_cast<int32_t>(systems.size()), &b));
        batch_.reset(b);
        for (HeatTransportSolver* s : solvers_) {
            buffers_.push_back(gpu::DeviceBuffer(3 * s->size() * sizeof(double)));
        }
    }

    std::vector<HeatStepResult> step(std::span<const Field2D<double>* const> temperatures_K,
                                     std::span<const Field2D<double>* const> forcings_W_m2,
                                     std::span<Field2D<double>* const> next_K, double relative_tolerance = 1.0e-10,
                                     int max_iterations = 1000) {
        const std::size_t count = solvers_.size();
        if (temperatures_K.size() != count || forcings_W_m2.size() != count || next_K.size() != count)
            throw std::invalid_argument("one field per solver is needed");
        const double dt = solvers_[0]->time_step();
        std::vector<const double*> t_old(count), forcing(count);
        std::vector<double*> t_new(count);
        for (std::size_t i = 0; i < count; ++i) {
            const std::size_t n = solvers_[i]->size();
            if (temperatures_K[i] == nullptr || forcings_W_m2[i] == nullptr || next_K[i] == nullptr)
                throw std::invalid_argument("null field");
            if (temperatures_K[i]->size() != n || forcings_W_m2[i]->size() != n)
                throw std::invalid_argument("transport temperature does not match the mesh");
            if (solvers_[i]->time_step() != dt) throw std::invalid_argument("batched solvers must share dt");
            double* base = static_cast<double*>(buffers_[i].get());
            gpu::check(psim_copy_to_device(gpu::context(), base, temperatures_K[i]->values().data(), n * 8));
            gpu::check(psim_copy_to_device(gpu::context(), base + n, forcings_W_m2[i]->values().data(), n * 8));
            t_old[i] = base;
            forcing[i] = base + n;
            t_new[i] = base + 2 * n;
        }
        std::vector<psim_solve_info> infos(count);
        gpu::check(psim_batch_heat_step(batch_.get(), nullptr, nullptr, t_old.data(), forcing.data(), dt,
                                        relative_tolerance, max_iterations, t_new.data(), infos.data()));
        std::vector<HeatStepResult> out(count);
        for (std::size_t i = 0; i < count; ++i) {
            const std::size_t n = solvers_[i]->size();
            *next_K[i] = Field2D<double>(n, 0.0);
            gpu::check(psim_copy_to_host(gpu::context(), next_K[i]->values().data(), t_new[i], n * 8));
            out[i] = {infos[i].iterations, infos[i].restarts, infos[i].v_cycles, infos[i].relative_residual,
                      infos[i].status == PSIM_OK};
        }
        return out;
    }

  private:
    std::vector<HeatTransportSolver*> solvers_;
    std::unique_ptr<psim_batch, detail::BatchFree> batch_;
    std::vector<gpu::DeviceBuffer> buffers_;
};

inline HeatStepResult solve_heat_step(const TransportGraph& graph, std::span<const double> conductance_per_entry,
                                      std::span<const double> capacity_area, const Field2D<double>& temperature_K,
                                      const Field2D<double>& forcing_W_m2, Field2D<double>& next_K,
                                      const HeatStepSettings& settings = {}) {
    HeatTransportSolver solver(graph, settings.mixed_precision);
    solver.set_conductance(conductance_per_entry, capacity_area, settings.dt_s);
    return solver.step(temperature_K, forcing_W_m2, next_K, settings.relative_tolerance, settings.max_iterations);
}

}  // namespace planetsim
