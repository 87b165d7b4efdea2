// Drop-in check of the C++20 layer (include/planetsim/): the reference's
// call sites compile unchanged against it and behave as the reference's
// tests expect (tests/conservation/test_operator_conservation.cpp,
// tests/regression/test_operator_determinism.cpp,
// tests/unit/test_heat_transport.cpp). `cpu` runs the host-only part,
// `gpu` the part that needs a device.
#include "planetsim/planetsim.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numbers>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

int failures = 0;

#define EXPECT(condition)                                                              \
    do {                                                                               \
        if (!(condition)) {                                                            \
            std::fprintf(stderr, "%s:%d: expectation failed: %s\n", __FILE__, __LINE__, #condition); \
            ++failures;                                                                \
        }                                                                              \
    } while (false)

template <class E, class F> bool throws_with(F&& f, const char* message) {
    try {
        f();
    } catch (const E& e) {
        return message == nullptr || std::string(e.what()).find(message) != std::string::npos;
    } catch (...) {
        return false;
    }
    return false;
}

// splitmix-based uniform numbers for test fields (not the RNG under test)
double uniform(std::uint64_t key) {
    key += 0x9e3779b97f4a7c15ULL;
    key = (key ^ (key >> 30U)) * 0xbf58476d1ce4e5b9ULL;
    key = (key ^ (key >> 27U)) * 0x94d049bb133111ebULL;
    key ^= key >> 31U;
    return static_cast<double>(key >> 11U) * (1.0 / 9007199254740992.0);
}

void host_part() {
    for (std::uint32_t level = 0; level <= 5; ++level) {
        const auto mesh = planetsim::make_icosphere(level, 6'371'000.0);
        EXPECT(mesh.cell_count() == 10U * (1U << (2U * level)) + 2U);
        EXPECT(mesh.edge_count() == 30U * (1U << (2U * level)));
        std::size_t pentagons = 0;
        double total = 0.0, carry = 0.0;
        for (const auto& cell : mesh.cells()) {
            pentagons += cell.is_pentagon() ? 1U : 0U;
            const double piece = cell.area_m2 - carry;
            const double next = total + piece;
            carry = (next - total) - piece;
            total = next;
        }
        EXPECT(pentagons == 12U);
        const double expected = 4.0 * std::numbers::pi_v<double> * 6'371'000.0 * 6'371'000.0;
        EXPECT(std::abs(total - expected) / expected <= 5e-14);
        // reciprocal edges (icosphere.cpp:383-387)
        for (const auto& cell : mesh.cells())
            for (const auto& ce : mesh.cell_edges(cell.id)) {
                const auto& edge = mesh.edge(ce.edge);
                EXPECT((edge.first_cell == cell.id && edge.second_cell == ce.neighbor) ||
                       (edge.second_cell == cell.id && edge.first_cell == ce.neighbor));
            }
        const auto graph = planetsim::mesh_transport_graph(mesh);
        EXPECT(graph.size() == mesh.cell_count() && graph.offset.back() == mesh.directed_edge_count());
    }
    // test_heat_transport.cpp:173-212: the agglomerated graph covers every
    // cell once, its areas sum to the sphere's, its weights are symmetric,
    // and groups hold about four cells; the cache hands out one graph.
    for (std::uint32_t level = 0; level <= 5; ++level) {
        const auto fine = planetsim::make_icosphere(level, 6'371'000.0);
        const std::vector<std::size_t>* map = nullptr;
        const auto& graph = planetsim::agglomerated_transport_graph(fine, map);
        const auto& coarse = planetsim::agglomerated_mesh(fine);
        EXPECT(level == 0 ? &coarse == &fine : coarse.subdivision() == level - 1);
        EXPECT(graph.size() == coarse.cell_count() && map->size() == fine.cell_count());
        const std::vector<std::size_t>* again = nullptr;
        EXPECT(&planetsim::agglomerated_transport_graph(fine, again) == &graph && again == map);
        double fine_area = 0.0, graph_area = 0.0;
        for (const auto& cell : fine.cells()) fine_area += cell.area_m2;
        for (const double a : graph.area_m2) graph_area += a;
        EXPECT(std::abs(graph_area - fine_area) <= 1e-12 * fine_area);
        bool covered = true, symmetric = true;
        for (const std::size_t g : *map) covered = covered && g < graph.size();
        for (std::size_t g = 0; g < graph.size(); ++g)
            for (std::size_t k = graph.offset[g]; k < graph.offset[g + 1]; ++k) {
                const std::size_t other = graph.neighbour[k];
                bool found = false;
                for (std::size_t j = graph.offset[other]; j < graph.offset[other + 1]; ++j)
                    found = found || (graph.neighbour[j] == g && graph.weight[j] == graph.weight[k]);
                symmetric = symmetric && found && other != g && graph.weight[k] > 0.0;
            }
        EXPECT(covered && symmetric);
        const double per_group = static_cast<double>(fine.cell_count()) / static_cast<double>(graph.size());
        EXPECT(level < 3 || (per_group > 3.9 && per_group < 4.1));
    }
    const auto mesh = planetsim::make_icosphere(2, 6'371'000.0);
    EXPECT(throws_with<std::invalid_argument>([&] { (void)planetsim::make_icosphere(2, -5.0); },
                                              "icosphere radius must be finite and positive"));
    EXPECT(throws_with<std::invalid_argument>(
        [&] { (void)planetsim::edge_orientation(mesh, planetsim::CellId{0}, planetsim::EdgeId{static_cast<std::uint32_t>(mesh.edge_count() - 1)}); },
        "cell is not incident to edge"));
    planetsim::Field2D<double> wrong(mesh.cell_count() + 1), out(mesh.cell_count());
    EXPECT(throws_with<std::invalid_argument>([&] { planetsim::laplacian(mesh, wrong, out); },
                                              "cell field size does not match the mesh"));
    planetsim::EdgeField<double> wrong_flux(mesh.edge_count() - 1);
    EXPECT(throws_with<std::invalid_argument>([&] { planetsim::divergence(mesh, wrong_flux, out); },
                                              "edge field size does not match the mesh"));
}

// Names the section that runs next, so that an exception names its place.
void stage(const char* name) { std::fprintf(stderr, "[device part] %s\n", name); }

void device_part() {
    const auto mesh = planetsim::make_icosphere(5, 6'371'000.0);
    const std::size_t n = mesh.cell_count();
    stage("operator conservation");
    // test_operator_conservation.cpp:36-57: constant fields give exactly 0.
    {
        const planetsim::Field2D<double> constant(n, 287.5);
        planetsim::Field2D<double> east(n, 1.0), north(n, 1.0), lap(n, 1.0), div(n, 1.0);
        planetsim::gradient(mesh, constant, east, north);
        planetsim::laplacian(mesh, constant, lap);
        planetsim::divergence(mesh, planetsim::EdgeField<double>(mesh.edge_count(), 0.0), div);
        bool zero = true;
        for (std::size_t c = 0; c < n; ++c) zero = zero && east[c] == 0.0 && north[c] == 0.0 && lap[c] == 0.0 && div[c] == 0.0;
        EXPECT(zero);
    }
    stage("operator determinism");
    // test_operator_determinism.cpp: bit-identical for any worker count;
    // the Laplacian is conservative and dissipative.
    {
        planetsim::Field2D<float> scalar(n);
        for (std::size_t c = 0; c < n; ++c) scalar[c] = static_cast<float>(uniform(11 * 1000003ULL + c));
        planetsim::EdgeField<float> flux(mesh.edge_count());
        for (std::size_t e = 0; e < mesh.edge_count(); ++e)
            flux[planetsim::EdgeId{static_cast<std::uint32_t>(e)}] = static_cast<float>(uniform(7 * 1000003ULL + e));
        planetsim::Field2D<float> e1(n), n1(n), d1(n), l1(n), e2(n), n2(n), d2(n), l2(n);
        planetsim::gradient(mesh, scalar, e1, n1, 1U);
        planetsim::divergence(mesh, flux, d1, 1U);
        planetsim::laplacian(mesh, scalar, l1, 1U);
        planetsim::gradient(mesh, scalar, e2, n2, 16U);
        planetsim::divergence(mesh, flux, d2, 16U);
        planetsim::laplacian(mesh, scalar, l2, 16U);
        EXPECT(std::memcmp(e1.values().data(), e2.values().data(), n * 4) == 0);
        EXPECT(std::memcmp(d1.values().data(), d2.values().data(), n * 4) == 0);
        EXPECT(std::memcmp(l1.values().data(), l2.values().data(), n * 4) == 0);
        planetsim::Field2D<double> phi(n), lap(n);
        for (std::size_t c = 0; c < n; ++c) phi[c] = uniform(3 * 1000003ULL + c) - 0.5;
        planetsim::laplacian(mesh, phi, lap);
        double sum = 0.0, abs_sum = 0.0, dissipation = 0.0;
        for (std::size_t c = 0; c < n; ++c) {
            sum += mesh.cells()[c].area_m2 * lap[c];
            abs_sum += mesh.cells()[c].area_m2 * std::abs(lap[c]);
            dissipation += mesh.cells()[c].area_m2 * phi[c] * lap[c];
        }
        EXPECT(std::abs(sum) <= 1e-13 * abs_sum);
        EXPECT(dissipation <= 0.0);
    }
    stage("multigrid and heat step");
    // GraphMultigrid and the heat step on the transport graph.
    {
        const auto graph = planetsim::mesh_transport_graph(mesh);
        const double dt = 2.63e6;
        std::vector<double> capacity(n), conductance(graph.neighbour.size()), diag(n), off(graph.neighbour.size());
        for (std::size_t c = 0; c < n; ++c) capacity[c] = (uniform(c) < 0.3 ? 1.0e7 : 4.0e8) * graph.area_m2[c];
        for (std::size_t c = 0; c < n; ++c) {
            double s = 0.0;
            for (std::size_t k = graph.offset[c]; k < graph.offset[c + 1U]; ++k) {
                conductance[k] = 0.5 * 6'371'000.0 * 6'371'000.0;   // stiff: Laplacian-dominated
                const double g = conductance[k] * graph.weight[k];
                s += g;
                off[k] = -dt * g;
            }
            diag[c] = capacity[c] + dt * s;
        }
        planetsim::GraphMultigrid multigrid(graph);
        EXPECT(throws_with<std::invalid_argument>([&] { multigrid.set_matrix(std::span(diag).first(n - 1), off); },
                                                  "multigrid matrix does not match the graph"));
        multigrid.set_matrix(diag, off);
        std::vector<double> rhs(n), z;
        for (std::size_t c = 0; c < n; ++c) rhs[c] = uniform(99 + c) - 0.5;
        multigrid.precondition(rhs, z, 4U);
        double rz = 0.0;
        for (std::size_t c = 0; c < n; ++c) rz += rhs[c] * z[c];
        EXPECT(z.size() == n && rz > 0.0);

        planetsim::Field2D<double> t_old(n), forcing(n), t_new;
        for (std::size_t c = 0; c < n; ++c) {
            t_old[c] = 288.0 - 40.0 * mesh.cells()[c].center_unit.z * mesh.cells()[c].center_unit.z;
            forcing[c] = 100.0 * std::sqrt(1.0 - mesh.cells()[c].center_unit.z * mesh.cells()[c].center_unit.z);
        }
        planetsim::HeatStepSettings settings;
        settings.relative_tolerance = 1e-10;
        const auto result = planetsim::solve_heat_step(graph, conductance, capacity, t_old, forcing, t_new, settings);
        EXPECT(result.converged && result.relative_residual <= 1e-10 && t_new.size() == n);
        // the same residual, recomputed here in FP64
        double rr = 0.0, bb = 0.0;
        for (std::size_t c = 0; c < n; ++c) {
            double ax = diag[c] * t_new[c];
            for (std::size_t k = graph.offset[c]; k < graph.offset[c + 1U]; ++k) ax += off[k] * t_new[graph.neighbour[k]];
            const double b = capacity[c] * t_old[c] + dt * graph.area_m2[c] * forcing[c];
            rr += (b - ax) * (b - ax);
            bb += b * b;
        }
        EXPECT(std::sqrt(rr / bb) <= 1.5e-10);
        // energy: sum (C A (T - T_old)) = dt sum A S (the Laplacian conserves)
        double stored = 0.0, supplied = 0.0;
        for (std::size_t c = 0; c < n; ++c) {
            stored += capacity[c] * (t_new[c] - t_old[c]);
            supplied += dt * graph.area_m2[c] * forcing[c];
        }
        EXPECT(std::abs(stored - supplied) <= 1e-8 * std::abs(supplied));
        // diffusion_source conserves energy: sum A H = 0 to rounding
        // (test_heat_transport.cpp:80-107 checks 1e-13 of sum A |H| on its
        // own fields; a rough field keeps the gross flux well above rounding)
        planetsim::Field2D<double> rough(n), source;
        for (std::size_t c = 0; c < n; ++c) rough[c] = 288.0 + 30.0 * (uniform(5 * 1000003ULL + c) - 0.5);
        planetsim::diffusion_source(graph, 3.5e12, rough, source);
        double net = 0.0, gross = 0.0;
        for (std::size_t c = 0; c < n; ++c) {
            net += graph.area_m2[c] * source[c];
            gross += graph.area_m2[c] * std::abs(source[c]);
        }
        EXPECT(std::abs(net) <= 1e-12 * gross);
    }
    stage("implicit transport");
    // solve_implicit_transport with a linear response (the reference's
    // check_linear, test_heat_transport.cpp:73-107).
    {
        const auto graph = planetsim::mesh_transport_graph(mesh);
        std::vector<double> base(n), slopes(n);
        for (std::size_t c = 0; c < n; ++c) {
            base[c] = 250.0 + 50.0 * uniform(61 * 1000003ULL + c);
            slopes[c] = c % 5U == 0U ? 0.0 : 0.01 + 0.2 * uniform(67 * 1000003ULL + c);
        }
        const planetsim::TransportResponse response = [&](const planetsim::Field2D<double>& source,
                                                          planetsim::Field2D<double>& mean,
                                                          planetsim::Field2D<double>& slope) {
            for (std::size_t c = 0; c < n; ++c) {
                mean[c] = base[c] + slopes[c] * source[c];
                slope[c] = slopes[c];
            }
        };
        const double conductance = 0.5 * 6'371'000.0 * 6'371'000.0;
        const planetsim::Field2D<double> floor(n, -std::numeric_limits<double>::infinity());
        const auto result = planetsim::solve_implicit_transport(graph, conductance, floor, response, {}, 4U);
        EXPECT(result.newton_iterations <= 3);
        EXPECT(result.consistency_residual_W_m2 <= 1e-6);
        EXPECT(std::abs(result.sum_W) <= 1e-13 * result.absolute_sum_W);
        EXPECT(result.dissipation_W_K <= 0.0);
        EXPECT(result.source_W_m2.size() == n && result.mean_K.size() == n);
        EXPECT(throws_with<std::invalid_argument>(
            [&] { (void)planetsim::solve_implicit_transport(graph, 0.0, floor, response); },
            "transport conductance must be finite and positive"));
        const planetsim::TransportResponse failing = [](const planetsim::Field2D<double>&,
                                                        planetsim::Field2D<double>&, planetsim::Field2D<double>&) {
            throw std::runtime_error("column did not converge");
        };
        EXPECT(throws_with<std::runtime_error>(
            [&] { (void)planetsim::solve_implicit_transport(graph, conductance, floor, failing); },
            "column did not converge"));
        // the same linear response as a device-resident 2-point table:
        // T(h) = base + s h through (0, base) and (100, base + 100 s)
        planetsim::TransportResponseTable table;
        table.points = 2;
        table.source_W_m2.assign(2 * n, 0.0);
        table.mean_K.assign(2 * n, 0.0);
        for (std::size_t c = 0; c < n; ++c) {
            table.source_W_m2[n + c] = 100.0;
            table.mean_K[c] = base[c];
            table.mean_K[n + c] = base[c] + 100.0 * slopes[c];
        }
        const auto tabled = planetsim::solve_implicit_transport(graph, conductance, floor, table);
        EXPECT(tabled.consistency_residual_W_m2 <= 1e-6 && tabled.newton_iterations <= 3);
        double largest = 0.0;
        for (std::size_t c = 0; c < n; ++c)
            largest = std::max(largest, std::abs(tabled.source_W_m2[c] - result.source_W_m2[c]));
        EXPECT(largest <= 1e-5);
        table.points = 3;
        EXPECT(throws_with<std::invalid_argument>(
            [&] { (void)planetsim::solve_implicit_transport(graph, conductance, floor, table); },
            "transport response table does not match the mesh"));
    }
    stage("agglomerated transport");
    // The agglomerated transport as the reference's surface step runs it:
    // group sources broadcast to cells, a per-cell response, the groups'
    // area-weighted means back (surface_energy.cpp:1222-1300), on the
    // device maps.
    {
        const std::vector<std::size_t>* map = nullptr;
        const auto& graph = planetsim::agglomerated_transport_graph(mesh, map);
        const planetsim::GroupMaps maps(mesh);
        EXPECT(maps.cells() == n && maps.groups() == graph.size());
        std::vector<double> ones(n, 1.0), mean(graph.size());
        maps.mean(ones, mean);
        bool exact = true;
        for (const double m : mean) exact = exact && m == 1.0;
        EXPECT(exact);
        EXPECT(throws_with<std::invalid_argument>([&] { maps.mean(std::span(ones).first(n - 1), mean); },
                                                  "field size does not match the groups"));
        std::vector<double> base(n), cell_h(n), cell_t(2 * n), both(2 * graph.size());
        for (std::size_t c = 0; c < n; ++c) base[c] = 250.0 + 40.0 * uniform(900 + c);
        const planetsim::TransportResponse response = [&](const planetsim::Field2D<double>& source,
                                                          planetsim::Field2D<double>& out_mean,
                                                          planetsim::Field2D<double>& out_slope) {
            maps.broadcast(source.values(), cell_h);
            for (std::size_t c = 0; c < n; ++c) {
                const double slope = 0.02 + 0.01 * uniform(1900 + c);
                cell_t[c] = base[c] + slope * cell_h[c];
                cell_t[n + c] = slope;
            }
            maps.mean(cell_t, both);
            for (std::size_t g = 0; g < graph.size(); ++g) {
                out_mean[g] = both[g];
                out_slope[g] = both[graph.size() + g];
            }
        };
        planetsim::Field2D<double> floor(graph.size(), -std::numeric_limits<double>::infinity());
        const double conductance = 0.5 * 6'371'000.0 * 6'371'000.0;
        const auto result = planetsim::solve_implicit_transport(graph, conductance, floor, response);
        EXPECT(result.consistency_residual_W_m2 <= 1e-6);
        EXPECT(std::abs(result.sum_W) <= 1e-12 * result.absolute_sum_W);
        EXPECT(result.dissipation_W_K <= 0.0);
    }
    stage("incomplete lu, tracer, advection-diffusion");
    {
        // IncompleteLU with one-row blocks drops every coupling: the solve
        // divides by the scaled diagonal, exactly
        const auto graph = planetsim::mesh_transport_graph(mesh);
        std::vector<double> diag(n), off(graph.neighbour.size());
        for (std::size_t c = 0; c < n; ++c) diag[c] = graph.weight_sum[c] * (1.0 + uniform(3000 + c));
        for (std::size_t k = 0; k < off.size(); ++k) off[k] = -graph.weight[k];
        std::vector<std::size_t> singles(n + 1);
        for (std::size_t c = 0; c <= n; ++c) singles[c] = c;
        const planetsim::IncompleteLU isolated(graph.offset, graph.neighbour, diag, off, singles, 2.0, 1U);
        std::vector<double> x(n, 3.0);
        isolated.solve(x);
        bool exact = true;
        for (std::size_t c = 0; c < n; ++c) exact = exact && x[c] == 3.0 / (diag[c] * 2.0);
        EXPECT(exact);
        EXPECT(throws_with<std::invalid_argument>(
            [&] { (void)planetsim::IncompleteLU(graph.offset, graph.neighbour, std::span(diag).first(n - 1), off); },
            "incomplete LU: the pattern does not match"));
        // a tracer step conserves mass and keeps c >= 0
        planetsim::AdvectionDiffusion transport;
        transport.layers = 2;
        transport.conductance_W_K.resize(graph.neighbour.size());
        transport.outflow_kg_s.assign(2 * graph.neighbour.size(), 0.0);
        for (std::size_t c = 0; c < n; ++c)
            for (std::size_t k = graph.offset[c]; k < graph.offset[c + 1]; ++k) {
                const std::size_t b = graph.neighbour[k];
                const double face = uniform(7000 + std::min(c, b) * 7 + std::max(c, b));
                transport.conductance_W_K[k] = 1.0e9 * (0.5 + face);
                const double sign = c < b ? 1.0 : -1.0;
                transport.outflow_kg_s[k] = sign * 2.0e7 * (face - 0.5);
                transport.outflow_kg_s[graph.neighbour.size() + k] = -sign * 1.0e7 * (face - 0.3);
            }
        std::vector<double> mass(n), tracer(2 * n);
        for (std::size_t c = 0; c < n; ++c) {
            mass[c] = graph.area_m2[c] * 1.0e4;
            tracer[c] = 1.0e-3 * (1.0 + uniform(8000 + c));
            tracer[n + c] = 2.0e-3 * uniform(9000 + c);
        }
        const auto stepped = planetsim::solve_implicit_tracer(graph, transport, mass, tracer, 2.63e6);
        double before = 0.0, after = 0.0;
        bool positive = true;
        for (std::size_t i = 0; i < 2 * n; ++i) {
            before += tracer[i] * mass[i % n];
            after += stepped.tracer[i] * mass[i % n];
            positive = positive && stepped.tracer[i] >= 0.0;
        }
        EXPECT(positive && stepped.relative_residual <= 1e-11);
        EXPECT(std::abs(after - before - stepped.clipped_kg) <= 1e-12 * before);
        EXPECT(throws_with<std::invalid_argument>(
            [&] { (void)planetsim::solve_implicit_tracer(graph, transport, mass, tracer, -1.0); },
            "tracer transport inputs do not match the graph"));
        // the coupled transport converges and conserves energy
        for (auto& g : transport.conductance_W_K) g *= 1.0e4;
        const planetsim::AdvectionResponse response = [&](const planetsim::Field2D<double>& h,
                                                          planetsim::Field2D<double>& theta,
                                                          planetsim::Field2D<double>& slope, std::vector<double>& s,
                                                          std::vector<double>& ds) {
            for (std::size_t c = 0; c < n; ++c) {
                const double k = 0.02 + 0.05 * uniform(10000 + c);
                theta[c] = 260.0 + 20.0 * uniform(11000 + c) + k * h[c];
                slope[c] = k;
                for (std::size_t l = 0; l < 2; ++l) {
                    s[l * n + c] = 1004.64 * (250.0 + 10.0 * l + 0.5 * (theta[c] - 250.0));
                    ds[l * n + c] = 1004.64 * 0.5 * k;
                }
            }
        };
        const planetsim::Field2D<double> floor(n, -std::numeric_limits<double>::infinity());
        const auto coupled = planetsim::solve_implicit_advection_diffusion(graph, transport, floor, response);
        EXPECT(coupled.consistency_residual_W_m2 <= 1e-6 && coupled.newton_iterations >= 1);
        EXPECT(std::abs(coupled.sum_W) <= 1e-12 * coupled.absolute_sum_W);
        EXPECT(coupled.energy_J_kg.size() == 2 * n);
        const planetsim::AdvectionResponse short_response = [&](const planetsim::Field2D<double>&,
                                                                planetsim::Field2D<double>&, planetsim::Field2D<double>&,
                                                                std::vector<double>& s, std::vector<double>&) {
            s.resize(1);
        };
        EXPECT(throws_with<std::logic_error>(
            [&] { (void)planetsim::solve_implicit_advection_diffusion(graph, transport, floor, short_response); },
            "advection-diffusion response has the wrong size"));
    }
    stage("heat step batch");
    // HeatStepBatch: solvers of different graphs and precisions in one
    // launch give exactly what each solver's own step() gives.
    {
        const auto small = planetsim::make_icosphere(3, 6'371'000.0);
        const planetsim::TransportGraph graphs[2] = {planetsim::mesh_transport_graph(mesh),
                                                     planetsim::mesh_transport_graph(small)};
        const double dt = 2.63e6;
        std::vector<planetsim::HeatTransportSolver> solvers;
        solvers.emplace_back(graphs[0], false);
        solvers.emplace_back(graphs[1], true);
        std::vector<planetsim::Field2D<double>> t_old, forcing, alone(2), batched(2);
        for (int q = 0; q < 2; ++q) {
            const auto& g = graphs[q];
            const std::size_t m = g.size();
            std::vector<double> capacity(m), conductance(g.neighbour.size());
            for (std::size_t c = 0; c < m; ++c) capacity[c] = (uniform(17 * q + c) < 0.3 ? 1.0e7 : 4.0e8) * g.area_m2[c];
            for (std::size_t k = 0; k < conductance.size(); ++k) conductance[k] = 1.0e6 * (1.0 + uniform(31 * q + k));
            solvers[static_cast<std::size_t>(q)].set_conductance(conductance, capacity, dt);
            planetsim::Field2D<double> t(m), f(m);
            for (std::size_t c = 0; c < m; ++c) {
                t[c] = 280.0 + 10.0 * uniform(41 * q + c);
                f[c] = 50.0 + 20.0 * uniform(53 * q + c);
            }
            t_old.push_back(std::move(t));
            forcing.push_back(std::move(f));
        }
        for (int q = 0; q < 2; ++q) {
            const auto r = solvers[static_cast<std::size_t>(q)].step(t_old[q], forcing[q], alone[q]);
            EXPECT(r.converged);
        }
        planetsim::HeatStepBatch batch({&solvers[0], &solvers[1]});
        const planetsim::Field2D<double>* ins[2] = {&t_old[0], &t_old[1]};
        const planetsim::Field2D<double>* fs[2] = {&forcing[0], &forcing[1]};
        planetsim::Field2D<double>* outs[2] = {&batched[0], &batched[1]};
        const auto results = batch.step(ins, fs, outs);
        for (int q = 0; q < 2; ++q) {
            EXPECT(results[static_cast<std::size_t>(q)].converged);
            EXPECT(batched[q].size() == alone[q].size() &&
                   std::memcmp(batched[q].values().data(), alone[q].values().data(), alone[q].size() * 8) == 0);
        }
        EXPECT(throws_with<std::invalid_argument>([&] { (void)batch.step(std::span(ins, 1), fs, outs); },
                                                  "one field per solver is needed"));
    }
}

}  // namespace

// Runs one part: 0 when every expectation held, 1 on failed expectations,
// 2 on an unexpected exception.
extern "C" int dropin_run(const char* part_name) {
    const std::string part = part_name != nullptr ? part_name : "cpu";
    failures = 0;
    try {
        if (part == "cpu") host_part();
        if (part == "gpu") device_part();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "unexpected exception: %s\n", e.what());
        return 2;
    }
    std::printf("%s part: %d failure(s)\n", part.c_str(), failures);
    std::fflush(stdout);
    return failures == 0 ? 0 : 1;
}

// The program form; the shared-object form (-DDROPIN_LIBRARY) lets a test
// runner call dropin_run inside its own process.
#ifndef DROPIN_LIBRARY
int main(int argc, char** argv) { return dropin_run(argc > 1 ? argv[1] : "cpu"); }
#endif
