// The coupled climate transport (heat_transport.hpp:140-200): the eddies'
// diffusion of the columns' theta with per-face conductances plus the
// upwind advection of each layer's dry static energy, and its implicit
// solve.
//
// advection_diffusion_source (heat_transport.cpp:774-816) is one device
// sweep, a thread per node folding its entries in CSR order and its layers
// in order, bit-identical to the reference. The Newton driver of
// solve_implicit_advection_diffusion (:1083-1318) runs on the host with the
// caller's column response as a callback (as the reference calls it); per
// Newton iteration the device factors the Jacobian's block ILU(0) (bit-
// identical to the reference's), sets the V-cycle of A / theta' + K (the
// Galerkin values and coarsest inverse of a cached system of the graph) and
// solves J delta = -F with psim_operator_solve in the bicgstab template's
// form (V-cycle on the area-scaled input, divided by the slopes, then the
// ILU of J on what is left). Merits and residual maxima are reduced on the
// host over 256-node blocks in block order, so they are the reference's
// bits for the same iterates; the iterates differ from the reference's only
// through the inner solve (Jacobi-smoothed V-cycle, other dot order).
#include "system.hpp"

#include <cmath>
#include <memory>
#include <vector>

namespace psim::dev {
namespace {

constexpr int kSourceThreads = 256;

__global__ __launch_bounds__(kSourceThreads) void advection_source_kernel(int n, int layers, int entries,
                                                                          const int32_t* row, const int32_t* col,
                                                                          const double* area, const double* conductance,
                                                                          const double* outflow, const double* latent,
                                                                          const double* mean, const double* energy,
                                                                          double* eddy_out, double* advective_out) {
    const int a = static_cast<int>(blockIdx.x) * kSourceThreads + static_cast<int>(threadIdx.x);
    if (a >= n) return;
    double eddy = 0.0, advective = 0.0;
    const double ma = mean[a];
    for (int k = row[a]; k < row[a + 1]; ++k) {
        const int b = col[k];
        eddy += conductance[k] * (mean[b] - ma);
        for (int l = 0; l < layers; ++l) {
            const double out = outflow[static_cast<size_t>(l) * entries + k];
            advective -= out * energy[static_cast<size_t>(l) * n + (out > 0.0 ? a : b)];
        }
    }
    double e = eddy / area[a];
    if (latent != nullptr) e -= latent[a];
    eddy_out[a] = e;
    advective_out[a] = advective / area[a];
}

}  // namespace
}  // namespace psim::dev

namespace {

constexpr size_t kBlock = 256;   // the graph's deterministic node blocks

// The driver's device state for a graph and layer count: the transport's
// arrays, the V-cycle system and the ILU of the Jacobian.
struct AdvectionWork {
    int32_t layers = 0;
    psim::Arena arena;
    double *conductance = nullptr, *outflow = nullptr, *latent = nullptr, *mean = nullptr, *energy = nullptr,
           *eddy = nullptr, *advective = nullptr;
    psim_system* multigrid = nullptr;
    psim_ilu* ilu = nullptr;
    ~AdvectionWork() {
        if (ilu != nullptr) psim_ilu_destroy(ilu);
        if (multigrid != nullptr) psim_system_destroy(multigrid);
        arena.release();
    }
};

AdvectionWork& work_of(psim_graph* g, int32_t layers) {
    auto* w = static_cast<AdvectionWork*>(g->advection_work.get());
    if (w != nullptr && w->layers == layers) return *w;
    auto fresh = std::make_shared<AdvectionWork>();
    const size_t n = static_cast<size_t>(g->n), E = static_cast<size_t>(g->nnz), L = static_cast<size_t>(layers);
    fresh->layers = layers;
    fresh->arena.plan(&fresh->conductance, E);
    fresh->arena.plan(&fresh->outflow, L * E);
    fresh->arena.plan(&fresh->latent, n);
    fresh->arena.plan(&fresh->mean, n);
    fresh->arena.plan(&fresh->energy, L * n);
    fresh->arena.plan(&fresh->eddy, n);
    fresh->arena.plan(&fresh->advective, n);
    fresh->arena.allocate();
    if (psim_system_create(g, PSIM_PLAIN_FP64, &fresh->multigrid) != PSIM_OK) throw psim::DeviceFailure(psim_last_error());
    const psim::HostGraph& hg = g->graph;
    std::vector<uint64_t> offset(hg.row.begin(), hg.row.end()), neighbour(hg.col.begin(), hg.col.end()), blocks{0};
    for (size_t first = 0; first < n; first += kBlock) blocks.push_back(std::min(n, first + kBlock));
    if (psim_ilu_create(g->ctx, n, offset.data(), neighbour.data(), blocks.data(), blocks.size() - 1, &fresh->ilu) !=
        PSIM_OK)
        throw psim::DeviceFailure(psim_last_error());
    g->advection_work = fresh;
    return *fresh;
}

void upload(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    psim::hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), "advection upload");
}

struct Driver {
    psim_graph* g;
    AdvectionWork& w;
    size_t n, E, L;
    bool latent;
    hipStream_t stream;

    // eddy and advective sources of (mean, energy), on the device
    void source(const std::vector<double>& mean, const std::vector<double>& energy, std::vector<double>& eddy,
                std::vector<double>& advective) {
        upload(w.mean, mean.data(), n * 8, stream);
        upload(w.energy, energy.data(), L * n * 8, stream);
        hipLaunchKernelGGL(psim::dev::advection_source_kernel,
                           dim3(static_cast<unsigned>((n + psim::dev::kSourceThreads - 1) / psim::dev::kSourceThreads)),
                           dim3(psim::dev::kSourceThreads), 0, stream, static_cast<int>(n), static_cast<int>(L),
                           static_cast<int>(E), g->views[0].row, g->csr_col, g->area, w.conductance, w.outflow,
                           latent ? w.latent : nullptr, w.mean, w.energy, w.eddy, w.advective);
        psim::hip_check(hipGetLastError(), "advection source launch");
        eddy.resize(n);
        advective.resize(n);
        psim::hip_check(hipMemcpyAsync(eddy.data(), w.eddy, n * 8, hipMemcpyDeviceToHost, stream), "advection copy");
        psim::hip_check(hipMemcpyAsync(advective.data(), w.advective, n * 8, hipMemcpyDeviceToHost, stream),
                        "advection copy");
        psim::hip_check(hipStreamSynchronize(stream), "advection source");
    }

    double max_abs(const std::vector<double>& v) const {
        double total = 0.0;
        for (size_t begin = 0; begin < n; begin += kBlock) {
            double largest = 0.0;
            for (size_t a = begin; a < std::min(n, begin + kBlock); ++a) largest = std::max(largest, std::abs(v[a]));
            total = std::max(total, largest);
        }
        return total;
    }
};

struct Iterate {
    std::vector<double> source, mean, slope, energy, energy_slope, residual;
    double merit = 0.0;
};

}  // namespace

extern "C" {

int psim_advection_source(psim_graph* g, uint32_t layers, const double* conductance_W_K, const double* outflow_kg_s,
                          const double* eddy_latent_W_m2, const double* mean_K, const double* energy_J_kg,
                          double* eddy_W_m2, double* advective_W_m2, int where) {
    return psim::guarded([&] {
        if (g == nullptr || conductance_W_K == nullptr || outflow_kg_s == nullptr || mean_K == nullptr ||
            energy_J_kg == nullptr || eddy_W_m2 == nullptr || advective_W_m2 == nullptr)
            throw std::invalid_argument("null argument");
        psim::check_where(where);
        if (layers == 0 || layers > 256) throw std::invalid_argument("advection-diffusion inputs do not match the graph");
        psim::hip_check(hipSetDevice(g->ctx->device), "hipSetDevice");
        const size_t n = static_cast<size_t>(g->n), E = static_cast<size_t>(g->nnz), L = layers;
        hipStream_t stream = g->ctx->stream;
        psim::Staging st;
        const auto in = [&](const double* p, size_t count) {
            return where == PSIM_HOST ? static_cast<const double*>(st.in(p, count * 8)) : p;
        };
        const double* c = in(conductance_W_K, E);
        const double* o = in(outflow_kg_s, L * E);
        const double* lat = eddy_latent_W_m2 != nullptr ? in(eddy_latent_W_m2, n) : nullptr;
        const double* m = in(mean_K, n);
        const double* e = in(energy_J_kg, L * n);
        double* eo = where == PSIM_HOST ? static_cast<double*>(st.out(n * 8)) : eddy_W_m2;
        double* ao = where == PSIM_HOST ? static_cast<double*>(st.out(n * 8)) : advective_W_m2;
        hipLaunchKernelGGL(psim::dev::advection_source_kernel,
                           dim3(static_cast<unsigned>((n + psim::dev::kSourceThreads - 1) / psim::dev::kSourceThreads)),
                           dim3(psim::dev::kSourceThreads), 0, stream, static_cast<int>(n), static_cast<int>(L),
                           static_cast<int>(E), g->views[0].row, g->csr_col, g->area, c, o, lat, m, e, eo, ao);
        psim::hip_check(hipGetLastError(), "advection source launch");
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(eddy_W_m2, eo, n * 8, hipMemcpyDeviceToHost, stream), "advection copy");
            psim::hip_check(hipMemcpyAsync(advective_W_m2, ao, n * 8, hipMemcpyDeviceToHost, stream), "advection copy");
            psim::hip_check(hipStreamSynchronize(stream), "advection source");
        }
    });
}

int psim_implicit_advection_diffusion(psim_graph* g, uint32_t layers, const double* conductance_W_K,
                                      const double* outflow_kg_s, const double* eddy_latent_W_m2,
                                      const double* source_floor_W_m2, psim_advection_response_fn response,
                                      void* user, const psim_transport_settings* settings,
                                      const double* initial_source_W_m2, double* source_W_m2, double* mean_K,
                                      double* energy_J_kg, psim_advection_result* result) {
    return psim::guarded([&] {
        if (g == nullptr || conductance_W_K == nullptr || outflow_kg_s == nullptr || source_floor_W_m2 == nullptr ||
            response == nullptr || source_W_m2 == nullptr || mean_K == nullptr || energy_J_kg == nullptr ||
            result == nullptr)
            throw std::invalid_argument("null argument");
        if (layers == 0 || layers > 256) throw std::invalid_argument("advection-diffusion inputs do not match the graph");
        const size_t n = static_cast<size_t>(g->n), E = static_cast<size_t>(g->nnz), L = layers;
        for (size_t k = 0; k < E; ++k)
            if (!(conductance_W_K[k] >= 0.0) || !std::isfinite(conductance_W_K[k]))
                throw std::invalid_argument("eddy conductances must be finite and non-negative");
        psim_transport_settings st;
        if (settings != nullptr)
            st = *settings;
        else
            psim_transport_settings_default(&st);
        std::lock_guard<std::mutex> hold(g->advection_mutex);
        psim::hip_check(hipSetDevice(g->ctx->device), "hipSetDevice");
        AdvectionWork& w = work_of(g, static_cast<int32_t>(layers));
        hipStream_t stream = g->ctx->stream;
        upload(w.conductance, conductance_W_K, E * 8, stream);
        upload(w.outflow, outflow_kg_s, L * E * 8, stream);
        if (eddy_latent_W_m2 != nullptr) upload(w.latent, eddy_latent_W_m2, n * 8, stream);
        Driver d{g, w, n, E, L, eddy_latent_W_m2 != nullptr, stream};
        const psim::HostGraph& hg = g->graph;
        const auto& area = hg.area;
        std::vector<double> eddy, advective;
        const auto evaluate = [&](Iterate& at) {
            at.mean.assign(n, 0.0);
            at.slope.assign(n, 0.0);
            at.energy.assign(L * n, 0.0);
            at.energy_slope.assign(L * n, 0.0);
            response(user, at.source.data(), at.mean.data(), at.slope.data(), at.energy.data(), at.energy_slope.data(),
                     n, layers);
            d.source(at.mean, at.energy, eddy, advective);
            at.residual.assign(n, 0.0);
            double merit = 0.0;
            for (size_t begin = 0; begin < n; begin += kBlock) {
                double part = 0.0;
                for (size_t a = begin; a < std::min(n, begin + kBlock); ++a) {
                    at.residual[a] = at.source[a] - (eddy[a] + advective[a]);
                    part += area[a] * at.residual[a] * at.residual[a];
                }
                merit += part;
            }
            at.merit = merit;
        };
        psim_advection_result res{};
        Iterate current, trial;
        current.source.assign(n, 0.0);
        for (size_t a = 0; a < n; ++a)
            current.source[a] = initial_source_W_m2 != nullptr ? std::max(initial_source_W_m2[a], source_floor_W_m2[a])
                                                               : std::max(0.0, source_floor_W_m2[a]);
        evaluate(current);
        constexpr double minimum_slope = 1.0e-12;
        std::vector<double> diagonal(n, 0.0), off(E, 0.0), rhs(n, 0.0), delta(n, 0.0), sym_diagonal(n, 0.0),
            sym_off(E, 0.0), divide(n, 0.0);
        int non_monotone = 0;
        for (int iteration = 0; iteration < st.max_newton_iterations; ++iteration) {
            if (d.max_abs(current.residual) <= st.newton_tolerance_W_m2) break;
            // J = I - dH/dh on the graph's pattern (heat_transport.cpp:1154-1174)
            for (size_t a = 0; a < n; ++a) {
                const double inverse_area = 1.0 / area[a];
                double own = 0.0;
                for (int32_t k = hg.row[a]; k < hg.row[a + 1]; ++k) {
                    const size_t b = static_cast<size_t>(hg.col[static_cast<size_t>(k)]);
                    const double gk = conductance_W_K[k];
                    own -= gk * current.slope[a];
                    double neighbour = gk * current.slope[b];
                    for (size_t l = 0; l < L; ++l) {
                        const double out = outflow_kg_s[l * E + static_cast<size_t>(k)];
                        if (out > 0.0)
                            own -= out * current.energy_slope[l * n + a];
                        else
                            neighbour -= out * current.energy_slope[l * n + b];
                    }
                    off[static_cast<size_t>(k)] = -neighbour * inverse_area;
                }
                diagonal[a] = 1.0 - own * inverse_area;
                rhs[a] = -current.residual[a];
            }
            // the V-cycle of A / theta' + K (:1191-1203) and the ILU of J
            for (size_t a = 0; a < n; ++a) {
                double sum = 0.0;
                for (int32_t k = hg.row[a]; k < hg.row[a + 1]; ++k) {
                    sum += conductance_W_K[k];
                    sym_off[static_cast<size_t>(k)] = -conductance_W_K[k];
                }
                divide[a] = std::max(current.slope[a], minimum_slope);
                sym_diagonal[a] = area[a] / divide[a] + sum;
            }
            if (psim_ilu_factor(w.ilu, diagonal.data(), off.data(), 1.0, PSIM_HOST) != PSIM_OK)
                throw std::runtime_error(psim_last_error());
            if (psim_system_set_matrix(w.multigrid, sym_diagonal.data(), n, sym_off.data(), E, PSIM_HOST) != PSIM_OK)
                throw psim::DeviceFailure(psim_last_error());
            const psim_preconditioner pre{w.multigrid, area.data(), divide.data(), w.ilu};
            psim_operator_info info{};
            if (psim_operator_solve(g, diagonal.data(), off.data(), &pre, PSIM_FORM_TRANSPORT, rhs.data(), delta.data(),
                                    st.cg_relative_tolerance, st.max_cg_iterations, &info, PSIM_HOST) != PSIM_OK)
                throw psim::DeviceFailure(psim_last_error());
            res.linear_iterations += info.iterations;
            // backtracking on the merit (:1229-1270)
            double step = 1.0;
            bool accepted = false;
            for (int halving = 0;; ++halving) {
                trial.source.assign(n, 0.0);
                for (size_t a = 0; a < n; ++a)
                    trial.source[a] = std::max(current.source[a] + step * delta[a], source_floor_W_m2[a]);
                evaluate(trial);
                if (trial.merit < current.merit) {
                    accepted = true;
                    break;
                }
                if (halving == st.max_line_search_halvings) {
                    if (d.max_abs(current.residual) > st.rounding_floor_W_m2 &&
                        non_monotone < st.max_non_monotone_steps) {
                        ++non_monotone;
                        for (size_t a = 0; a < n; ++a)
                            trial.source[a] = std::max(current.source[a] + delta[a], source_floor_W_m2[a]);
                        evaluate(trial);
                        accepted =
                            std::isfinite(trial.merit) && trial.merit <= st.max_non_monotone_growth * current.merit;
                    }
                    break;
                }
                step *= 0.5;
            }
            if (!accepted) break;
            std::swap(current, trial);
            ++res.newton_iterations;
        }
        // the conservative source of the last iterate (:1274-1296)
        d.source(current.mean, current.energy, eddy, advective);
        for (size_t a = 0; a < n; ++a) {
            source_W_m2[a] = eddy[a] + advective[a];
            res.sum_W += area[a] * source_W_m2[a];
            res.absolute_sum_W += area[a] * std::abs(source_W_m2[a]);
            res.eddy_absolute_sum_W += area[a] * std::abs(eddy[a]);
            res.advective_absolute_sum_W += area[a] * std::abs(advective[a]);
        }
        std::copy(current.mean.begin(), current.mean.end(), mean_K);
        std::copy(current.energy.begin(), current.energy.end(), energy_J_kg);
        res.consistency_residual_W_m2 = d.max_abs(current.residual);
        *result = res;
    });
}

}  // extern "C"
