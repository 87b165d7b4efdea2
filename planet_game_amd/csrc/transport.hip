// The implicit transport's Newton driver (heat_transport.cpp:595-772,
// solve_implicit_transport) over the device path.
//
// The caller's column response stays on the host (a callback, evaluated at
// every Newton iterate and line-search trial, as the reference does). The
// device does the work that scales with the mesh: the conservative
// diffusion K lap T(h) (bit-identical to the reference's diffusion_source),
// and the Newton system
//   (A / s) dT - K Q dT = -A F     (held cells, s <= 0: identity rows)
// assembled on the host in CSR order, set on a cached device system of the
// graph (Galerkin values and coarsest inverse on the device) and solved by
// the device PCG with the V-cycle. Merit, maximum residual and the final
// conservation sums are reduced on the host over 256-cell blocks in block
// order, as the reference reduces them, so they do not depend on anything
// but the iterates. The iterates differ from the reference's only through
// the inner solve (Jacobi-smoothed V-cycle on the device, Gauss-Seidel in
// the reference); both stop at the same Newton tolerance.
#include "system.hpp"

#include <cmath>
#include <vector>

namespace {

constexpr size_t kBlock = 256;   // the mesh's deterministic blocks

struct Driver {
    psim_graph* g;
    double k;
    const double* floor;
    psim_response_fn response;
    void* user;
    const double* latent;
    size_t n;
    std::vector<double> diffusion;

    void diffuse(const std::vector<double>& t, std::vector<double>& out) {
        out.resize(n);
        if (psim_diffusion_source(g, k, t.data(), out.data(), PSIM_HOST) != PSIM_OK)
            throw psim::DeviceFailure(psim_last_error());
    }

    template <class Body> double block_sum(const Body& body) const {
        double total = 0.0;
        for (size_t begin = 0; begin < n; begin += kBlock) total += body(begin, std::min(n, begin + kBlock));
        return total;
    }

    double max_abs(const std::vector<double>& v) const {
        double total = 0.0;
        for (size_t begin = 0; begin < n; begin += kBlock) {
            double largest = 0.0;
            for (size_t c = begin; c < std::min(n, begin + kBlock); ++c) largest = std::max(largest, std::abs(v[c]));
            total = std::max(total, largest);
        }
        return total;
    }

    // F(h) = h - K lap T(h) (+ latent) and the merit sum A F^2.
    double evaluate(const std::vector<double>& at, std::vector<double>& mean, std::vector<double>& slope,
                    std::vector<double>& residual) {
        response(user, at.data(), mean.data(), slope.data(), n);
        diffuse(mean, diffusion);
        const auto& area = g->graph.area;
        return block_sum([&](size_t begin, size_t end) {
            double merit = 0.0;
            for (size_t c = begin; c < end; ++c) {
                residual[c] = at[c] - diffusion[c];
                if (latent != nullptr) residual[c] += latent[c];
                merit += area[c] * residual[c] * residual[c];
            }
            return merit;
        });
    }

    // heat_transport.cpp:350-363 on the host, the solve on the device.
    int newton_system(const std::vector<double>& slope, const std::vector<double>& rhs, std::vector<double>& solution,
                      const psim_transport_settings& st) {
        const psim::HostGraph& hg = g->graph;
        std::vector<double> diag(n), off(hg.col.size()), b(n);
        std::vector<uint8_t> hold(n);
        for (size_t c = 0; c < n; ++c) {
            const bool free = slope[c] > 0.0;
            hold[c] = free ? 0 : 1;
            diag[c] = free ? hg.area[c] / slope[c] + k * hg.weight_sum[c] : 1.0;
            for (int32_t q = hg.row[c]; q < hg.row[c + 1]; ++q)
                off[static_cast<size_t>(q)] = free && slope[static_cast<size_t>(hg.col[static_cast<size_t>(q)])] > 0.0
                                                  ? -k * hg.weight[static_cast<size_t>(q)]
                                                  : 0.0;
            b[c] = free ? rhs[c] : 0.0;
        }
        psim_system* sys = g->newton;
        if (psim_system_set_matrix(sys, diag.data(), n, off.data(), off.size(), PSIM_HOST) != PSIM_OK)
            throw psim::DeviceFailure(psim_last_error());
        solution.resize(n);
        psim_solve_info info{};
        const int status =
            psim_system_pcg(sys, b.data(), hold.data(), solution.data(), st.cg_relative_tolerance,
                            st.max_cg_iterations, &info, PSIM_HOST);
        if (status != PSIM_OK && status != PSIM_NOT_CONVERGED) throw psim::DeviceFailure(psim_last_error());
        return info.iterations;
    }
};

}  // namespace

extern "C" {

void psim_transport_settings_default(psim_transport_settings* s) {
    if (s == nullptr) return;
    *s = psim_transport_settings{};
    s->max_newton_iterations = 30;
    s->max_line_search_halvings = 12;
    s->rounding_floor_W_m2 = 1e-3;
    s->max_non_monotone_steps = 4;
    s->max_non_monotone_growth = 100.0;
    s->newton_tolerance_W_m2 = 1e-6;
    s->cg_relative_tolerance = 1e-6;
    s->max_cg_iterations = 2000;
}

int psim_implicit_transport(psim_graph* g, double conductance_W_K, const double* source_floor_W_m2,
                            psim_response_fn response, void* user, const psim_transport_settings* settings,
                            const double* latent_W_m2, double* source_W_m2, double* mean_K,
                            psim_transport_result* result) {
    return psim::guarded([&] {
        if (g == nullptr || source_floor_W_m2 == nullptr || response == nullptr || source_W_m2 == nullptr ||
            mean_K == nullptr || result == nullptr)
            throw std::invalid_argument("null argument");
        if (!std::isfinite(conductance_W_K) || !(conductance_W_K > 0.0))
            throw std::invalid_argument("transport conductance must be finite and positive");
        psim_transport_settings st;
        if (settings != nullptr)
            st = *settings;
        else
            psim_transport_settings_default(&st);
        std::lock_guard<std::mutex> hold(g->newton_mutex);
        if (g->newton == nullptr) {
            psim_system* s = nullptr;
            if (psim_system_create(g, PSIM_PLAIN_FP64, &s) != PSIM_OK) throw psim::DeviceFailure(psim_last_error());
            g->newton = s;
        }
        const size_t n = static_cast<size_t>(g->n);
        Driver d{g, conductance_W_K, source_floor_W_m2, response, user, latent_W_m2, n, {}};
        const auto& area = g->graph.area;
        psim_transport_result res{};
        std::vector<double> source(n), mean(n, 0.0), slope(n, 0.0), residual(n, 0.0), rhs(n, 0.0), delta_mean,
            delta_source, trial(n, 0.0), trial_mean(n, 0.0), trial_slope(n, 0.0), trial_residual;
        for (size_t c = 0; c < n; ++c) source[c] = std::max(0.0, source_floor_W_m2[c]);
        double merit = d.evaluate(source, mean, slope, residual);
        int non_monotone = 0;
        for (int iteration = 0; iteration < st.max_newton_iterations; ++iteration) {
            if (d.max_abs(residual) <= st.newton_tolerance_W_m2) break;
            for (size_t c = 0; c < n; ++c) rhs[c] = -residual[c] * area[c];
            res.cg_iterations += d.newton_system(slope, rhs, delta_mean, st);
            d.diffuse(delta_mean, delta_source);
            for (size_t c = 0; c < n; ++c) delta_source[c] -= residual[c];
            // backtracking: halve the step until the merit falls
            double step = 1.0, trial_merit = 0.0;
            for (int halving = 0;; ++halving) {
                for (size_t c = 0; c < n; ++c)
                    trial[c] = std::max(source[c] + step * delta_source[c], source_floor_W_m2[c]);
                trial_residual.assign(n, 0.0);
                trial_merit = d.evaluate(trial, trial_mean, trial_slope, trial_residual);
                if (trial_merit < merit) {
                    residual = trial_residual;
                    break;
                }
                if (halving == st.max_line_search_halvings) {
                    // no descent: the rounding floor, or a kink of the response,
                    // crossed by a bounded number of full steps
                    if (d.max_abs(residual) <= st.rounding_floor_W_m2 ||
                        non_monotone == st.max_non_monotone_steps) {
                        step = 0.0;
                    } else {
                        ++non_monotone;
                        step = 1.0;
                        for (size_t c = 0; c < n; ++c)
                            trial[c] = std::max(source[c] + delta_source[c], source_floor_W_m2[c]);
                        trial_merit = d.evaluate(trial, trial_mean, trial_slope, trial_residual);
                        if (std::isfinite(trial_merit) && trial_merit <= st.max_non_monotone_growth * merit)
                            residual = trial_residual;
                        else
                            step = 0.0;
                    }
                    break;
                }
                step *= 0.5;
            }
            if (step == 0.0) break;
            std::swap(source, trial);
            std::swap(mean, trial_mean);
            std::swap(slope, trial_slope);
            merit = trial_merit;
            ++res.newton_iterations;
        }
        // the transport applied: the conservative diffusion of the last T
        std::vector<double> out;
        d.diffuse(mean, out);
        if (latent_W_m2 != nullptr)
            for (size_t c = 0; c < n; ++c) out[c] -= latent_W_m2[c];
        res.consistency_residual_W_m2 = d.max_abs(residual);
        for (size_t begin = 0; begin < n; begin += kBlock) {
            double sum = 0.0, absolute = 0.0, dissipation = 0.0;
            for (size_t c = begin; c < std::min(n, begin + kBlock); ++c) {
                const double flux = area[c] * out[c];
                sum += flux;
                absolute += std::abs(flux);
                dissipation += flux * mean[c];
            }
            res.sum_W += sum;
            res.absolute_sum_W += absolute;
            res.dissipation_W_K += dissipation;
        }
        std::copy(out.begin(), out.end(), source_W_m2);
        std::copy(mean.begin(), mean.end(), mean_K);
        *result = res;
    });
}

}  // extern "C"
