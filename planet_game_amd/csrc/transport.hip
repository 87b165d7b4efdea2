// The implicit transport's Newton driver (heat_transport.cpp:595-772,
// solve_implicit_transport) with every field in HBM.
//
// The host runs the Newton loop and its scalar decisions; the device does
// the field work:
//   - the response T(h), s(h): a table held on the device (one kernel), or
//     the caller's host callback (the iterate goes down, T and s come back);
//   - the conservative diffusion K lap T (the bit-identical
//     diffusion_source sweep);
//   - F = h - K lap T (+ latent) with, per 256-cell block, the merit sum
//     A F^2 folded in cell order and the block's largest |F|: the host folds
//     the block partials in block order, as the reference reduces them, so
//     merit and residual norms are the reference's bits for the same
//     iterates;
//   - the Newton system (A / s) dT - K Q dT = -A F (held cells, s <= 0:
//     identity rows) assembled in place into a cached per-graph system's
//     matrix, its Galerkin values and coarsest inverse (set_matrix), and the
//     device PCG with the V-cycle;
//   - the line-search trials h = max(h + step (K lap dT - F), floor) and the
//     final conservation sums.
// Iterates differ from the reference's only through the inner solve
// (Jacobi-smoothed V-cycle here, Gauss-Seidel there); both stop at the same
// Newton tolerance.
#include "system.hpp"

#include <cmath>
#include <memory>
#include <vector>

namespace psim::dev {
namespace {

constexpr int kBlockCells = 256;   // the mesh's deterministic cell blocks

__device__ __forceinline__ double larger(double a, double b) { return a < b ? b : a; }   // std::max(a, b)

__global__ __launch_bounds__(kBlockCells) void start_source_kernel(int n, const double* floor, double* source) {
    const int c = static_cast<int>(blockIdx.x) * kBlockCells + static_cast<int>(threadIdx.x);
    if (c < n) source[c] = larger(0.0, floor[c]);
}

// Piecewise-linear response per node: `points` (h, T) pairs stored point-
// major, h increasing. The segment is the last one whose left end is <= h
// (the first below the table, the last above it).
__global__ __launch_bounds__(kBlockCells) void table_response_kernel(int n, int points, const double* table_h,
                                                                     const double* table_t, const double* at,
                                                                     double* mean, double* slope) {
    const int c = static_cast<int>(blockIdx.x) * kBlockCells + static_cast<int>(threadIdx.x);
    if (c >= n) return;
    const double h = at[c];
    const size_t stride = static_cast<size_t>(n);
    int k = 0;
    while (k + 2 < points && h >= table_h[static_cast<size_t>(k + 1) * stride + c]) ++k;
    const double h0 = table_h[static_cast<size_t>(k) * stride + c];
    const double h1 = table_h[static_cast<size_t>(k + 1) * stride + c];
    const double t0 = table_t[static_cast<size_t>(k) * stride + c];
    const double t1 = table_t[static_cast<size_t>(k + 1) * stride + c];
    const double s = (t1 - t0) / (h1 - h0);
    mean[c] = t0 + s * (h - h0);
    slope[c] = s;
}

// F = h - K lap T (+ latent); per block: sum of A F^2 in cell order and the
// largest |F| (std::max from 0, so NaN entries are passed over).
__global__ __launch_bounds__(kBlockCells) void residual_kernel(int n, const double* at, const double* diffusion,
                                                               const double* latent, const double* area,
                                                               double* residual, double* block_merit,
                                                               double* block_max) {
    __shared__ double term[kBlockCells], size[kBlockCells];
    const int base = static_cast<int>(blockIdx.x) * kBlockCells;
    const int t = static_cast<int>(threadIdx.x), c = base + t;
    if (c < n) {
        double r = at[c] - diffusion[c];
        if (latent != nullptr) r += latent[c];
        residual[c] = r;
        term[t] = area[c] * r * r;
        size[t] = fabs(r);
    }
    __syncthreads();
    if (t == 0) {
        double merit = 0.0, largest = 0.0;
        const int end = min(kBlockCells, n - base);
        for (int j = 0; j < end; ++j) {
            merit += term[j];
            largest = larger(largest, size[j]);
        }
        block_merit[blockIdx.x] = merit;
        block_max[blockIdx.x] = largest;
    }
}

// The Newton matrix of heat_transport.cpp:350-363 in CSR order, its
// right-hand side -A F and the held rows.
__global__ __launch_bounds__(kBlockCells) void newton_matrix_kernel(int n, double k, const int32_t* row,
                                                                    const int32_t* col, const double* weight,
                                                                    const double* weight_sum, const double* area,
                                                                    const double* slope, const double* residual,
                                                                    double* diag, double* off, double* b,
                                                                    uint8_t* hold) {
    const int c = static_cast<int>(blockIdx.x) * kBlockCells + static_cast<int>(threadIdx.x);
    if (c >= n) return;
    const bool free = slope[c] > 0.0;
    hold[c] = free ? 0 : 1;
    diag[c] = free ? area[c] / slope[c] + k * weight_sum[c] : 1.0;
    for (int q = row[c]; q < row[c + 1]; ++q) off[q] = free && slope[col[q]] > 0.0 ? -k * weight[q] : 0.0;
    b[c] = free ? -residual[c] * area[c] : 0.0;
}

// trial = max(h + step (K lap dT - F), floor).
__global__ __launch_bounds__(kBlockCells) void trial_kernel(int n, double step, const double* source,
                                                            const double* delta_diffusion, const double* residual,
                                                            const double* floor, double* trial) {
    const int c = static_cast<int>(blockIdx.x) * kBlockCells + static_cast<int>(threadIdx.x);
    if (c >= n) return;
    const double delta = delta_diffusion[c] - residual[c];
    trial[c] = larger(source[c] + step * delta, floor[c]);
}

// The transport applied, K lap T - latent, and per block the sums of A H,
// |A H| and A H T in cell order.
__global__ __launch_bounds__(kBlockCells) void final_sums_kernel(int n, const double* diffusion, const double* latent,
                                                                 const double* area, const double* mean, double* out,
                                                                 double* block_sums) {
    __shared__ double flux[kBlockCells], heat[kBlockCells];
    const int base = static_cast<int>(blockIdx.x) * kBlockCells;
    const int t = static_cast<int>(threadIdx.x), c = base + t;
    if (c < n) {
        double h = diffusion[c];
        if (latent != nullptr) h -= latent[c];
        out[c] = h;
        flux[t] = area[c] * h;
        heat[t] = flux[t] * mean[c];
    }
    __syncthreads();
    if (t == 0) {
        double sum = 0.0, absolute = 0.0, dissipation = 0.0;
        const int end = min(kBlockCells, n - base);
        for (int j = 0; j < end; ++j) {
            sum += flux[j];
            absolute += fabs(flux[j]);
            dissipation += heat[j];
        }
        double* out_sums = block_sums + 3 * static_cast<size_t>(blockIdx.x);
        out_sums[0] = sum;
        out_sums[1] = absolute;
        out_sums[2] = dissipation;
    }
}

}  // namespace
}  // namespace psim::dev

namespace {

using psim::dev::kBlockCells;

// Device vectors of one graph's transport solves, made on first use and
// kept with the graph (calls on a graph are serialised by its mutex).
struct Work {
    psim::Arena arena;
    int32_t n = 0, nnz = 0, blocks = 0;
    double *source = nullptr, *mean = nullptr, *slope = nullptr, *residual = nullptr;
    double *trial = nullptr, *trial_mean = nullptr, *trial_slope = nullptr, *trial_residual = nullptr;
    double *diffusion = nullptr, *delta_mean = nullptr, *delta_diffusion = nullptr, *b = nullptr;
    double* weight_sum = nullptr;
    double* partial = nullptr;     // 3 per block
    uint8_t* hold = nullptr;
    std::vector<double> host_partial;
    ~Work() { arena.release(); }
};

Work& work_of(psim_graph* g) {
    if (!g->transport_work) {
        auto w = std::make_shared<Work>();
        const size_t n = static_cast<size_t>(g->n);
        w->n = g->n;
        w->nnz = g->nnz;
        w->blocks = (g->n + kBlockCells - 1) / kBlockCells;
        for (double** v : {&w->source, &w->mean, &w->slope, &w->residual, &w->trial, &w->trial_mean, &w->trial_slope,
                           &w->trial_residual, &w->diffusion, &w->delta_mean, &w->delta_diffusion, &w->b,
                           &w->weight_sum})
            w->arena.plan(v, n);
        w->arena.plan(&w->partial, 3 * static_cast<size_t>(w->blocks));
        w->arena.plan(&w->hold, n);
        w->arena.allocate();
        psim::hip_check(hipMemcpy(w->weight_sum, g->graph.weight_sum.data(), n * 8, hipMemcpyHostToDevice),
                        "transport upload");
        w->host_partial.resize(3 * static_cast<size_t>(w->blocks));
        g->transport_work = w;
    }
    return *static_cast<Work*>(g->transport_work.get());
}

// The response of an iterate: a device table or a host callback.
struct Response {
    psim_response_fn callback = nullptr;
    void* user = nullptr;
    uint32_t points = 0;
    const double* table_h = nullptr;   // device
    const double* table_t = nullptr;
    std::vector<double> at, mean, slope;   // host buffers of the callback
};

struct Driver {
    psim_graph* g;
    Work& w;
    double k;
    const double* floor;    // device
    const double* latent;   // device or null
    Response& response;
    hipStream_t stream;
    dim3 grid, block{kBlockCells};

    void diffuse(const double* t, double* out) const {
        psim::dev::launch_diffusion(g->n, g->views[0].row, g->csr_col, g->weight, g->area, k, t, out, stream);
    }

    void respond(const double* at, double* mean, double* slope) {
        if (response.callback == nullptr) {
            hipLaunchKernelGGL(psim::dev::table_response_kernel, grid, block, 0, stream, w.n,
                               static_cast<int>(response.points), response.table_h, response.table_t, at, mean,
                               slope);
            psim::hip_check(hipGetLastError(), "response launch");
            return;
        }
        const size_t bytes = static_cast<size_t>(w.n) * 8;
        psim::hip_check(hipMemcpyAsync(response.at.data(), at, bytes, hipMemcpyDeviceToHost, stream), "iterate copy");
        psim::hip_check(hipStreamSynchronize(stream), "iterate copy");
        response.callback(response.user, response.at.data(), response.mean.data(), response.slope.data(),
                          static_cast<uint64_t>(w.n));
        psim::hip_check(hipMemcpyAsync(mean, response.mean.data(), bytes, hipMemcpyHostToDevice, stream),
                        "response copy");
        psim::hip_check(hipMemcpyAsync(slope, response.slope.data(), bytes, hipMemcpyHostToDevice, stream),
                        "response copy");
    }

    // Block partials down, folded in block order.
    void fetch_partials(size_t count) {
        psim::hip_check(hipMemcpyAsync(w.host_partial.data(), w.partial, count * 8, hipMemcpyDeviceToHost, stream),
                        "partial copy");
        psim::hip_check(hipStreamSynchronize(stream), "transport step");
    }

    // Response, diffusion and F at `at`: returns the merit, sets `largest`.
    double evaluate(const double* at, double* mean, double* slope, double* residual, double& largest) {
        respond(at, mean, slope);
        diffuse(mean, w.diffusion);
        hipLaunchKernelGGL(psim::dev::residual_kernel, grid, block, 0, stream, w.n, at, w.diffusion, latent, g->area,
                           residual, w.partial, w.partial + w.blocks);
        psim::hip_check(hipGetLastError(), "residual launch");
        fetch_partials(2 * static_cast<size_t>(w.blocks));
        double merit = 0.0;
        largest = 0.0;
        for (int b = 0; b < w.blocks; ++b) {
            merit += w.host_partial[static_cast<size_t>(b)];
            largest = std::max(largest, w.host_partial[static_cast<size_t>(w.blocks + b)]);
        }
        return merit;
    }

    // The Newton system on the graph's cached system: matrix in place,
    // set_matrix, then PCG into delta_mean. Returns the CG iterations.
    int newton_system(const double* slope, const double* residual, const psim_transport_settings& st) {
        psim_system* s = g->newton;
        hipLaunchKernelGGL(psim::dev::newton_matrix_kernel, grid, block, 0, stream, w.n, k, g->views[0].row,
                           g->csr_col, g->weight, w.weight_sum, g->area, slope, residual, s->host.diag64[0],
                           s->csr_stage, w.b, w.hold);
        psim::hip_check(hipGetLastError(), "newton matrix launch");
        psim::dev::Launch p{};
        p.op = psim::dev::kOpSetMatrix | psim::dev::kOpCsrValues;
        p.in = s->csr_stage;
        psim::sys::launch(s, p);
        s->matrix_set = true;
        s->heat_matrix = false;
        psim::dev::Launch q{};
        q.op = psim::dev::kOpPcg;
        q.tolerance = st.cg_relative_tolerance;
        q.max_iterations = st.max_cg_iterations;
        q.in = w.b;
        q.out = w.delta_mean;
        q.hold = w.hold;
        psim::sys::launch(s, q);
        return psim::sys::fetch_result(s).iterations;
    }

    void trial(double step, const double* source, const double* residual, double* out) const {
        hipLaunchKernelGGL(psim::dev::trial_kernel, grid, block, 0, stream, w.n, step, source, w.delta_diffusion,
                           residual, floor, out);
        psim::hip_check(hipGetLastError(), "trial launch");
    }

    // heat_transport.cpp:615-767 on the device vectors; source and mean are
    // left in w.diffusion (the transport applied) and w.mean.
    psim_transport_result run(const psim_transport_settings& st) {
        psim_transport_result res{};
        double *source = w.source, *mean = w.mean, *slope = w.slope, *residual = w.residual;
        double *t_source = w.trial, *t_mean = w.trial_mean, *t_slope = w.trial_slope, *t_residual = w.trial_residual;
        hipLaunchKernelGGL(psim::dev::start_source_kernel, grid, block, 0, stream, w.n, floor, source);
        psim::hip_check(hipGetLastError(), "transport start launch");
        double largest = 0.0;
        double merit = evaluate(source, mean, slope, residual, largest);
        int non_monotone = 0;
        for (int iteration = 0; iteration < st.max_newton_iterations; ++iteration) {
            if (largest <= st.newton_tolerance_W_m2) break;
            res.cg_iterations += newton_system(slope, residual, st);
            diffuse(w.delta_mean, w.delta_diffusion);
            // backtracking: halve the step until the merit falls
            double step = 1.0, trial_merit = 0.0, trial_largest = 0.0;
            for (int halving = 0;; ++halving) {
                trial(step, source, residual, t_source);
                trial_merit = evaluate(t_source, t_mean, t_slope, t_residual, trial_largest);
                if (trial_merit < merit) {
                    std::swap(residual, t_residual);
                    largest = trial_largest;
                    break;
                }
                if (halving == st.max_line_search_halvings) {
                    // no descent: the rounding floor, or a kink of the response,
                    // crossed by a bounded number of full steps
                    if (largest <= st.rounding_floor_W_m2 || non_monotone == st.max_non_monotone_steps) {
                        step = 0.0;
                    } else {
                        ++non_monotone;
                        step = 1.0;
                        trial(1.0, source, residual, t_source);
                        trial_merit = evaluate(t_source, t_mean, t_slope, t_residual, trial_largest);
                        if (std::isfinite(trial_merit) && trial_merit <= st.max_non_monotone_growth * merit) {
                            std::swap(residual, t_residual);
                            largest = trial_largest;
                        } else {
                            step = 0.0;
                        }
                    }
                    break;
                }
                step *= 0.5;
            }
            if (step == 0.0) break;
            std::swap(source, t_source);
            std::swap(mean, t_mean);
            std::swap(slope, t_slope);
            merit = trial_merit;
            ++res.newton_iterations;
        }
        // the transport applied: the conservative diffusion of the last T
        diffuse(mean, w.delta_diffusion);
        res.consistency_residual_W_m2 = largest;
        hipLaunchKernelGGL(psim::dev::final_sums_kernel, grid, block, 0, stream, w.n, w.delta_diffusion, latent,
                           g->area, mean, w.diffusion, w.partial);
        psim::hip_check(hipGetLastError(), "transport sums launch");
        fetch_partials(3 * static_cast<size_t>(w.blocks));
        for (int b = 0; b < w.blocks; ++b) {
            const double* p = &w.host_partial[3 * static_cast<size_t>(b)];
            res.sum_W += p[0];
            res.absolute_sum_W += p[1];
            res.dissipation_W_K += p[2];
        }
        final_mean = mean;
        return res;
    }

    const double* final_mean = nullptr;
};

void check_settings(double conductance) {
    if (!std::isfinite(conductance) || !(conductance > 0.0))
        throw std::invalid_argument("transport conductance must be finite and positive");
}

psim_transport_settings settings_or_default(const psim_transport_settings* settings) {
    psim_transport_settings st;
    if (settings != nullptr)
        st = *settings;
    else
        psim_transport_settings_default(&st);
    return st;
}

// Runs the driver for a graph under its mutex, staging host arrays.
void transport(psim_graph* g, double conductance, const double* floor, Response& response,
               const psim_transport_settings* settings, const double* latent, double* source_out, double* mean_out,
               psim_transport_result* result, int where) {
    const psim_transport_settings st = settings_or_default(settings);
    std::lock_guard<std::mutex> hold(g->newton_mutex);
    psim::hip_check(hipSetDevice(g->ctx->device), "hipSetDevice");
    if (g->newton == nullptr) {
        psim_system* s = nullptr;
        if (psim_system_create(g, PSIM_PLAIN_FP64, &s) != PSIM_OK) throw psim::DeviceFailure(psim_last_error());
        g->newton = s;
    }
    Work& w = work_of(g);
    hipStream_t stream = g->ctx->stream;
    const size_t bytes = static_cast<size_t>(g->n) * 8;
    psim::Staging staging;
    const double* d_floor = where == PSIM_HOST ? static_cast<const double*>(staging.in(floor, bytes)) : floor;
    const double* d_latent = latent;
    if (latent != nullptr && where == PSIM_HOST) d_latent = static_cast<const double*>(staging.in(latent, bytes));
    if (response.callback == nullptr && where == PSIM_HOST) {
        const size_t table_bytes = bytes * response.points;
        response.table_h = static_cast<const double*>(staging.in(response.table_h, table_bytes));
        response.table_t = static_cast<const double*>(staging.in(response.table_t, table_bytes));
    }
    if (response.callback != nullptr) {
        response.at.resize(static_cast<size_t>(g->n));
        response.mean.resize(static_cast<size_t>(g->n));
        response.slope.resize(static_cast<size_t>(g->n));
    }
    Driver d{g, w, conductance, d_floor, d_latent, response, stream,
             dim3(static_cast<unsigned>(w.blocks))};
    *result = d.run(st);
    const hipMemcpyKind kind = where == PSIM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
    psim::hip_check(hipMemcpyAsync(source_out, w.diffusion, bytes, kind, stream), "transport copy");
    psim::hip_check(hipMemcpyAsync(mean_out, d.final_mean, bytes, kind, stream), "transport copy");
    psim::hip_check(hipStreamSynchronize(stream), "implicit transport");
}

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
        check_settings(conductance_W_K);
        Response r;
        r.callback = response;
        r.user = user;
        transport(g, conductance_W_K, source_floor_W_m2, r, settings, latent_W_m2, source_W_m2, mean_K, result,
                  PSIM_HOST);
    });
}

int psim_implicit_transport_table(psim_graph* g, double conductance_W_K, const double* source_floor_W_m2,
                                  const psim_response_table* table, const psim_transport_settings* settings,
                                  const double* latent_W_m2, double* source_W_m2, double* mean_K,
                                  psim_transport_result* result, int where) {
    return psim::guarded([&] {
        if (g == nullptr || source_floor_W_m2 == nullptr || table == nullptr || table->source_W_m2 == nullptr ||
            table->mean_K == nullptr || source_W_m2 == nullptr || mean_K == nullptr || result == nullptr)
            throw std::invalid_argument("null argument");
        psim::check_where(where);
        check_settings(conductance_W_K);
        if (table->points < 2 || table->points > 4096) throw std::invalid_argument("response table needs 2 to 4096 points");
        Response r;
        r.points = table->points;
        r.table_h = table->source_W_m2;
        r.table_t = table->mean_K;
        transport(g, conductance_W_K, source_floor_W_m2, r, settings, latent_W_m2, source_W_m2, mean_K, result, where);
    });
}

}  // extern "C"
