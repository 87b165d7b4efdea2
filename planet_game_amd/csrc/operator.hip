// Linear solves with an operator that differs from the multigrid's matrix,
// as the other GraphMultigrid consumers run them:
//   balanced_circulation.cpp:316-444   BiCGSTAB on the non-symmetric flux
//       operator; preconditioner: the V-cycle of the symmetric part, then
//       block ILU(0) of the full operator on what the V-cycle leaves;
//       breakdown and the iteration cap are errors; the true residual is
//       reported;
//   heat_transport.cpp:1176-1226       the advection-diffusion Newton
//       system: the same, with the V-cycle's input scaled by the areas and
//       its output divided by the slopes, in the bicgstab template of
//       heat_transport.cpp:820-897 (stops on breakdown or the cap).
// The operator is a diagonal plus off-diagonals in the graph's CSR order.
// The V-cycle is a psim_system (its matrix set by psim_system_set_matrix)
// run as one launch of the persistent program; the ILU(0) is a factored
// psim_ilu on the same rows. The BiCGSTAB loop runs on the host over device
// vectors: its dot products are per-256-row-block partials reduced in a
// fixed order on the device and added in block order on the host.
#include "system.hpp"
#include "wave.hpp"

#include <cmath>
#include <memory>
#include <vector>

namespace psim::dev {
namespace {

constexpr int kOpThreads = 256;

// y = D x + O x (CSR, entries in order).
__global__ __launch_bounds__(kOpThreads) void csr_apply_kernel(int n, const int32_t* row, const int32_t* col,
                                                               const double* diag, const double* off, const double* x,
                                                               double* y) {
    const int i = static_cast<int>(blockIdx.x) * kOpThreads + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    double sum = diag[i] * x[i];
    for (int k = row[i]; k < row[i + 1]; ++k) sum += off[k] * x[col[k]];
    y[i] = sum;
}

// out = a + c1 (b - c2 c)  (kind 1), a - c1 b (kind 2), a - b (kind 3),
// a + b (kind 4), a * b (kind 5), a / b (kind 6), a (kind 0).
__global__ __launch_bounds__(kOpThreads) void combine_kernel(int n, int kind, const double* a, const double* b,
                                                             const double* c, double c1, double c2, double* out) {
    const int i = static_cast<int>(blockIdx.x) * kOpThreads + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    double v = a[i];
    switch (kind) {
        case 1: v = a[i] + c1 * (b[i] - c2 * c[i]); break;
        case 2: v = a[i] - c1 * b[i]; break;
        case 3: v = a[i] - b[i]; break;
        case 4: v = a[i] + b[i]; break;
        case 5: v = b[i] * a[i]; break;
        case 6: v = a[i] / b[i]; break;
        default: break;
    }
    out[i] = v;
}

// x += c1 a + c2 b (c2 b skipped when b is null).
__global__ __launch_bounds__(kOpThreads) void update_kernel(int n, double* x, const double* a, const double* b,
                                                            double c1, double c2) {
    const int i = static_cast<int>(blockIdx.x) * kOpThreads + static_cast<int>(threadIdx.x);
    if (i >= n) return;
    x[i] += b != nullptr ? c1 * a[i] + c2 * b[i] : c1 * a[i];
}

// Per block of 256 rows: sum of u v (fixed order: wave reductions, then
// waves in order).
__global__ __launch_bounds__(kOpThreads) void dot_kernel(int n, const double* u, const double* v, double* partial) {
    __shared__ double red[4];
    const int i = static_cast<int>(blockIdx.x) * kOpThreads + static_cast<int>(threadIdx.x);
    double s = i < n ? u[i] * v[i] : 0.0;
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

}  // namespace
}  // namespace psim::dev

namespace {

struct Vectors {
    psim::Arena arena;
    int32_t n = 0, blocks = 0;
    double *diag = nullptr, *off = nullptr, *b = nullptr, *x = nullptr, *r = nullptr, *r0 = nullptr, *p = nullptr,
           *v = nullptr, *ph = nullptr, *s = nullptr, *sh = nullptr, *t = nullptr, *work = nullptr, *leftover = nullptr,
           *pre = nullptr, *post = nullptr, *partial = nullptr;
    std::vector<double> host_partial;
    ~Vectors() { arena.release(); }
};

Vectors& vectors_of(psim_graph* g) {
    if (!g->operator_work) {
        auto w = std::make_shared<Vectors>();
        const size_t n = static_cast<size_t>(g->n);
        w->n = g->n;
        w->blocks = (g->n + psim::dev::kOpThreads - 1) / psim::dev::kOpThreads;
        w->arena.plan(&w->off, static_cast<size_t>(g->nnz));
        for (double** v : {&w->diag, &w->b, &w->x, &w->r, &w->r0, &w->p, &w->v, &w->ph, &w->s, &w->sh, &w->t, &w->work,
                           &w->leftover, &w->pre, &w->post})
            w->arena.plan(v, n);
        w->arena.plan(&w->partial, static_cast<size_t>(w->blocks));
        w->arena.allocate();
        w->host_partial.resize(static_cast<size_t>(w->blocks));
        g->operator_work = w;
    }
    return *static_cast<Vectors*>(g->operator_work.get());
}

struct Solver {
    psim_graph* g;
    Vectors& w;
    const psim_preconditioner* pre;
    hipStream_t stream;
    dim3 grid;

    void apply(const double* x, double* y) const {
        hipLaunchKernelGGL(psim::dev::csr_apply_kernel, grid, dim3(psim::dev::kOpThreads), 0, stream, w.n,
                           g->views[0].row, g->csr_col, w.diag, w.off, x, y);
        psim::hip_check(hipGetLastError(), "operator apply");
    }
    void combine(int kind, const double* a, const double* b, const double* c, double c1, double c2, double* out) const {
        hipLaunchKernelGGL(psim::dev::combine_kernel, grid, dim3(psim::dev::kOpThreads), 0, stream, w.n, kind, a, b, c,
                           c1, c2, out);
        psim::hip_check(hipGetLastError(), "operator combine");
    }
    void update(double* x, const double* a, const double* b, double c1, double c2) const {
        hipLaunchKernelGGL(psim::dev::update_kernel, grid, dim3(psim::dev::kOpThreads), 0, stream, w.n, x, a, b, c1, c2);
        psim::hip_check(hipGetLastError(), "operator update");
    }
    double dot(const double* u, const double* v) {
        hipLaunchKernelGGL(psim::dev::dot_kernel, grid, dim3(psim::dev::kOpThreads), 0, stream, w.n, u, v, w.partial);
        psim::hip_check(hipGetLastError(), "operator dot");
        psim::hip_check(hipMemcpyAsync(w.host_partial.data(), w.partial, static_cast<size_t>(w.blocks) * 8,
                                       hipMemcpyDeviceToHost, stream),
                        "operator dot");
        psim::hip_check(hipStreamSynchronize(stream), "operator dot");
        double total = 0.0;
        for (const double p : w.host_partial) total += p;
        return total;
    }

    // out = M in: the V-cycle (input scaled, output divided, when given),
    // then the ILU(0) of the operator on what it leaves; without a V-cycle,
    // the ILU(0) alone.
    void precondition(const double* in, double* out) {
        if (pre->multigrid == nullptr) {
            combine(0, in, nullptr, nullptr, 0.0, 0.0, out);
            if (pre->ilu != nullptr && psim_ilu_solve(pre->ilu, out, PSIM_DEVICE) != PSIM_OK)
                throw psim::DeviceFailure(psim_last_error());
            return;
        }
        const double* v_in = in;
        if (pre->pre_scale != nullptr) {
            combine(5, in, w.pre, nullptr, 0.0, 0.0, w.work);
            v_in = w.work;
        }
        if (psim_system_precondition(pre->multigrid, v_in, out, PSIM_DEVICE) != PSIM_OK)
            throw psim::DeviceFailure(psim_last_error());
        if (pre->post_divide != nullptr) combine(6, out, w.post, nullptr, 0.0, 0.0, out);
        if (pre->ilu != nullptr) {
            apply(out, w.leftover);
            combine(3, in, w.leftover, nullptr, 0.0, 0.0, w.leftover);
            if (psim_ilu_solve(pre->ilu, w.leftover, PSIM_DEVICE) != PSIM_OK) throw psim::DeviceFailure(psim_last_error());
            combine(4, out, w.leftover, nullptr, 0.0, 0.0, out);
        }
    }
};

}  // namespace

extern "C" {

int psim_operator_solve(psim_graph* g, const double* diagonal, const double* off, const psim_preconditioner* pre,
                        int form, const double* b, double* x, double relative_tolerance, int32_t max_iterations,
                        psim_operator_info* info, int where) {
    return psim::guarded([&] {
        if (g == nullptr || diagonal == nullptr || off == nullptr || pre == nullptr || b == nullptr || x == nullptr ||
            info == nullptr)
            throw std::invalid_argument("null argument");
        psim::check_where(where);
        if (form != PSIM_FORM_TRANSPORT && form != PSIM_FORM_BALANCE) throw std::invalid_argument("unknown solve form");
        if (!(relative_tolerance >= 0.0) || max_iterations < 0) throw std::invalid_argument("invalid solver settings");
        if (pre->multigrid != nullptr && pre->multigrid->graph != g)
            throw std::invalid_argument("the multigrid belongs to another graph");
        if (pre->multigrid == nullptr && (pre->pre_scale != nullptr || pre->post_divide != nullptr))
            throw std::invalid_argument("scaling needs a multigrid");
        std::lock_guard<std::mutex> hold(g->newton_mutex);
        psim::hip_check(hipSetDevice(g->ctx->device), "hipSetDevice");
        Vectors& w = vectors_of(g);
        hipStream_t stream = g->ctx->stream;
        const size_t n = static_cast<size_t>(g->n), bytes = n * 8;
        const hipMemcpyKind kind = where == PSIM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        psim::hip_check(hipMemcpyAsync(w.diag, diagonal, bytes, kind, stream), "operator copy");
        psim::hip_check(hipMemcpyAsync(w.off, off, static_cast<size_t>(g->nnz) * 8, kind, stream), "operator copy");
        psim::hip_check(hipMemcpyAsync(w.b, b, bytes, kind, stream), "operator copy");
        if (pre->pre_scale != nullptr)
            psim::hip_check(hipMemcpyAsync(w.pre, pre->pre_scale, bytes, kind, stream), "operator copy");
        if (pre->post_divide != nullptr)
            psim::hip_check(hipMemcpyAsync(w.post, pre->post_divide, bytes, kind, stream), "operator copy");
        Solver S{g, w, pre, stream, dim3(static_cast<unsigned>(w.blocks))};
        psim_operator_info out{};
        // x = 0, r = r0 = b
        psim::hip_check(hipMemsetAsync(w.x, 0, bytes, stream), "operator init");
        psim::hip_check(hipMemsetAsync(w.p, 0, bytes, stream), "operator init");
        psim::hip_check(hipMemsetAsync(w.v, 0, bytes, stream), "operator init");
        S.combine(0, w.b, nullptr, nullptr, 0.0, 0.0, w.r);
        S.combine(0, w.b, nullptr, nullptr, 0.0, 0.0, w.r0);
        const double b_norm = std::sqrt(S.dot(w.b, w.b));
        const double target = relative_tolerance * b_norm;
        int iteration = 0;
        int status = PSIM_OK;
        if (b_norm > 0.0) {
            double rho = 1.0, alpha = 1.0, omega = 1.0, r_norm = b_norm;
            for (;;) {
                if (form == PSIM_FORM_BALANCE) {
                    if (!(r_norm > target)) break;
                    if (iteration == max_iterations) {
                        status = PSIM_NOT_CONVERGED;
                        break;
                    }
                } else if (iteration >= max_iterations) {
                    status = PSIM_NOT_CONVERGED;
                    break;
                }
                const double rho1 = S.dot(w.r0, w.r);
                if (rho1 == 0.0 || omega == 0.0) {
                    status = form == PSIM_FORM_BALANCE ? PSIM_RUNTIME_ERROR : PSIM_NOT_CONVERGED;
                    break;
                }
                const double beta = rho1 / rho * (alpha / omega);
                rho = rho1;
                S.combine(1, w.r, w.p, w.v, beta, omega, w.p);
                S.precondition(w.p, w.ph);
                S.apply(w.ph, w.v);
                const double r0v = S.dot(w.r0, w.v);
                if (r0v == 0.0 && form == PSIM_FORM_TRANSPORT) {
                    status = PSIM_NOT_CONVERGED;
                    break;
                }
                alpha = rho / r0v;
                S.combine(2, w.r, w.v, nullptr, alpha, 0.0, w.s);
                ++iteration;
                const double s_norm = std::sqrt(S.dot(w.s, w.s));
                if (s_norm <= target) {
                    S.update(w.x, w.ph, nullptr, alpha, 0.0);
                    r_norm = s_norm;
                    break;
                }
                S.precondition(w.s, w.sh);
                S.apply(w.sh, w.t);
                const double tt = S.dot(w.t, w.t);
                const double ts = S.dot(w.t, w.s);
                omega = form == PSIM_FORM_BALANCE ? ts / tt : (tt > 0.0 ? ts / tt : 0.0);
                S.update(w.x, w.ph, w.sh, alpha, omega);
                S.combine(2, w.s, w.t, nullptr, omega, 0.0, w.r);
                r_norm = std::sqrt(S.dot(w.r, w.r));
                if (form == PSIM_FORM_TRANSPORT && r_norm <= target) break;
            }
            if (form == PSIM_FORM_TRANSPORT && status == PSIM_NOT_CONVERGED && r_norm <= target) status = PSIM_OK;
        }
        // the true residual b - A x
        S.apply(w.x, w.t);
        S.combine(3, w.b, w.t, nullptr, 0.0, 0.0, w.t);
        out.iterations = iteration;
        out.status = status;
        out.relative_residual = b_norm > 0.0 ? std::sqrt(S.dot(w.t, w.t)) / b_norm : 0.0;
        psim::hip_check(
            hipMemcpyAsync(x, w.x, bytes, where == PSIM_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, stream),
            "operator copy");
        psim::hip_check(hipStreamSynchronize(stream), "operator solve");
        *info = out;
    });
}

}  // extern "C"
