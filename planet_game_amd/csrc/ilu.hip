// Block ILU(0) (incomplete_lu.cpp:40-146) on the device, and the implicit
// tracer step built on it (solve_implicit_tracer, heat_transport.cpp:901-1081).
//
// ILU(0) restricted to fixed blocks of rows: a row keeps its (scaled)
// diagonal and the off-diagonals whose column lies in its own block, sorted
// by column; each block is factored by IKJ elimination on that pattern and
// solved by forward (unit lower) then backward substitution. The pattern,
// and for every block a level schedule of its rows, are built once on the
// host: a row's lower level is one more than the highest lower level among
// the rows it eliminates with (its lower columns), its upper level likewise
// over its upper columns. One workgroup owns one block and walks the levels
// with a workgroup barrier between them; the rows of a level are
// independent, and each row does its own operations in the reference's
// order, so factors and solves are bit-identical to the reference's.
//
// The tracer step's BiCGSTAB (heat_transport.cpp:820-897, block ILU(0) as
// the right preconditioner) runs as one cooperative launch: workgroup b
// owns block b's rows for every phase, the ILU solves need no grid barrier
// (a block's rows are its own), and five grid barriers per iteration carry
// the SpMV's neighbour reads and the dot products. Dot products reduce each
// block in a fixed order and then the block partials in block order, so
// every workgroup computes the same scalars. The stacked system's assembly
// and the flux form of the solution run on the host, in the reference's
// order.
#include "system.hpp"
#include "wave.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

namespace psim::dev {
namespace {

constexpr int kIluThreads = 256;
constexpr int kGridWords = 64;   // [0] generation, [16] and [24] arrivals by parity, [32] abort

struct IluDev {
    int32_t n, blocks;
    const int32_t* block_start;   // blocks + 1
    const int32_t* start;         // n + 1
    const int32_t* column;
    const int32_t* diag;          // position of the diagonal entry
    double* value;
    const int32_t* lrow;          // rows in (block, lower level) order
    const int32_t* llev;          // level boundaries in lrow
    const int32_t* lblk;          // blocks + 1 boundaries in llev
    const int32_t* urow;
    const int32_t* ulev;
    const int32_t* ublk;
    int32_t* status;              // bit 0: zero pivot
};

__global__ __launch_bounds__(kIluThreads) void ilu_load_kernel(int entries, const int32_t* source, const double* diag,
                                                               const double* off, double scale, double* value) {
    const int e = static_cast<int>(blockIdx.x) * kIluThreads + static_cast<int>(threadIdx.x);
    if (e >= entries) return;
    const int src = source[e];
    value[e] = src < 0 ? diag[-src - 1] * scale : off[src];
}

// IKJ elimination of row i on its pattern (incomplete_lu.cpp:109-134).
__device__ void ilu_factor_row(const IluDev& F, int i) {
    const int r0 = F.start[i], r1 = F.start[i + 1];
    for (int q = r0; q < F.diag[i]; ++q) {
        const int j = F.column[q];
        const double pivot = F.value[F.diag[j]];
        if (pivot == 0.0) {
            atomicOr(F.status, 1);
            return;
        }
        const double vq = F.value[q] / pivot;
        F.value[q] = vq;
        for (int t = F.diag[j] + 1; t < F.start[j + 1]; ++t) {
            const int c = F.column[t];
            for (int at = r0; at < r1; ++at)
                if (F.column[at] == c) {
                    F.value[at] -= vq * F.value[t];
                    break;
                }
        }
    }
}

__global__ __launch_bounds__(kIluThreads) void ilu_factor_kernel(IluDev F) {
    const int b = static_cast<int>(blockIdx.x);
    for (int lv = F.lblk[b]; lv < F.lblk[b + 1]; ++lv) {
        for (int k = F.llev[lv] + static_cast<int>(threadIdx.x); k < F.llev[lv + 1]; k += kIluThreads)
            ilu_factor_row(F, F.lrow[k]);
        __syncthreads();
    }
}

// x <- (LU)^-1 x on block b's rows (incomplete_lu.cpp:136-150). Ends with
// a workgroup barrier.
__device__ void ilu_solve_block(const IluDev& F, int b, double* x) {
    for (int lv = F.lblk[b]; lv < F.lblk[b + 1]; ++lv) {
        for (int k = F.llev[lv] + static_cast<int>(threadIdx.x); k < F.llev[lv + 1]; k += kIluThreads) {
            const int i = F.lrow[k];
            double sum = x[i];
            for (int q = F.start[i]; q < F.diag[i]; ++q) sum -= F.value[q] * x[F.column[q]];
            x[i] = sum;
        }
        __syncthreads();
    }
    for (int lv = F.ublk[b]; lv < F.ublk[b + 1]; ++lv) {
        for (int k = F.ulev[lv] + static_cast<int>(threadIdx.x); k < F.ulev[lv + 1]; k += kIluThreads) {
            const int i = F.urow[k];
            double sum = x[i];
            for (int q = F.diag[i] + 1; q < F.start[i + 1]; ++q) sum -= F.value[q] * x[F.column[q]];
            x[i] = sum / F.value[F.diag[i]];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kIluThreads) void ilu_solve_kernel(IluDev F, double* x) {
    ilu_solve_block(F, static_cast<int>(blockIdx.x), x);
}

// Static LDS of the staged cooperative solve; blocks that need more use the
// unstaged form.
constexpr size_t kIluStagedBytes = 61440;

// Block b's factors, level lists and solve vector staged in LDS for
// the cooperative solve (the pattern holds in-block columns only, so every
// index is local to the block and fits 16 bits; the host checks both).
// Layout: value[E], x[R], then column[E], start[R + 1], diag[R], lrow[R],
// urow[R] as int16.
struct IluLds {
    double* value;
    double* x;
    int16_t* column;
    int16_t* start;
    int16_t* diag;
    int16_t* lrow;
    int16_t* urow;
};

__device__ IluLds ilu_stage(const IluDev& F, int b, char* lds) {
    const int tid = static_cast<int>(threadIdx.x);
    const int first = F.block_start[b], rows = F.block_start[b + 1] - first;
    const int e0 = F.start[first], entries = F.start[first + rows] - e0;
    const int lbase = F.llev[F.lblk[b]], ubase = F.ulev[F.ublk[b]];
    IluLds L;
    L.value = reinterpret_cast<double*>(lds);
    L.x = L.value + entries;
    L.column = reinterpret_cast<int16_t*>(L.x + rows);
    L.start = L.column + entries;
    L.diag = L.start + rows + 1;
    L.lrow = L.diag + rows;
    L.urow = L.lrow + rows;
    for (int q = tid; q < entries; q += kIluThreads) {
        L.value[q] = F.value[e0 + q];
        L.column[q] = static_cast<int16_t>(F.column[e0 + q] - first);
    }
    for (int k = tid; k <= rows; k += kIluThreads) L.start[k] = static_cast<int16_t>(F.start[first + k] - e0);
    for (int k = tid; k < rows; k += kIluThreads) {
        L.diag[k] = static_cast<int16_t>(F.diag[first + k] - e0);
        L.lrow[k] = static_cast<int16_t>(F.lrow[lbase + k] - first);
        L.urow[k] = static_cast<int16_t>(F.urow[ubase + k] - first);
    }
    __syncthreads();
    return L;
}

// ilu_solve_block on the staged block: the same rows per level, the same
// operations in the same order, with every operand in LDS. x moves in and
// out of LDS around the walk. Ends with a workgroup barrier.
__device__ void ilu_solve_block_lds(const IluDev& F, const IluLds& L, int b, double* x) {
    const int tid = static_cast<int>(threadIdx.x);
    const int first = F.block_start[b], rows = F.block_start[b + 1] - first;
    for (int k = tid; k < rows; k += kIluThreads) L.x[k] = x[first + k];
    __syncthreads();
    const int lbase = F.llev[F.lblk[b]];
    for (int lv = F.lblk[b]; lv < F.lblk[b + 1]; ++lv) {
        for (int k = F.llev[lv] - lbase + tid; k < F.llev[lv + 1] - lbase; k += kIluThreads) {
            const int i = L.lrow[k];
            double sum = L.x[i];
            for (int q = L.start[i]; q < L.diag[i]; ++q) sum -= L.value[q] * L.x[L.column[q]];
            L.x[i] = sum;
        }
        __syncthreads();
    }
    const int ubase = F.ulev[F.ublk[b]];
    for (int lv = F.ublk[b]; lv < F.ublk[b + 1]; ++lv) {
        for (int k = F.ulev[lv] - ubase + tid; k < F.ulev[lv + 1] - ubase; k += kIluThreads) {
            const int i = L.urow[k];
            double sum = L.x[i];
            for (int q = L.diag[i] + 1; q < L.start[i + 1]; ++q) sum -= L.value[q] * L.x[L.column[q]];
            L.x[i] = sum / L.value[L.diag[i]];
        }
        __syncthreads();
    }
    for (int k = tid; k < rows; k += kIluThreads) x[first + k] = L.x[k];
    __syncthreads();
}

// Grid barrier of a cooperative launch. Each workgroup issues one
// agent-scope release fence and a relaxed arrival; the last arrival adds an
// acq_rel fence (ordering every arrival's writes before its own generation
// bump), clears the other episode's counter and bumps the generation the
// others poll with relaxed loads, each then issuing one acquire fence.
// Arrival counters alternate between two words by the parity of the
// generation, so the clear never races with an arrival. Waits are bounded
// by the wall clock; a timeout sets the sticky abort word, which every
// waiter also polls, so a broken launch drains at once.
//
// Words: [0] generation, [16 + 8 p] arrivals of parity p, [32] abort.
__device__ bool grid_sync(uint32_t* words, uint32_t blocks, long long limit) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        ok = 1;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t gen = __hip_atomic_load(&words[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t p = gen & 1u;
        const uint32_t arrived =
            __hip_atomic_fetch_add(&words[16 + 8 * p], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (arrived == blocks) {
            __hip_atomic_store(&words[16 + 8 * (p ^ 1u)], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            __hip_atomic_store(&words[0], gen + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            const long long t0 = static_cast<long long>(wall_clock64());
            while (__hip_atomic_load(&words[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) {
                if (__hip_atomic_load(&words[32], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                    static_cast<long long>(wall_clock64()) - t0 > limit) {
                    __hip_atomic_store(&words[32], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
    }
    __syncthreads();
    return ok != 0;
}

// Fixed-order workgroup sum of K values: wave reductions (wave_sum), then
// waves in order.
template <int K> __device__ void block_sums(double (&v)[K], double* red) {
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < K; ++k) red[k * 4 + wave] = v[k];
    __syncthreads();
    for (int k = 0; k < K; ++k) v[k] = ((red[k * 4 + 0] + red[k * 4 + 1]) + red[k * 4 + 2]) + red[k * 4 + 3];
    __syncthreads();
}

struct CsrOp {
    const int32_t* row;   // n + 1
    const int32_t* col;
    const double* diag;
    const double* off;
};

__device__ inline double apply_row(const CsrOp& A, const double* x, int i) {
    double sum = A.diag[i] * x[i];
    for (int k = A.row[i]; k < A.row[i + 1]; ++k) sum += A.off[k] * x[A.col[k]];
    return sum;
}

struct Krylov {
    const double* b;
    double *x, *r, *r0, *p, *v, *ph, *s, *sh, *t;
    double* partial;        // 2 x blocks x 2
    uint32_t* words;        // grid barrier
    double tolerance;
    int32_t max_iterations;
    int32_t* iterations;    // out
    int32_t* failed;        // out: barrier timeout
    long long limit;
};

// Each workgroup's K partials go to the parity buffer; after the grid
// barrier every workgroup adds all blocks' partials in block order.
template <int K>
__device__ bool reduce(const Krylov& S, int& parity, double (&v)[K], double* red, int blocks) {
    block_sums<K>(v, red);
    double* part = S.partial + static_cast<size_t>(parity) * blocks * 2;
    parity ^= 1;
    if (threadIdx.x == 0)
        for (int k = 0; k < K; ++k) part[blockIdx.x * 2 + k] = v[k];
    if (!grid_sync(S.words, blocks, S.limit)) return false;
    for (int k = 0; k < K; ++k) {
        double total = 0.0;
        for (int q = 0; q < blocks; ++q) total += part[q * 2 + k];
        v[k] = total;
    }
    return true;
}

// Right-preconditioned BiCGSTAB from x = 0 (heat_transport.cpp:820-897)
// with block ILU(0): the reference's recurrences and stopping tests.
template <bool kStaged> struct IluShared {
    double red[8];
};
template <> struct IluShared<true> {
    double red[8];
    alignas(16) char staged[kIluStagedBytes];
};

// kStaged: the block's factors live in LDS for the whole launch
// (ilu_stage), and each preconditioner application walks them there.
template <bool kStaged>
__global__ __launch_bounds__(kIluThreads) void ilu_bicgstab_kernel(IluDev F, CsrOp A, Krylov S) {
    __shared__ IluShared<kStaged> shared;
    double* red = shared.red;
    const int b = static_cast<int>(blockIdx.x), blocks = static_cast<int>(gridDim.x);
    IluLds staged{};
    if constexpr (kStaged) staged = ilu_stage(F, b, shared.staged);
    const auto precondition = [&](double* x) {
        if constexpr (kStaged)
            ilu_solve_block_lds(F, staged, b, x);
        else
            ilu_solve_block(F, b, x);
    };
    const int first = F.block_start[b], last = F.block_start[b + 1];
    const int tid = static_cast<int>(threadIdx.x);
    int parity = 0;
    bool ok = true;
    double d1[1];
    d1[0] = 0.0;
    for (int i = first + tid; i < last; i += kIluThreads) {
        const double bi = S.b[i];
        S.x[i] = 0.0;
        S.r[i] = bi;
        S.r0[i] = bi;
        S.p[i] = 0.0;
        S.v[i] = 0.0;
        d1[0] += bi * bi;
    }
    ok = reduce<1>(S, parity, d1, red, blocks);
    const double b_norm = sqrt(d1[0]);
    int iteration = 0;
    if (ok && b_norm > 0.0) {
        double rho = 1.0, alpha = 1.0, omega = 1.0, rho1 = d1[0];
        const double target = S.tolerance * b_norm;
        while (iteration < S.max_iterations) {
            if (rho1 == 0.0 || omega == 0.0) break;
            const double beta = rho1 / rho * (alpha / omega);
            rho = rho1;
            for (int i = first + tid; i < last; i += kIluThreads) {
                const double pi = S.r[i] + beta * (S.p[i] - omega * S.v[i]);
                S.p[i] = pi;
                S.ph[i] = pi;
            }
            __syncthreads();
            precondition(S.ph);
            if (!(ok = grid_sync(S.words, blocks, S.limit))) break;
            d1[0] = 0.0;
            for (int i = first + tid; i < last; i += kIluThreads) {
                const double vi = apply_row(A, S.ph, i);
                S.v[i] = vi;
                d1[0] += S.r0[i] * vi;
            }
            if (!(ok = reduce<1>(S, parity, d1, red, blocks))) break;
            if (d1[0] == 0.0) break;
            alpha = rho / d1[0];
            d1[0] = 0.0;
            for (int i = first + tid; i < last; i += kIluThreads) {
                const double si = S.r[i] - alpha * S.v[i];
                S.s[i] = si;
                S.sh[i] = si;
                d1[0] += si * si;
            }
            if (!(ok = reduce<1>(S, parity, d1, red, blocks))) break;
            ++iteration;
            if (sqrt(d1[0]) <= target) {
                for (int i = first + tid; i < last; i += kIluThreads) S.x[i] += alpha * S.ph[i];
                break;
            }
            precondition(S.sh);
            if (!(ok = grid_sync(S.words, blocks, S.limit))) break;
            double d2[2] = {0.0, 0.0};
            for (int i = first + tid; i < last; i += kIluThreads) {
                const double ti = apply_row(A, S.sh, i);
                S.t[i] = ti;
                d2[0] += ti * ti;
                d2[1] += ti * S.s[i];
            }
            if (!(ok = reduce<2>(S, parity, d2, red, blocks))) break;
            omega = d2[0] > 0.0 ? d2[1] / d2[0] : 0.0;
            d2[0] = d2[1] = 0.0;
            for (int i = first + tid; i < last; i += kIluThreads) {
                S.x[i] += alpha * S.ph[i] + omega * S.sh[i];
                const double ri = S.s[i] - omega * S.t[i];
                S.r[i] = ri;
                d2[0] += ri * ri;
                d2[1] += S.r0[i] * ri;
            }
            if (!(ok = reduce<2>(S, parity, d2, red, blocks))) break;
            if (sqrt(d2[0]) <= target) break;
            rho1 = d2[1];
        }
    }
    if (b == 0 && tid == 0) {
        *S.iterations = iteration;
        if (!ok) *S.failed = 1;
    }
}

}  // namespace
}  // namespace psim::dev

// ---------------------------------------------------------------- host side

struct psim_ilu {
    psim_ctx* ctx = nullptr;
    int32_t n = 0, blocks = 0, entries = 0, off_count = 0;
    size_t staged_lds = 0;   // LDS the staged cooperative solve needs (0: indices exceed 16 bits)
    psim::Arena arena;
    int32_t *block_start = nullptr, *start = nullptr, *column = nullptr, *diag = nullptr, *source = nullptr;
    double* value = nullptr;
    int32_t *lrow = nullptr, *llev = nullptr, *lblk = nullptr, *urow = nullptr, *ulev = nullptr, *ublk = nullptr;
    int32_t* status = nullptr;
    bool factored = false;
    ~psim_ilu() { arena.release(); }

    psim::dev::IluDev view() const {
        return psim::dev::IluDev{n, blocks, block_start, start, column, diag, value, lrow, llev, lblk,
                                 urow,  ulev,  ublk,  status};
    }
};

namespace {

// Rows of every block grouped by level, in (block, level, row) order.
void schedule(const std::vector<int32_t>& level, const std::vector<int32_t>& block_start, std::vector<int32_t>& rows,
              std::vector<int32_t>& levels, std::vector<int32_t>& per_block) {
    const int blocks = static_cast<int>(block_start.size()) - 1;
    rows.clear();
    levels.assign(1, 0);
    per_block.assign(1, 0);
    for (int b = 0; b < blocks; ++b) {
        int deepest = -1;
        for (int i = block_start[b]; i < block_start[b + 1]; ++i) deepest = std::max(deepest, level[i]);
        std::vector<int32_t> count(static_cast<size_t>(deepest + 2), 0);
        for (int i = block_start[b]; i < block_start[b + 1]; ++i) ++count[static_cast<size_t>(level[i]) + 1];
        for (int l = 0; l <= deepest; ++l) count[static_cast<size_t>(l) + 1] += count[static_cast<size_t>(l)];
        const size_t base = rows.size();
        rows.resize(base + static_cast<size_t>(block_start[b + 1] - block_start[b]));
        std::vector<int32_t> next(count.begin(), count.end() - 1);
        for (int i = block_start[b]; i < block_start[b + 1]; ++i) rows[base + next[static_cast<size_t>(level[i])]++] = i;
        for (int l = 0; l <= deepest; ++l)
            levels.push_back(static_cast<int32_t>(base) + count[static_cast<size_t>(l) + 1]);
        per_block.push_back(static_cast<int32_t>(levels.size()) - 1);
    }
}

std::unique_ptr<psim_ilu> make_ilu(psim_ctx* ctx, size_t n, const uint64_t* offset, const uint64_t* neighbour,
                                   const std::vector<int32_t>& block_start) {
    if (n == 0 || n >= (1ULL << 30)) throw std::invalid_argument("incomplete LU: the pattern does not match");
    if (block_start.size() < 2 || block_start.front() != 0 || static_cast<size_t>(block_start.back()) != n)
        throw std::invalid_argument("incomplete LU: the pattern does not match");
    for (size_t b = 0; b + 1 < block_start.size(); ++b)
        if (block_start[b + 1] <= block_start[b]) throw std::invalid_argument("incomplete LU: empty block");
    std::vector<int32_t> block_of(n);
    for (size_t b = 0; b + 1 < block_start.size(); ++b)
        for (int32_t i = block_start[b]; i < block_start[b + 1]; ++i) block_of[static_cast<size_t>(i)] = static_cast<int32_t>(b);
    std::vector<int32_t> start(n + 1, 0), column, diag(n), source;
    std::vector<std::pair<int32_t, int32_t>> row;   // (column, source)
    for (size_t i = 0; i < n; ++i) {
        row.assign(1, {static_cast<int32_t>(i), -static_cast<int32_t>(i) - 1});
        for (uint64_t k = offset[i]; k < offset[i + 1]; ++k) {
            if (neighbour[k] >= n) throw std::invalid_argument("incomplete LU: the pattern does not match");
            if (block_of[neighbour[k]] == block_of[i])
                row.push_back({static_cast<int32_t>(neighbour[k]), static_cast<int32_t>(k)});
        }
        std::sort(row.begin(), row.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (const auto& [c, src] : row) {
            if (c == static_cast<int32_t>(i)) diag[i] = static_cast<int32_t>(column.size());
            column.push_back(c);
            source.push_back(src);
        }
        start[i + 1] = static_cast<int32_t>(column.size());
    }
    std::vector<int32_t> lower(n, 0), upper(n, 0);
    for (size_t i = 0; i < n; ++i)
        for (int32_t q = start[i]; q < diag[i]; ++q) lower[i] = std::max(lower[i], lower[static_cast<size_t>(column[static_cast<size_t>(q)])] + 1);
    for (size_t i = n; i-- > 0;)
        for (int32_t q = diag[i] + 1; q < start[i + 1]; ++q)
            upper[i] = std::max(upper[i], upper[static_cast<size_t>(column[static_cast<size_t>(q)])] + 1);
    std::vector<int32_t> lrow, llev, lblk, urow, ulev, ublk;
    schedule(lower, block_start, lrow, llev, lblk);
    schedule(upper, block_start, urow, ulev, ublk);
    // the schedules cover every row once, inside its block, after the rows
    // it depends on (what the kernels' indexing assumes)
    const auto check_schedule = [&](const std::vector<int32_t>& rows, const std::vector<int32_t>& levels,
                                    const std::vector<int32_t>& per_block, bool lower_side) {
        if (rows.size() != n || per_block.size() != block_start.size() || per_block.back() + 1 != static_cast<int32_t>(levels.size()) ||
            levels.back() != static_cast<int32_t>(n))
            throw std::logic_error("incomplete LU: schedule does not cover the rows");
        std::vector<int32_t> at_level(n, -1);
        for (size_t b = 0; b + 1 < per_block.size(); ++b)
            for (int32_t lv = per_block[b]; lv < per_block[b + 1]; ++lv) {
                if (levels[static_cast<size_t>(lv)] > levels[static_cast<size_t>(lv) + 1])
                    throw std::logic_error("incomplete LU: schedule out of order");
                for (int32_t k = levels[static_cast<size_t>(lv)]; k < levels[static_cast<size_t>(lv) + 1]; ++k) {
                    const int32_t i = rows[static_cast<size_t>(k)];
                    if (i < block_start[b] || i >= block_start[b + 1] || at_level[static_cast<size_t>(i)] != -1)
                        throw std::logic_error("incomplete LU: schedule row outside its block");
                    at_level[static_cast<size_t>(i)] = lv;
                }
            }
        for (size_t i = 0; i < n; ++i) {
            const int32_t q0 = lower_side ? start[i] : diag[i] + 1, q1 = lower_side ? diag[i] : start[i + 1];
            for (int32_t q = q0; q < q1; ++q)
                if (at_level[static_cast<size_t>(column[static_cast<size_t>(q)])] >= at_level[i])
                    throw std::logic_error("incomplete LU: schedule breaks a dependency");
        }
    };
    check_schedule(lrow, llev, lblk, true);
    check_schedule(urow, ulev, ublk, false);
    auto f = std::make_unique<psim_ilu>();
    f->ctx = ctx;
    f->n = static_cast<int32_t>(n);
    f->blocks = static_cast<int32_t>(block_start.size()) - 1;
    f->entries = static_cast<int32_t>(column.size());
    for (size_t b = 0; b + 1 < block_start.size(); ++b) {
        const size_t rows = static_cast<size_t>(block_start[b + 1] - block_start[b]);
        const size_t ents = static_cast<size_t>(start[static_cast<size_t>(block_start[b + 1])] -
                                                start[static_cast<size_t>(block_start[b])]);
        if (rows >= 32767 || ents >= 32767) {
            f->staged_lds = 0;
            break;
        }
        const size_t bytes = 8 * (ents + rows) + 2 * (ents + 4 * rows + 1);
        f->staged_lds = std::max(f->staged_lds, (bytes + 255) / 256 * 256);
    }
    f->off_count = static_cast<int32_t>(offset[n]);
    psim::Arena& A = f->arena;
    A.plan(&f->block_start, block_start.size());
    A.plan(&f->start, start.size());
    A.plan(&f->column, column.size());
    A.plan(&f->diag, diag.size());
    A.plan(&f->source, source.size());
    A.plan(&f->value, column.size());
    A.plan(&f->lrow, lrow.size());
    A.plan(&f->llev, llev.size());
    A.plan(&f->lblk, lblk.size());
    A.plan(&f->urow, urow.size());
    A.plan(&f->ulev, ulev.size());
    A.plan(&f->ublk, ublk.size());
    A.plan(&f->status, 1);
    psim::hip_check(hipSetDevice(ctx->device), "hipSetDevice");
    A.allocate();
    const auto put = [](void* dst, const std::vector<int32_t>& v) {
        psim::hip_check(hipMemcpy(dst, v.data(), v.size() * 4, hipMemcpyHostToDevice), "incomplete LU upload");
    };
    put(f->block_start, block_start);
    put(f->start, start);
    put(f->column, column);
    put(f->diag, diag);
    put(f->source, source);
    put(f->lrow, lrow);
    put(f->llev, llev);
    put(f->lblk, lblk);
    put(f->urow, urow);
    put(f->ulev, ulev);
    put(f->ublk, ublk);
    psim::hip_check(hipMemset(f->status, 0, 4), "incomplete LU status");
    return f;
}

// Values from the operator and the factorization, on the context's stream;
// throws on a zero pivot (synchronises).
void factor(psim_ilu* f, const double* diagonal, const double* off, double scale) {
    hipStream_t stream = f->ctx->stream;
    psim::hip_check(hipMemsetAsync(f->status, 0, 4, stream), "incomplete LU status");
    hipLaunchKernelGGL(psim::dev::ilu_load_kernel, dim3((f->entries + psim::dev::kIluThreads - 1) / psim::dev::kIluThreads),
                       dim3(psim::dev::kIluThreads), 0, stream, f->entries, f->source, diagonal, off, scale, f->value);
    psim::hip_check(hipGetLastError(), "incomplete LU load");
    hipLaunchKernelGGL(psim::dev::ilu_factor_kernel, dim3(f->blocks), dim3(psim::dev::kIluThreads), 0, stream, f->view());
    psim::hip_check(hipGetLastError(), "incomplete LU factor");
    int32_t status = 0;
    psim::hip_check(hipMemcpyAsync(&status, f->status, 4, hipMemcpyDeviceToHost, stream), "incomplete LU status");
    psim::hip_check(hipStreamSynchronize(stream), "incomplete LU factor");
    f->factored = false;
    if (status != 0) throw std::runtime_error("incomplete LU: zero pivot");
    f->factored = true;
}

}  // namespace

extern "C" {

int psim_ilu_create(psim_ctx* ctx, uint64_t n, const uint64_t* offset, const uint64_t* neighbour,
                    const uint64_t* block_start, uint64_t blocks, psim_ilu** out) {
    return psim::guarded([&] {
        if (ctx == nullptr || offset == nullptr || neighbour == nullptr || block_start == nullptr || out == nullptr)
            throw std::invalid_argument("null argument");
        if (blocks == 0 || blocks > n) throw std::invalid_argument("incomplete LU: the pattern does not match");
        if (n >= (1ULL << 30)) throw std::invalid_argument("incomplete LU: the pattern does not match");
        (void)psim::checked_graph_entries(n, offset);   // rows are read by these offsets
        std::vector<int32_t> bs(blocks + 1);
        for (uint64_t b = 0; b <= blocks; ++b) {
            if (block_start[b] > n) throw std::invalid_argument("incomplete LU: the pattern does not match");
            bs[b] = static_cast<int32_t>(block_start[b]);
        }
        *out = make_ilu(ctx, n, offset, neighbour, bs).release();
    });
}

void psim_ilu_destroy(psim_ilu* f) { delete f; }

int psim_ilu_factor(psim_ilu* f, const double* diagonal, const double* off, double diagonal_scale, int where) {
    return psim::guarded([&] {
        if (f == nullptr || diagonal == nullptr || off == nullptr) throw std::invalid_argument("null argument");
        psim::check_where(where);
        psim::hip_check(hipSetDevice(f->ctx->device), "hipSetDevice");
        psim::Staging st;
        const double* d = where == PSIM_HOST ? static_cast<const double*>(st.in(diagonal, static_cast<size_t>(f->n) * 8)) : diagonal;
        const double* o = off;
        if (where == PSIM_HOST && f->off_count > 0)
            o = static_cast<const double*>(st.in(off, static_cast<size_t>(f->off_count) * 8));
        factor(f, d, o, diagonal_scale);
    });
}

int psim_ilu_solve(psim_ilu* f, double* x, int where) {
    return psim::guarded([&] {
        if (f == nullptr || x == nullptr) throw std::invalid_argument("null argument");
        psim::check_where(where);
        if (!f->factored) throw std::logic_error("incomplete LU: not factored");
        psim::hip_check(hipSetDevice(f->ctx->device), "hipSetDevice");
        hipStream_t stream = f->ctx->stream;
        const size_t bytes = static_cast<size_t>(f->n) * 8;
        psim::Staging st;
        double* d = where == PSIM_HOST ? static_cast<double*>(st.in(x, bytes)) : x;
        hipLaunchKernelGGL(psim::dev::ilu_solve_kernel, dim3(f->blocks), dim3(psim::dev::kIluThreads), 0, stream,
                           f->view(), d);
        psim::hip_check(hipGetLastError(), "incomplete LU solve");
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(x, d, bytes, hipMemcpyDeviceToHost, stream), "incomplete LU copy");
            psim::hip_check(hipStreamSynchronize(stream), "incomplete LU solve");
        }
    });
}

}  // extern "C"

// ------------------------------------------------------------- tracer step

namespace {

// The stacked system of a graph and layer count (heat_transport.cpp:
// 941-958): node-major rows a L + l; a row's entries are the node's graph
// row for its layer, then the layer below, then the layer above. Its block
// ILU pattern keeps whole columns (blocks of 256 nodes), and the device
// vectors of its BiCGSTAB.
struct Stacked {
    int32_t layers = 0, rows = 0, entries = 0;
    std::vector<uint64_t> offset, neighbour;
    std::unique_ptr<psim_ilu> ilu;
    psim::Arena arena;
    int32_t *row = nullptr, *col = nullptr;
    double *diag = nullptr, *off = nullptr, *b = nullptr;
    double *x = nullptr, *r = nullptr, *r0 = nullptr, *p = nullptr, *v = nullptr, *ph = nullptr, *s = nullptr,
           *sh = nullptr, *t = nullptr, *partial = nullptr;
    uint32_t* words = nullptr;
    int32_t* out = nullptr;   // iterations, failed
    ~Stacked() { arena.release(); }
};

struct StackedCache {
    std::vector<std::unique_ptr<Stacked>> entries;
};

Stacked& stacked_of(psim_graph* g, int32_t layers) {
    if (!g->tracer_work) g->tracer_work = std::make_shared<StackedCache>();
    auto& cache = *static_cast<StackedCache*>(g->tracer_work.get());
    for (auto& e : cache.entries)
        if (e->layers == layers) return *e;
    auto s = std::make_unique<Stacked>();
    const psim::HostGraph& hg = g->graph;
    const size_t nodes = static_cast<size_t>(hg.n), L = static_cast<size_t>(layers);
    s->layers = layers;
    s->rows = static_cast<int32_t>(nodes * L);
    s->offset.assign(nodes * L + 1, 0);
    for (size_t a = 0; a < nodes; ++a)
        for (size_t l = 0; l < L; ++l) {
            const size_t row = a * L + l;
            for (int32_t k = hg.row[a]; k < hg.row[a + 1]; ++k)
                s->neighbour.push_back(static_cast<uint64_t>(hg.col[static_cast<size_t>(k)]) * L + l);
            if (l > 0) s->neighbour.push_back(row - 1);
            if (l + 1 < L) s->neighbour.push_back(row + 1);
            s->offset[row + 1] = s->neighbour.size();
        }
    s->entries = static_cast<int32_t>(s->neighbour.size());
    std::vector<int32_t> block_start{0};
    for (size_t first = 0; first < nodes; first += 256) block_start.push_back(static_cast<int32_t>(std::min(nodes, first + 256) * L));
    s->ilu = make_ilu(g->ctx, nodes * L, s->offset.data(), s->neighbour.data(), block_start);
    const size_t rows = nodes * L, blocks = block_start.size() - 1;
    psim::Arena& A = s->arena;
    A.plan(&s->row, rows + 1);
    A.plan(&s->col, s->neighbour.size());
    A.plan(&s->diag, rows);
    A.plan(&s->off, s->neighbour.size());
    for (double** v : {&s->b, &s->x, &s->r, &s->r0, &s->p, &s->v, &s->ph, &s->s, &s->sh, &s->t}) A.plan(v, rows);
    A.plan(&s->partial, 2 * blocks * 2);
    A.plan(&s->words, psim::dev::kGridWords);
    A.plan(&s->out, 2);
    A.allocate();
    std::vector<int32_t> row32(s->offset.begin(), s->offset.end()), col32(s->neighbour.begin(), s->neighbour.end());
    psim::hip_check(hipMemcpy(s->row, row32.data(), row32.size() * 4, hipMemcpyHostToDevice), "tracer upload");
    psim::hip_check(hipMemcpy(s->col, col32.data(), col32.size() * 4, hipMemcpyHostToDevice), "tracer upload");
    psim::hip_check(hipMemset(s->words, 0, psim::dev::kGridWords * 4), "tracer barrier");
    cache.entries.push_back(std::move(s));
    return *cache.entries.back();
}

// BiCGSTAB with the block ILU on the device: b, x device vectors of the
// stacked system (x = solution). Returns the iterations.
int ilu_bicgstab(psim_ctx* ctx, Stacked& s, double tolerance, int max_iterations) {
    psim_ilu& f = *s.ilu;
    psim::dev::CsrOp A{s.row, s.col, s.diag, s.off};
    psim::dev::Krylov K{s.b,       s.x,    s.r,         s.r0,           s.p,           s.v,    s.ph,
                        s.s,       s.sh,   s.t,         s.partial,      s.words,       tolerance,
                        max_iterations, s.out, s.out + 1, static_cast<long long>(ctx->wall_clock_hz)};
    psim::dev::IluDev F = f.view();
    hipStream_t stream = ctx->stream;
    psim::hip_check(hipMemsetAsync(s.out, 0, 8, stream), "tracer result");
    void* args[] = {&F, &A, &K};
    // the staged form when the largest block's factors fit its LDS area
    const bool staged = f.staged_lds > 0 && f.staged_lds <= psim::dev::kIluStagedBytes;
    const void* kernel = staged ? reinterpret_cast<const void*>(&psim::dev::ilu_bicgstab_kernel<true>)
                                : reinterpret_cast<const void*>(&psim::dev::ilu_bicgstab_kernel<false>);
    psim::hip_check(hipLaunchCooperativeKernel(kernel, dim3(f.blocks), dim3(psim::dev::kIluThreads), args, 0, stream),
                    "tracer solve launch");
    int32_t out[2] = {0, 0};
    psim::hip_check(hipMemcpyAsync(out, s.out, 8, hipMemcpyDeviceToHost, stream), "tracer result");
    psim::hip_check(hipStreamSynchronize(stream), "tracer solve");
    if (out[1] != 0) {
        psim::hip_check(hipMemset(s.words, 0, psim::dev::kGridWords * 4), "tracer barrier reset");
        throw psim::DeviceFailure("tracer solve: grid barrier timed out");
    }
    return out[0];
}

}  // namespace

extern "C" {

int psim_tracer_step(psim_graph* g, uint32_t layers, const double* conductance_W_K, const double* outflow_kg_s,
                     double heat_capacity_J_kg_K, const double* layer_mass_kg, const double* tracer, double dt_s,
                     double relative_tolerance, int32_t max_iterations, double* tracer_out,
                     psim_tracer_result* result) {
    return psim::guarded([&] {
        if (g == nullptr || conductance_W_K == nullptr || outflow_kg_s == nullptr || layer_mass_kg == nullptr ||
            tracer == nullptr || tracer_out == nullptr || result == nullptr)
            throw std::invalid_argument("null argument");
        if (layers == 0 || layers > 256 || !(dt_s > 0.0))
            throw std::invalid_argument("tracer transport inputs do not match the graph");
        if (!(relative_tolerance >= 0.0) || max_iterations < 0) throw std::invalid_argument("invalid solver settings");
        std::lock_guard<std::mutex> hold(g->newton_mutex);
        psim::hip_check(hipSetDevice(g->ctx->device), "hipSetDevice");
        const psim::HostGraph& hg = g->graph;
        const size_t nodes = static_cast<size_t>(hg.n), E = hg.col.size(), L = layers;
        if (nodes * L >= (1ULL << 30)) throw std::invalid_argument("tracer transport too large");
        Stacked& s = stacked_of(g, static_cast<int32_t>(layers));
        const double exchange = 1.0 / (heat_capacity_J_kg_K * static_cast<double>(L));
        psim_tracer_result res{};
        // vertical fluxes W_{l+1/2} closing each layer's horizontal divergence
        std::vector<double> vertical(nodes * L, 0.0);
        for (size_t a = 0; a < nodes; ++a) {
            double w = 0.0;
            for (size_t l = 0; l < L; ++l) {
                for (int32_t k = hg.row[a]; k < hg.row[a + 1]; ++k) w -= outflow_kg_s[l * E + static_cast<size_t>(k)];
                vertical[a * L + l] = l + 1 < L ? w : 0.0;
            }
            res.column_divergence_kg_s = std::max(res.column_divergence_kg_s, std::abs(w));
        }
        // the stacked backward-Euler system (heat_transport.cpp:959-987)
        const size_t rows = nodes * L;
        std::vector<double> diag(rows, 0.0), off(s.neighbour.size(), 0.0), rhs(rows, 0.0), start(rows, 0.0);
        for (size_t a = 0; a < nodes; ++a) {
            const double rate = layer_mass_kg[a] / dt_s;
            for (size_t l = 0; l < L; ++l) {
                const size_t row = a * L + l;
                double own = rate;
                size_t at = s.offset[row];
                for (int32_t k = hg.row[a]; k < hg.row[a + 1]; ++k, ++at) {
                    const double flow = outflow_kg_s[l * E + static_cast<size_t>(k)];
                    const double eddy = conductance_W_K[static_cast<size_t>(k)] * exchange;
                    own += std::max(flow, 0.0) + eddy;
                    off[at] = -std::max(-flow, 0.0) - eddy;
                }
                const double below = l > 0 ? vertical[a * L + l - 1] : 0.0;
                const double above = vertical[row];
                if (l > 0) {
                    own += std::max(-below, 0.0);
                    off[at++] = -std::max(below, 0.0);
                }
                if (l + 1 < L) {
                    own += std::max(above, 0.0);
                    off[at++] = -std::max(-above, 0.0);
                }
                diag[row] = own;
                rhs[row] = rate * tracer[l * nodes + a];
                start[row] = tracer[l * nodes + a];
            }
        }
        const auto apply = [&](const std::vector<double>& x, std::vector<double>& y) {
            for (size_t row = 0; row < rows; ++row) {
                double sum = diag[row] * x[row];
                for (uint64_t k = s.offset[row]; k < s.offset[row + 1]; ++k) sum += off[k] * x[s.neighbour[k]];
                y[row] = sum;
            }
        };
        // solve for the change from the start: A delta = b - A c0
        std::vector<double> applied(rows, 0.0), residual(rows, 0.0), delta(rows, 0.0);
        apply(start, applied);
        for (size_t row = 0; row < rows; ++row) residual[row] = rhs[row] - applied[row];
        hipStream_t stream = g->ctx->stream;
        psim::hip_check(hipMemcpyAsync(s.diag, diag.data(), rows * 8, hipMemcpyHostToDevice, stream), "tracer upload");
        psim::hip_check(hipMemcpyAsync(s.off, off.data(), off.size() * 8, hipMemcpyHostToDevice, stream), "tracer upload");
        psim::hip_check(hipMemcpyAsync(s.b, residual.data(), rows * 8, hipMemcpyHostToDevice, stream), "tracer upload");
        factor(s.ilu.get(), s.diag, s.off, 1.0);
        res.iterations = ilu_bicgstab(g->ctx, s, relative_tolerance, max_iterations);
        psim::hip_check(hipMemcpy(delta.data(), s.x, rows * 8, hipMemcpyDeviceToHost), "tracer copy");
        std::vector<double> solved(rows, 0.0);
        for (size_t row = 0; row < rows; ++row) solved[row] = std::max(0.0, start[row] + delta[row]);
        apply(solved, applied);
        double rn = 0.0, bn = 0.0;
        for (size_t row = 0; row < rows; ++row) {
            rn += (rhs[row] - applied[row]) * (rhs[row] - applied[row]);
            bn += rhs[row] * rhs[row];
        }
        res.relative_residual = bn > 0.0 ? std::sqrt(rn / bn) : 0.0;
        // the flux form: each face once, from its lower node, and the
        // vertical fluxes (heat_transport.cpp:1041-1068)
        std::vector<double> change(rows, 0.0);
        for (size_t a = 0; a < nodes; ++a)
            for (size_t l = 0; l < L; ++l) {
                const size_t row = a * L + l;
                for (int32_t k = hg.row[a]; k < hg.row[a + 1]; ++k) {
                    const size_t b = static_cast<size_t>(hg.col[static_cast<size_t>(k)]);
                    if (b < a) continue;
                    const size_t other = b * L + l;
                    const double flow = outflow_kg_s[l * E + static_cast<size_t>(k)];
                    const double eddy = conductance_W_K[static_cast<size_t>(k)] * exchange;
                    const double flux = (flow > 0.0 ? flow * solved[row] : flow * solved[other]) +
                                        eddy * (solved[row] - solved[other]);
                    change[row] -= flux;
                    change[other] += flux;
                }
                if (l + 1 < L) {
                    const double w = vertical[row];
                    const double flux = w > 0.0 ? w * solved[row] : w * solved[row + 1];
                    change[row] -= flux;
                    change[row + 1] += flux;
                }
            }
        for (size_t a = 0; a < nodes; ++a)
            for (size_t l = 0; l < L; ++l) {
                double value = tracer[l * nodes + a] + dt_s * change[a * L + l] / layer_mass_kg[a];
                if (value < 0.0) {
                    res.clipped_kg -= value * layer_mass_kg[a];
                    value = 0.0;
                }
                tracer_out[l * nodes + a] = value;
            }
        *result = res;
    });
}

}  // extern "C"
