// The persistent solver program for gfx950.
//
// One launch runs whole operations (set_matrix, a V-cycle, a complete
// BiCGSTAB solve with its true-residual check) for a batch of instances.
// Each instance owns a group of workgroups; the group walks the
// instance's phases (RHS assembly, SpMV, V-cycle levels, vector updates)
// tile by tile, 256 rows per tile and one row per thread, and meets at a
// group barrier between phases that read what the previous one wrote. With
// one workgroup per instance the barrier is a plain __syncthreads.
//
// The algorithm is the reference's (heat_transport.cpp:820-897 BiCGSTAB,
// right-preconditioned, from x = 0; V(1,1) aggregation multigrid of
// :288-326 with Galerkin values of :157-171), with these device forms:
//   * smoothing is Chebyshev-weighted Jacobi (weight 2/(a+b) on the
//     Gershgorin interval of D^-1 A, lower end clipped at b/4) instead of
//     block Gauss-Seidel, and the pre-smoothing, residual and restriction
//     of a level are one sweep: residuals land in LDS in aggregate-major
//     order and each coarse row sums its fine rows in ascending order;
//   * prolongation is fused with post-smoothing;
//   * the coarsest level is solved with its dense inverse (Gauss-Jordan
//     at set_matrix time) as a GEMV;
//   * inner products reduce each 256-row tile with a fixed butterfly, then
//     the tile partials in a fixed order, so every workgroup of a group
//     computes the same scalars and results do not depend on the group
//     size or on the batch;
//   * after BiCGSTAB stops, the FP64 residual b - A x is recomputed and the
//     solve restarts from x if it is above tolerance (the mixed-precision
//     form relies on this refinement).
#include "device.hpp"

namespace psim::dev {
namespace {

// ------------------------------------------------------------ reductions

__device__ inline double wave_sum(double v) {
    // xor butterfly: every lane ends with the same bits
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

struct Lds {
    double red[kPartials][4];   // per-wave partials
    double bcast[kPartials];
    double tile[kTile];         // restriction staging
    double colk[kMaxCoarse];    // sweep pivot column
};

// Dynamic LDS: the packed coarsest matrix during set_matrix launches.
__device__ inline double* dynamic_lds() {
    extern __shared__ double dyn[];
    return dyn;
}

// Sum of K per-thread values over the workgroup (one 256-row tile), fixed
// order: wave butterflies, then waves 0..3 left to right. Thread 0 returns
// the totals.
template <int K> __device__ inline void tile_sums(const double (&v)[K], double (&out)[K], Lds& lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const double w = wave_sum(v[q]);
        if (lane == 0) lds.red[q][wave] = w;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) out[q] = ((lds.red[q][0] + lds.red[q][1]) + lds.red[q][2]) + lds.red[q][3];
    }
    __syncthreads();
}

__device__ inline double tile_max(double v, Lds& lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const double w = wave_max(v);
    if (lane == 0) lds.red[0][wave] = w;
    __syncthreads();
    double out = 0.0;
    if (threadIdx.x == 0) out = fmax(fmax(lds.red[0][0], lds.red[0][1]), fmax(lds.red[0][2], lds.red[0][3]));
    __syncthreads();
    return out;
}

// Total of slot q over all tiles (fixed order), broadcast to the workgroup.
template <int K> __device__ inline void group_totals(const double* partial, int tiles, double (&out)[K], Lds& lds) {
    if (threadIdx.x < 64) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            double s = 0.0;
            for (int t = threadIdx.x; t < tiles; t += 64) s += partial[t * kPartials + q];
            s = wave_sum(s);
            if (threadIdx.x == 0) lds.bcast[q] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) out[q] = lds.bcast[q];
    __syncthreads();
}

__device__ inline double group_max(const double* partial, int tiles, Lds& lds) {
    if (threadIdx.x < 64) {
        double m = 0.0;
        for (int t = threadIdx.x; t < tiles; t += 64) m = fmax(m, partial[t * kPartials]);
        m = wave_max(m);
        if (threadIdx.x == 0) lds.bcast[0] = m;
    }
    __syncthreads();
    const double out = lds.bcast[0];
    __syncthreads();
    return out;
}

// ------------------------------------------------------------ group barrier

// Two-level counting barrier over the instance's workgroups. Workgroup g
// arrives on leaf counter g % kLeaves (separate 128-byte lines, so the
// arrivals do not serialise on one address); the last arrival of a leaf
// arrives at the root; the last root arrival bumps the generation word that
// every thread 0 polls. Writes are published with one agent-scope release
// fence per workgroup and read after one acquire fence; the poll itself
// uses relaxed (coherent, non-invalidating) loads. The wait is bounded by
// the 100 MHz wall clock (about one second); a timeout is recorded and is
// sticky, so a broken group drains within seconds and the host reports it.
//
// Barrier words (uint32, zeroed at creation, left zeroed by each episode):
//   [0]       generation
//   [32]      root arrivals
//   [64 + 32 l] arrivals of leaf l
constexpr long long kBarrierTicks = 100000000;
constexpr int kLeaves = 8;
constexpr int kBarrierWords = 64 + 32 * kLeaves;

// Phase trace: workgroup 0 of a traced instance stamps the wall clock at
// every group barrier (after it completes) into I.trace.
__device__ inline void trace_stamp(const Instance& I) {
    if (I.trace != nullptr && threadIdx.x == 0 && static_cast<int>(blockIdx.x) == I.group_first) {
        const uint32_t slot = I.trace_count[0];
        if (slot < static_cast<uint32_t>(I.trace_capacity)) I.trace[slot] = static_cast<uint64_t>(wall_clock64());
        I.trace_count[0] = slot + 1;
    }
}

__device__ void group_sync(const Instance& I, int& timeouts) {
    __syncthreads();
    const int G = I.group_size;
    if (G == 1) {
        trace_stamp(I);
        return;
    }
    if (threadIdx.x == 0 && timeouts == 0) {
        uint32_t* gen = I.barrier;
        uint32_t* root = I.barrier + 32;
        const int g = static_cast<int>(blockIdx.x) - I.group_first;
        const int leaf = g % kLeaves;
        const int leaves = G < kLeaves ? G : kLeaves;
        // workgroups on this leaf: g' = leaf, leaf + kLeaves, ... < G
        const uint32_t leaf_size = static_cast<uint32_t>((G - leaf + kLeaves - 1) / kLeaves);
        uint32_t* mine = I.barrier + 64 + 32 * leaf;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const uint32_t seen = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool last = false;
        if (__hip_atomic_fetch_add(mine, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1 == leaf_size) {
            __hip_atomic_store(mine, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_fetch_add(root, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1 ==
                static_cast<uint32_t>(leaves)) {
                __hip_atomic_store(root, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(gen, seen + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                last = true;
            }
        }
        if (!last) {
            const long long start = wall_clock64();
            while (__hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == seen) {
                __builtin_amdgcn_s_sleep(1);
                if (wall_clock64() - start > kBarrierTicks) {
                    timeouts = 1;
                    break;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    trace_stamp(I);
    __syncthreads();
}

// ------------------------------------------------------------ typed views

template <class VT> struct Values;
template <> struct Values<double> {
    const Instance& I;
    __device__ const double* diag(int j) const { return I.diag64[j]; }
    __device__ const double* dinv(int j) const { return I.dinv64[j]; }
    __device__ const double* off(int j) const { return I.off64[j]; }
};
template <> struct Values<float> {
    const Instance& I;
    __device__ const float* diag(int j) const { return I.diag32[j]; }
    __device__ const float* dinv(int j) const { return I.dinv32[j]; }
    __device__ const float* off(int j) const { return I.off32[j]; }
};
template <class VT> __device__ inline Values<VT> values_of(const Instance& I) { return Values<VT>{I}; }

// ------------------------------------------------------------ V-cycle phases

// Pre-smoothing from zero, residual and restriction of level j in one sweep
// over the level's restriction tiles.
template <class VT>
__device__ void down_phase(const Instance& I, const Values<VT>& V, int j, const VT* rhs, VT* coarse_rhs, int g, int G,
                           Lds& lds) {
    const LevelView& L = I.lv[j];
    const VT om = static_cast<VT>(I.omega[j]);
    const VT* diag = V.diag(j);
    const VT* dinv = V.dinv(j);
    const VT* off = V.off(j);
    for (int tile = g; tile < L.down_tiles; tile += G) {
        const int c0 = L.tile_start[tile], c1 = L.tile_start[tile + 1];
        const int base = L.agg_start[c0], count = L.agg_start[c1] - base;
        if (static_cast<int>(threadIdx.x) < count) {
            const int i = L.agg_rows[base + threadIdx.x];
            const VT xi = om * dinv[i] * rhs[i];
            VT acc = rhs[i] - diag[i] * xi;
            for (int k = L.row[i]; k < L.row[i + 1]; ++k) {
                const int c = L.col[k];
                acc -= off[k] * (om * dinv[c] * rhs[c]);
            }
            lds.tile[threadIdx.x] = static_cast<double>(acc);
        }
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < c1 - c0) {
            const int c = c0 + threadIdx.x;
            double s = 0.0;
            for (int q = L.agg_start[c] - base; q < L.agg_start[c + 1] - base; ++q) s += lds.tile[q];
            coarse_rhs[c] = static_cast<VT>(s);
        }
        __syncthreads();
    }
}

// Dense solve of the coarsest level with its inverse (FP64 GEMV).
template <class VT>
__device__ void coarse_phase(const Instance& I, int j, const VT* rhs, VT* out, int g, int G) {
    const int n = I.lv[j].n;
    const double* inv = I.cinv;
    const int tiles = (n + kTile - 1) / kTile;
    for (int tile = g; tile < tiles; tile += G) {
        const int i = tile * kTile + threadIdx.x;
        if (i < n) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += inv[static_cast<size_t>(k) * n + i] * static_cast<double>(rhs[k]);
            out[i] = static_cast<VT>(s);
        }
    }
}

// Prolongation of the coarse correction fused with post-smoothing.
template <class VT>
__device__ void up_phase(const Instance& I, const Values<VT>& V, int j, const VT* rhs, const VT* corr, VT* out, int g,
                         int G) {
    const LevelView& L = I.lv[j];
    const VT om = static_cast<VT>(I.omega[j]);
    const VT* diag = V.diag(j);
    const VT* dinv = V.dinv(j);
    const VT* off = V.off(j);
    const int tiles = (L.n + kTile - 1) / kTile;
    for (int tile = g; tile < tiles; tile += G) {
        const int i = tile * kTile + threadIdx.x;
        if (i < L.n) {
            const VT x2 = om * dinv[i] * rhs[i] + corr[L.agg[i]];
            VT acc = rhs[i] - diag[i] * x2;
            for (int k = L.row[i]; k < L.row[i + 1]; ++k) {
                const int c = L.col[k];
                acc -= off[k] * (om * dinv[c] * rhs[c] + corr[L.agg[c]]);
            }
            out[i] = x2 + om * dinv[i] * acc;
        }
    }
}

template <class VT>
__device__ void vcycle(const Instance& I, const Values<VT>& V, const VT* rhs0, VT* out0, int g, int G, Lds& lds,
                       int& timeouts) {
    const int J = I.depth;
    if (J == 1) {
        coarse_phase<VT>(I, 0, rhs0, out0, g, G);
        group_sync(I, timeouts);
        return;
    }
    for (int j = 0; j + 1 < J; ++j) {
        const VT* rhs = j == 0 ? rhs0 : static_cast<const VT*>(I.lvl_b[j]);
        down_phase<VT>(I, V, j, rhs, static_cast<VT*>(I.lvl_b[j + 1]), g, G, lds);
        group_sync(I, timeouts);
    }
    coarse_phase<VT>(I, J - 1, static_cast<const VT*>(I.lvl_b[J - 1]), static_cast<VT*>(I.lvl_e[J - 1]), g, G);
    group_sync(I, timeouts);
    for (int j = J - 2; j >= 0; --j) {
        const VT* rhs = j == 0 ? rhs0 : static_cast<const VT*>(I.lvl_b[j]);
        VT* out = j == 0 ? out0 : static_cast<VT*>(I.lvl_e[j]);
        up_phase<VT>(I, V, j, rhs, static_cast<const VT*>(I.lvl_e[j + 1]), out, g, G);
        group_sync(I, timeouts);
    }
}

// ------------------------------------------------------------ set_matrix

// Inverse of the coarsest level by the symmetric sweep operator (one
// workgroup). SWEEP(k) maps a symmetric A to B with
//   B_kk = -1/A_kk, B_ik = A_ik / A_kk, B_ij = A_ij - A_ik A_kj / A_kk,
// keeps symmetry, and sweeping every pivot of an SPD matrix gives -A^-1
// (no pivoting needed). Only the packed lower triangle is stored, in
// `tri` (LDS when it fits, else global scratch); the result is written
// negated and fully expanded, column k contiguous, to I.cinv.
__device__ inline int tri_index(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

__device__ void coarse_inverse(const Instance& I, double* tri, Lds& lds) {
    const int j = I.depth - 1;
    const LevelView& L = I.lv[j];
    const int n = L.n;
    const int packed = n * (n + 1) / 2;
    const double* diag = I.diag64[j];
    const double* off = I.off64[j];
    for (int e = threadIdx.x; e < packed; e += blockDim.x) tri[e] = 0.0;
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        tri[tri_index(r, r)] = diag[r];
        for (int k = L.row[r]; k < L.row[r + 1]; ++k)
            if (L.col[k] < r) tri[tri_index(r, L.col[k])] = off[k];
    }
    __syncthreads();
    // each thread owns a contiguous run of packed entries
    const int per = (packed + blockDim.x - 1) / blockDim.x;
    const int e0 = min(packed, static_cast<int>(threadIdx.x) * per), e1 = min(packed, e0 + per);
    int i0 = static_cast<int>((sqrt(8.0 * e0 + 1.0) - 1.0) * 0.5);
    while (i0 * (i0 + 1) / 2 > e0) --i0;
    while ((i0 + 1) * (i0 + 2) / 2 <= e0) ++i0;
    const int j0 = e0 - i0 * (i0 + 1) / 2;
    for (int k = 0; k < n; ++k) {
        for (int c = threadIdx.x; c < n; c += blockDim.x) lds.colk[c] = tri[tri_index(c, k)];
        __syncthreads();
        const double d = lds.colk[k];
        const double dinv = 1.0 / d;
        int i = i0, jj = j0;
        for (int e = e0; e < e1; ++e) {
            if (i == k && jj == k)
                tri[e] = -dinv;
            else if (i == k)
                tri[e] = lds.colk[jj] * dinv;
            else if (jj == k)
                tri[e] = lds.colk[i] * dinv;
            else
                tri[e] -= lds.colk[i] * lds.colk[jj] * dinv;
            if (++jj > i) {
                ++i;
                jj = 0;
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int col = e / n, row = e - col * n;
        I.cinv[e] = -tri[tri_index(row, col)];
    }
}

__device__ void set_matrix(const Instance& I, const Launch& P, int g, int G, Lds& lds, int& timeouts, int& parity) {
    const int J = I.depth;
    if (P.op & kOpHeatAssemble) {
        // diag = C A + dt sum_k G_k, off_k = -dt G_k, G_k = K_k w_k.
        const LevelView& L = I.lv[0];
        for (int tile = g; tile < I.tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            if (i < I.n) {
                double s = 0.0;
                for (int k = L.row[i]; k < L.row[i + 1]; ++k) {
                    const double gk = P.entry_conductance[k] * I.weight[k];
                    s += gk;
                    I.off64[0][k] = -P.dt * gk;
                }
                I.diag64[0][i] = P.capacity_area[i] + P.dt * s;
            }
        }
        group_sync(I, timeouts);
    }
    for (int j = 0; j < J; ++j) {
        const LevelView& L = I.lv[j];
        double* part = I.partial + parity * I.tiles * kPartials;
        parity ^= 1;
        const int tiles = (L.n + kTile - 1) / kTile;
        // inverse diagonal, FP32 copies, Gershgorin ratio sum |off| / diag
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            double ratio = 0.0;
            if (i < L.n) {
                const double d = I.diag64[j][i];
                const double di = 1.0 / d;
                I.dinv64[j][i] = di;
                double s = 0.0;
                for (int k = L.row[i]; k < L.row[i + 1]; ++k) {
                    s += fabs(I.off64[j][k]);
                    if (I.off32[j] != nullptr) I.off32[j][k] = static_cast<float>(I.off64[j][k]);
                }
                ratio = s * di;
                if (I.diag32[j] != nullptr) {
                    I.diag32[j][i] = static_cast<float>(d);
                    I.dinv32[j][i] = static_cast<float>(di);
                }
            }
            const double m = tile_max(ratio, lds);
            if (threadIdx.x == 0) part[tile * kPartials] = m;
        }
        // Galerkin values of level j+1 from level j's, in the reference's order
        if (j + 1 < J) {
            const int nc = L.n_next, zc = I.lv[j + 1].nnz;
            const int work_tiles = (nc + zc + kTile - 1) / kTile;
            for (int tile = g; tile < work_tiles; tile += G) {
                const int e = tile * kTile + threadIdx.x;
                if (e < nc) {
                    double s = 0.0;
                    for (int q = L.gd_start[e]; q < L.gd_start[e + 1]; ++q) {
                        const int code = L.gd[q];
                        s += code < 0 ? I.diag64[j][-code - 1] : I.off64[j][code];
                    }
                    I.diag64[j + 1][e] = s;
                } else if (e < nc + zc) {
                    const int t = e - nc;
                    double s = 0.0;
                    for (int q = L.go_start[t]; q < L.go_start[t + 1]; ++q) s += I.off64[j][L.go[q]];
                    I.off64[j + 1][t] = s;
                }
            }
        }
        group_sync(I, timeouts);
        const double rho = group_max(part, tiles, lds);
        // Chebyshev degree-1 weight on [a, b] = [max(1 - rho, (1 + rho) / 4), 1 + rho]
        const double hi = 1.0 + rho;
        const double lo = fmax(1.0 - rho, 0.25 * hi);
        if (g == 0 && threadIdx.x == 0) I.omega[j] = 2.0 / (lo + hi);
    }
    if (g == 0) coarse_inverse(I, I.tri_lds ? dynamic_lds() : I.tri, lds);
    group_sync(I, timeouts);
}

// ------------------------------------------------------------ BiCGSTAB

template <class VT>
__device__ void solve(const Instance& I, const Launch& P, int g, int G, Lds& lds, int& timeouts, int& parity) {
    const Values<VT> V = values_of<VT>(I);
    const LevelView& L0 = I.lv[0];
    const int n = I.n, tiles = I.tiles;
    VT* r = static_cast<VT*>(I.r);
    VT* r0 = static_cast<VT*>(I.r0);
    VT* p = static_cast<VT*>(I.p);
    VT* v = static_cast<VT*>(I.v);
    VT* ph = static_cast<VT*>(I.ph);
    VT* s = static_cast<VT*>(I.s);
    VT* sh = static_cast<VT*>(I.sh);
    VT* t = static_cast<VT*>(I.t);
    double* x = I.x;
    double* b = I.b;
    const VT* diag = V.diag(0);
    const VT* off = V.off(0);
    const auto next_partials = [&]() {
        double* part = I.partial + parity * tiles * kPartials;
        parity ^= 1;
        return part;
    };

    // b (assembled or given), x = 0, r = r0 = b, p = v = 0.
    double* part = next_partials();
    for (int tile = g; tile < tiles; tile += G) {
        const int i = tile * kTile + threadIdx.x;
        double q[2] = {0.0, 0.0};
        if (i < n) {
            double bi;
            if (P.op & kOpHeatRhs)
                bi = P.capacity_area[i] * P.t_old[i] + P.dt * I.area[i] * P.forcing[i];
            else
                bi = P.in[i];
            b[i] = bi;
            x[i] = 0.0;
            const VT ri = static_cast<VT>(bi);
            r[i] = ri;
            r0[i] = ri;
            p[i] = VT(0);
            v[i] = VT(0);
            q[0] = bi * bi;
            q[1] = static_cast<double>(ri) * static_cast<double>(ri);
        }
        double tot[2];
        tile_sums<2>(q, tot, lds);
        if (threadIdx.x == 0) {
            part[tile * kPartials] = tot[0];
            part[tile * kPartials + 1] = tot[1];
        }
    }
    group_sync(I, timeouts);
    double tot2[2];
    group_totals<2>(part, tiles, tot2, lds);
    const double b_norm = sqrt(tot2[0]);
    double rho1 = tot2[1];
    double r_norm = b_norm;
    int iterations = 0, restarts = 0, v_cycles = 0;
    int status = 0;
    const double target = P.tolerance * b_norm;
    bool first = true;

    if (b_norm > 0.0) {
        for (;;) {
            if (!first) {
                // FP64 true residual r = b - A x; restart from x if needed.
                part = next_partials();
                for (int tile = g; tile < tiles; tile += G) {
                    const int i = tile * kTile + threadIdx.x;
                    double q[2] = {0.0, 0.0};
                    if (i < n) {
                        double ax = I.diag64[0][i] * x[i];
                        for (int k = L0.row[i]; k < L0.row[i + 1]; ++k) ax += I.off64[0][k] * x[L0.col[k]];
                        const double ri = b[i] - ax;
                        const VT rv = static_cast<VT>(ri);
                        r[i] = rv;
                        r0[i] = rv;
                        p[i] = VT(0);
                        v[i] = VT(0);
                        q[0] = ri * ri;
                        q[1] = static_cast<double>(rv) * static_cast<double>(rv);
                    }
                    double tot[2];
                    tile_sums<2>(q, tot, lds);
                    if (threadIdx.x == 0) {
                        part[tile * kPartials] = tot[0];
                        part[tile * kPartials + 1] = tot[1];
                    }
                }
                group_sync(I, timeouts);
                group_totals<2>(part, tiles, tot2, lds);
                r_norm = sqrt(tot2[0]);
                rho1 = tot2[1];
                if (r_norm <= target) break;
                if (iterations >= P.max_iterations || restarts >= P.max_restarts) {
                    status = kNotConverged;
                    break;
                }
                ++restarts;
            }
            first = false;
            const double inner_target = fmax(target, P.inner_relative * r_norm);
            double rho = 1.0, alpha = 1.0, omega = 1.0;
            while (iterations < P.max_iterations) {
                if (rho1 == 0.0 || omega == 0.0) break;
                const double beta = rho1 / rho * (alpha / omega);
                rho = rho1;
                for (int tile = g; tile < tiles; tile += G) {
                    const int i = tile * kTile + threadIdx.x;
                    if (i < n)
                        p[i] = static_cast<VT>(static_cast<double>(r[i]) +
                                               beta * (static_cast<double>(p[i]) - omega * static_cast<double>(v[i])));
                }
                group_sync(I, timeouts);
                vcycle<VT>(I, V, p, ph, g, G, lds, timeouts);
                ++v_cycles;
                // v = A ph, (r0, v)
                part = next_partials();
                for (int tile = g; tile < tiles; tile += G) {
                    const int i = tile * kTile + threadIdx.x;
                    double q[1] = {0.0};
                    if (i < n) {
                        VT acc = diag[i] * ph[i];
                        for (int k = L0.row[i]; k < L0.row[i + 1]; ++k) acc += off[k] * ph[L0.col[k]];
                        v[i] = acc;
                        q[0] = static_cast<double>(r0[i]) * static_cast<double>(acc);
                    }
                    double tot[1];
                    tile_sums<1>(q, tot, lds);
                    if (threadIdx.x == 0) part[tile * kPartials] = tot[0];
                }
                group_sync(I, timeouts);
                double r0v[1];
                group_totals<1>(part, tiles, r0v, lds);
                if (r0v[0] == 0.0) break;
                alpha = rho / r0v[0];
                // s = r - alpha v, (s, s)
                part = next_partials();
                for (int tile = g; tile < tiles; tile += G) {
                    const int i = tile * kTile + threadIdx.x;
                    double q[1] = {0.0};
                    if (i < n) {
                        const double si = static_cast<double>(r[i]) - alpha * static_cast<double>(v[i]);
                        const VT sv = static_cast<VT>(si);
                        s[i] = sv;
                        q[0] = static_cast<double>(sv) * static_cast<double>(sv);
                    }
                    double tot[1];
                    tile_sums<1>(q, tot, lds);
                    if (threadIdx.x == 0) part[tile * kPartials] = tot[0];
                }
                group_sync(I, timeouts);
                double ss[1];
                group_totals<1>(part, tiles, ss, lds);
                ++iterations;
                if (sqrt(ss[0]) <= inner_target) {
                    for (int tile = g; tile < tiles; tile += G) {
                        const int i = tile * kTile + threadIdx.x;
                        if (i < n) x[i] += alpha * static_cast<double>(ph[i]);
                    }
                    group_sync(I, timeouts);
                    break;
                }
                vcycle<VT>(I, V, s, sh, g, G, lds, timeouts);
                ++v_cycles;
                // t = A sh, (t, t), (t, s)
                part = next_partials();
                for (int tile = g; tile < tiles; tile += G) {
                    const int i = tile * kTile + threadIdx.x;
                    double q[2] = {0.0, 0.0};
                    if (i < n) {
                        VT acc = diag[i] * sh[i];
                        for (int k = L0.row[i]; k < L0.row[i + 1]; ++k) acc += off[k] * sh[L0.col[k]];
                        t[i] = acc;
                        const double td = static_cast<double>(acc);
                        q[0] = td * td;
                        q[1] = td * static_cast<double>(s[i]);
                    }
                    double tot[2];
                    tile_sums<2>(q, tot, lds);
                    if (threadIdx.x == 0) {
                        part[tile * kPartials] = tot[0];
                        part[tile * kPartials + 1] = tot[1];
                    }
                }
                group_sync(I, timeouts);
                double tt[2];
                group_totals<2>(part, tiles, tt, lds);
                omega = tt[0] > 0.0 ? tt[1] / tt[0] : 0.0;
                // x += alpha ph + omega sh, r = s - omega t, (r, r), (r0, r)
                part = next_partials();
                for (int tile = g; tile < tiles; tile += G) {
                    const int i = tile * kTile + threadIdx.x;
                    double q[2] = {0.0, 0.0};
                    if (i < n) {
                        x[i] += alpha * static_cast<double>(ph[i]) + omega * static_cast<double>(sh[i]);
                        const double ri = static_cast<double>(s[i]) - omega * static_cast<double>(t[i]);
                        const VT rv = static_cast<VT>(ri);
                        r[i] = rv;
                        q[0] = static_cast<double>(rv) * static_cast<double>(rv);
                        q[1] = static_cast<double>(r0[i]) * static_cast<double>(rv);
                    }
                    double tot[2];
                    tile_sums<2>(q, tot, lds);
                    if (threadIdx.x == 0) {
                        part[tile * kPartials] = tot[0];
                        part[tile * kPartials + 1] = tot[1];
                    }
                }
                group_sync(I, timeouts);
                double rr[2];
                group_totals<2>(part, tiles, rr, lds);
                rho1 = rr[1];
                if (sqrt(rr[0]) <= inner_target) break;
            }
        }
    }
    if (P.op & kOpHeatRhs) {
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            if (i < n) P.t_new[i] = x[i];
        }
    }
    if (g == 0 && threadIdx.x == 0) {
        I.result->iterations = iterations;
        I.result->restarts = restarts;
        I.result->status = status;
        I.result->relative_residual = b_norm > 0.0 ? r_norm / b_norm : 0.0;
        I.result->b_norm = b_norm;
        I.result->v_cycles = v_cycles;
    }
}

// ------------------------------------------------------------ the kernel

__global__ __launch_bounds__(kTile) void program_kernel(const Instance* instances, const int32_t* block_instance,
                                                        const Launch* launches, Launch single) {
    __shared__ Lds lds;
    __shared__ Launch params;
    const int index = block_instance[blockIdx.x];
    const Instance& I = instances[index];
    if (threadIdx.x == 0) params = launches != nullptr ? launches[index] : single;
    __syncthreads();
    const Launch& P = params;
    const int g = static_cast<int>(blockIdx.x) - I.group_first;
    const int G = I.group_size;
    int timeouts = 0;
    int parity = 0;
    if (I.trace != nullptr && threadIdx.x == 0 && g == 0) {
        I.trace_count[0] = 1;
        I.trace[0] = static_cast<uint64_t>(wall_clock64());
    }
    if (P.op & kOpSetMatrix) set_matrix(I, P, g, G, lds, timeouts, parity);
    if (P.op & kOpPrecondition) {
        if (I.precision == 0) {
            vcycle<double>(I, values_of<double>(I), P.in, P.out, g, G, lds, timeouts);
        } else {
            float* pin = static_cast<float*>(I.p);
            float* pout = static_cast<float*>(I.ph);
            for (int tile = g; tile < I.tiles; tile += G) {
                const int i = tile * kTile + threadIdx.x;
                if (i < I.n) pin[i] = static_cast<float>(P.in[i]);
            }
            group_sync(I, timeouts);
            vcycle<float>(I, values_of<float>(I), pin, pout, g, G, lds, timeouts);
            for (int tile = g; tile < I.tiles; tile += G) {
                const int i = tile * kTile + threadIdx.x;
                if (i < I.n) P.out[i] = static_cast<double>(pout[i]);
            }
        }
    }
    if (P.op & kOpApply) {
        const LevelView& L = I.lv[0];
        for (int tile = g; tile < I.tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            if (i < I.n) {
                double acc = I.diag64[0][i] * P.in[i];
                for (int k = L.row[i]; k < L.row[i + 1]; ++k) acc += I.off64[0][k] * P.in[L.col[k]];
                P.out[i] = acc;
            }
        }
    }
    if (P.op & kOpSolve) {
        if (I.precision == 0)
            solve<double>(I, P, g, G, lds, timeouts, parity);
        else
            solve<float>(I, P, g, G, lds, timeouts, parity);
    }
    if (timeouts != 0 && threadIdx.x == 0) atomicAdd(&I.result->barrier_timeouts, timeouts);
}

}  // namespace

hipError_t launch_program(const Instance* instances, const int32_t* block_instance, const Launch* launches,
                          const Launch& single, int32_t blocks, size_t dynamic_lds, hipStream_t stream,
                          bool cooperative) {
    if (dynamic_lds > 0) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&program_kernel),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 static_cast<int>(dynamic_lds));
        if (e != hipSuccess) return e;
    }
    if (cooperative) {
        Launch copy = single;
        void* args[] = {const_cast<Instance**>(&instances), const_cast<int32_t**>(&block_instance),
                        const_cast<Launch**>(&launches), &copy};
        return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&program_kernel), dim3(blocks), dim3(kTile),
                                          args, static_cast<unsigned>(dynamic_lds), stream);
    }
    hipLaunchKernelGGL(program_kernel, dim3(blocks), dim3(kTile), dynamic_lds, stream, instances, block_instance,
                       launches, single);
    return hipGetLastError();
}

}  // namespace psim::dev

namespace psim::dev {
int program_blocks_per_cu(size_t dynamic_lds) {
    int blocks = 0;
    if (dynamic_lds > 0 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&program_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(dynamic_lds)) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, program_kernel, kTile, dynamic_lds) != hipSuccess)
        return 0;
    return blocks;
}
}  // namespace psim::dev
