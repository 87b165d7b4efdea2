// The persistent solver program for gfx950.
//
// One launch runs whole operations (set_matrix, a V-cycle, a complete
// BiCGSTAB solve with its true-residual check, or a chain of heat steps)
// for a batch of instances. Each instance owns a group of workgroups; the
// group walks the instance's phases tile by tile (256 rows per tile, one
// row per thread) and meets at a group barrier between phases that read
// what the previous one wrote. With one workgroup per instance the barrier
// is a plain __syncthreads.
//
// The algorithm is the reference's (heat_transport.cpp:820-897 BiCGSTAB,
// right-preconditioned, from x = 0; V(1,1) aggregation multigrid of
// :288-326 with Galerkin values of :157-171), in these device forms:
//   * smoothing is Chebyshev-weighted Jacobi (weight 2/(a+b) on the
//     Gershgorin interval [a, b] of D^-1 A, a clipped at b/4) instead of
//     block Gauss-Seidel; pre-smoothing from zero, the residual and the
//     restriction of a level are one sweep: residuals are staged in LDS in
//     aggregate-major order and each coarse row sums its fine rows in
//     ascending order (the reference's restriction order);
//   * prolongation is fused with post-smoothing, and the coarsest level's
//     dense solve (a GEMV with its inverse, built at set_matrix time by the
//     symmetric sweep operator) is recomputed into LDS by each workgroup of
//     the last prolongation sweep instead of costing a phase of its own;
//   * the vector updates feeding a V-cycle (p = r + beta (p - omega v),
//     s = r - alpha v) are evaluated inside its first sweep (neighbours
//     recompute them; each row stores its own), and the x update after a
//     half-step exit is fused with the FP64 true-residual SpMV;
//   * inner products reduce each wave with DPP and swizzle moves in a fixed
//     order (wave.hpp), then the waves and the tile partials in a fixed order, so every workgroup computes the same
//     scalars and results do not depend on the group size or the batch;
//   * BiCGSTAB stops on the recursive residual as the reference does, then
//     the FP64 residual b - A x is recomputed; above tolerance the solve
//     restarts from x (the mixed-precision form relies on this refinement).
#include "device.hpp"
#include "wave.hpp"

#include <type_traits>

namespace psim::dev {
namespace {

// A pointer every lane holds, rebuilt in scalar registers: descriptor reads
// through it are scalar loads (a function argument arrives in vector
// registers otherwise).
template <class T> __device__ inline const T* uniform_ptr(const T* p) {
    const uint64_t bits = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(bits));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(bits >> 32));
    return reinterpret_cast<const T*>((static_cast<uint64_t>(hi) << 32) | lo);
}
template <class T> __device__ inline T* uniform_ptr(T* p) {
    return const_cast<T*>(uniform_ptr(static_cast<const T*>(p)));
}

// Every array the program touches lives in device global memory. Pointers
// that come out of a descriptor are generic (flat) pointers to the
// compiler, and flat accesses cost twice: they count against the LDS and
// scalar-load counter too, so each wait for an LDS or descriptor load also
// waits for every flat load in flight, and atomic flat loads are issued one
// wait at a time. The program therefore reads and writes its arrays through
// global-address-space pointers (global_load / global_store): independent
// loads stay in flight together.
template <class T> using gptr_t = __attribute__((address_space(1))) T*;
template <class T> __device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}
template <class T> __device__ __forceinline__ __attribute__((address_space(1))) T* gptr(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// ------------------------------------------------------------ reductions

// Values that a later phase reads from other workgroups are written with
// agent-scope (coherent) stores: they go through to memory instead of
// staying dirty in this XCD's L2. A fenced group barrier then has (almost)
// nothing left to write back; a lean one (group_sync<true>) only waits for
// each thread's stores, and the readers load them with fetch<true>.
template <class T> __device__ __forceinline__ void publish(T* p, T v) {
    __hip_atomic_store(gptr(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A value another workgroup wrote before the last group barrier. kFresh:
// an agent-scope (coherent) load, which reads past this XCD's L2, for the
// phases after a lean barrier (group_sync<true>), which does not invalidate
// the L2; otherwise a plain load (the barrier's acquire fence invalidated
// the caches).
template <bool kFresh, class T> __device__ __forceinline__ T fetch(const T* p) {
    if constexpr (kFresh)
        return __hip_atomic_load(gptr(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        return *gptr(p);
}

// The value type behind a pointer of any address space (int32, float and
// double arrays only: the unary plus would promote narrower integers).
template <class P> using elem_t = decltype(+(*static_cast<P>(nullptr)));

// A plain load of element i (read-only operator and graph data).
template <class P> __device__ __forceinline__ elem_t<P> ld(P p, int i) {
    using T = elem_t<P>;
    static_assert(sizeof(T) == sizeof(*p), "array element type");
    return ((const __attribute__((address_space(1))) T*)p)[i];
}

// The coarsest inverse's MFMA blocks: 16 x 16, kMaxBlocks of them along a
// side for up to kBlockedCoarse rows.
constexpr int kMB = kSweepBlock;
constexpr int kMaxBlocks = (kBlockedCoarse + kMB - 1) / kMB;
typedef double v4d __attribute__((ext_vector_type(4)));

struct Lds {
    double red[kPartials][4];        // per-wave partials
    double bcast[kPartials];
    union {
        struct {
            double tile[2][kTile];       // restriction staging (two tiles in flight)
            double cvec[kMaxCoarse];     // coarse right-hand side, then solution
            double crhs[kMaxCoarse];     // coarse right-hand side (FP64 staging)
            double colk[kMaxCoarse];     // sweep pivot column (set_matrix, large coarsest)
        };
        // set_matrix, coarsest of at most kBlockedCoarse rows (block sweep
        // with MFMA): the block column K before the step and its product with
        // the swept pivot block, both as column-major 16 x 16 blocks (entry
        // (r, c) of block i at [i][c * 16 + r], so MFMA operand reads are
        // consecutive across lanes), and the pivot block under its scalar
        // sweeps, double-buffered, row-major
        struct {
            double colT[kMaxBlocks][kMB * kMB];
            double dT[kMaxBlocks][kMB * kMB];
            double piv[2][kMB * kMB];
        };
    };
};

// Dynamic LDS: the packed coarsest matrix during set_matrix launches.
__device__ inline double* dynamic_lds() {
    extern __shared__ double dyn[];
    return dyn;
}

// Sum of K per-thread values over the workgroup, fixed order: wave_sum in
// each wave, then waves 0..3 left to right. Thread 0 gets the totals.
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

template <int K> __device__ inline void store_partials(double* part, int tile, const double (&v)[K], Lds& lds) {
    double tot[K];
    tile_sums<K>(v, tot, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < K; ++q) publish(&part[tile * kPartials + q], tot[q]);
    }
}

// Partials of R tiles at once (one reduction pass, one pair of barriers):
// tile tl[r] gets exactly what store_partials<K> would give it.
template <int K, int R>
__device__ inline void store_partials_n(double* part, const int (&tl)[R], const double (&v)[R][K], Lds& lds) {
    static_assert(K * R <= kPartials, "reduction slots");
    if constexpr (R == 1) {
        store_partials<K>(part, tl[0], v[0], lds);
    } else {
        double flat[K * R], tot[K * R];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int q = 0; q < K; ++q) flat[r * K + q] = v[r][q];
        tile_sums<K * R>(flat, tot, lds);
        if (threadIdx.x == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int q = 0; q < K; ++q) publish(&part[tl[r] * kPartials + q], tot[r * K + q]);
        }
    }
}

// This workgroup's tiles of a sweep (g, g + G, g + 2G, ...), two at a time
// while two remain, so that each thread keeps two rows' loads in flight;
// f is called with the tile list (one or two tiles). Pairs hold two rows'
// slots in registers, which FP32 sweeps afford and FP64 sweeps afford only
// in the program instance with a 256-VGPR budget (two workgroups per CU:
// kW == 2); elsewhere FP64 sweeps go one tile at a time.
template <bool Pairs = true, class F> __device__ inline void tile_loop(int tiles, int g, int G, const F& f) {
    if constexpr (!Pairs) {
        for (int t = g; t < tiles; t += G) {
            const int tl[1] = {t};
            f(tl);
        }
        return;
    }
    for (int t = g; t < tiles; t += 2 * G) {
        if (t + G < tiles) {
            const int tl[2] = {t, t + G};
            f(tl);
        } else {
            const int tl[1] = {t};
            f(tl);
        }
    }
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

// Totals of slots 0..K-1 over `tiles` tiles (fixed order), to every thread.
// Lane l adds tiles l, l + 64, l + 128, ... in ascending order; the loads of
// kU such tiles and all K slots go out together before any is added (a lane
// past the last tile loads the last tile again and discards it, so no load
// sits behind a branch).
template <int K, bool kF = false>
__device__ inline void group_totals(const double* partial, int tiles, double (&out)[K], Lds& lds) {
    if (threadIdx.x < 64) {
        constexpr int kU = 4;
        double s[K];
#pragma unroll
        for (int q = 0; q < K; ++q) s[q] = 0.0;
        for (int t0 = threadIdx.x; t0 < tiles; t0 += 64 * kU) {
            double v[kU][K];
#pragma unroll
            for (int u = 0; u < kU; ++u)
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const int t = min(t0 + 64 * u, tiles - 1);
                    v[u][q] = fetch<kF>(&partial[t * kPartials + q]);
                }
#pragma unroll
            for (int u = 0; u < kU; ++u)
#pragma unroll
                for (int q = 0; q < K; ++q)
                    if (t0 + 64 * u < tiles) s[q] += v[u][q];
        }
#pragma unroll
        for (int q = 0; q < K; ++q) {
            const double w = wave_sum(s[q]);
            if (threadIdx.x == 0) lds.bcast[q] = w;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < K; ++q) out[q] = lds.bcast[q];
    __syncthreads();
}

// Like group_totals, the partials are read with fetch<kF>: after a lean
// barrier (kF) they must be coherent loads; set_matrix, the only caller,
// meets at the fenced barrier and reads them with plain loads.
template <bool kF>
__device__ inline double group_max(const double* partial, int tiles, Lds& lds) {
    if (threadIdx.x < 64) {
        double m = 0.0;
        for (int t = threadIdx.x; t < tiles; t += 64) m = fmax(m, fetch<kF>(&partial[t * kPartials]));
        m = wave_max(m);
        if (threadIdx.x == 0) lds.bcast[0] = m;
    }
    __syncthreads();
    const double out = lds.bcast[0];
    __syncthreads();
    return out;
}

// ------------------------------------------------------------ group barrier

// The group this workgroup belongs to while it serves an instance: its
// workgroups, barrier words, result record and phase trace. Set by the
// kernel entry (from the instance descriptor, or from a server slot) and
// read from LDS, so every use is a uniform LDS read.
struct Group {
    int32_t first, size;
    uint32_t* barrier;
    SolveResult* result;
    uint64_t* trace;
    uint32_t* trace_count;
    int32_t trace_capacity;
    uint32_t episode;   // group barriers this workgroup has entered in this launch
};
__shared__ Group grp;

__device__ inline void set_group(int32_t first, int32_t size, uint32_t* barrier, SolveResult* result, uint64_t* trace,
                                 uint32_t* trace_count, int32_t trace_capacity) {
    if (threadIdx.x == 0) grp = Group{first, size, barrier, result, trace, trace_count, trace_capacity, 0u};
    __syncthreads();
}

// Phase trace: workgroup 0 of a traced group stamps the wall clock at every
// group barrier (after it completes).
__device__ inline void trace_stamp() {
    if (grp.trace != nullptr && threadIdx.x == 0 && static_cast<int>(blockIdx.x) == grp.first) {
        const uint32_t slot = grp.trace_count[0];
        if (slot < static_cast<uint32_t>(grp.trace_capacity)) grp.trace[slot] = static_cast<uint64_t>(wall_clock64());
        grp.trace_count[0] = slot + 1;
    }
}

// Counting barrier over the instance's workgroups. Every barrier of a
// launch is an episode k = 1, 2, ... that all workgroups of the group enter
// in the same order. Workgroup g adds 1 to leaf counter g % kLeaves (eight
// counters on separate 256-byte lines, so arrivals do not serialise on one
// address) with an atomic whose result it does not wait for; lanes 0..7 of
// its first wave then poll the eight leaves until each holds at least
// k x (the leaf's share of the group). The counters only grow within a
// launch, so no arrival is waited on for more than its own trip to memory
// and back: one round trip from the last arrival to every waiter's release,
// where a tree with a generation word takes four (leaf, root, clear,
// generation). A counter is at most one episode ahead of a waiter (a
// workgroup enters episode k + 1 only after k completes, which needs this
// waiter's arrival), so the comparison is wrap-safe. The last workgroup to
// leave the launch (group_exit) clears the counters, so every launch starts
// from zero whatever the previous group size was. The wait is bounded by
// the 100 MHz wall clock (about one second); a timeout is recorded and is
// sticky, so a broken group drains within seconds and the host reports it.
//
// Ordering (fenced form): thread 0 issues one agent-scope release fence
// before its relaxed arrival and the polling lanes one acquire fence after
// the wait. Every arrival is an RMW, so the value a waiter reads continues
// the release sequence of every earlier arrival on that leaf, and reading
// all eight leaves at their targets synchronises with every workgroup's
// release: fence-fence synchronisation orders all writes before the
// barrier before all reads after it.
//
// Lean form (kLean, the solve and PCG phases, whose values for other
// workgroups are all written with coherent stores and read with coherent
// loads): each thread waits for its own stores (s_waitcnt vmcnt(0)) instead
// of the release fence, and no acquire fence follows the wait, so this
// XCD's L2 keeps the read-only operator data across phases.
//
// Barrier words (uint32, zeroed at creation and at every launch's end):
//   [32]         workgroups that have left the launch (group_exit)
//   [96 + 64 l]  arrivals on leaf l in this launch
constexpr long long kBarrierTicks = 100000000;
constexpr int kLeaves = 8;
static_assert(96 + 64 * kLeaves == kBarrierWordCount, "barrier layout");

template <bool kLean>
__device__ __noinline__ int group_sync_call(int timeouts);

// The barrier's code is called, not inlined: one copy of it stays in the
// instruction cache for every phase of a step.
template <bool kLean = false>
__device__ __forceinline__ void group_sync(int& timeouts) {
    timeouts = group_sync_call<kLean>(timeouts);
}

template <bool kLean>
__device__ __noinline__ int group_sync_call(int timeouts) {
    if constexpr (kLean) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const int G = grp.size;
    if (G == 1) {
        trace_stamp();
        return timeouts;
    }
    if (threadIdx.x < 64) {
        const uint32_t k = grp.episode + 1;
        const int g = static_cast<int>(blockIdx.x) - grp.first;
        const int lane = static_cast<int>(threadIdx.x);
        if (lane == 0) {
            if constexpr (!kLean) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            (void)__hip_atomic_fetch_add(grp.barrier + 96 + 64 * (g % kLeaves), 1u, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
            grp.episode = k;
        }
        bool late = false;
        if (lane < kLeaves && lane < G && timeouts == 0) {
            const uint32_t target = k * static_cast<uint32_t>((G - lane + kLeaves - 1) / kLeaves);
            const uint32_t* leaf = grp.barrier + 96 + 64 * lane;
            // the clock is read every kClockPolls polls only (its read waits
            // on the scalar memory path, which the next LDS or scalar access
            // would wait for too); a wait that ends sooner never reads it
            constexpr uint32_t kClockPolls = 64;
            uint32_t polls = 0;
            long long start = 0;
            while (static_cast<int32_t>(__hip_atomic_load(const_cast<uint32_t*>(leaf), __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT) -
                                        target) < 0) {
                __builtin_amdgcn_s_sleep(1);
                if (++polls % kClockPolls == 0) {
                    const long long now = wall_clock64();
                    if (polls == kClockPolls) {
                        start = now;
                    } else if (now - start > kBarrierTicks) {
                        late = true;
                        break;
                    }
                }
            }
        }
        if (__any(late)) timeouts = 1;
        if constexpr (!kLean) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    trace_stamp();
    __syncthreads();
    return timeouts;
}

// A workgroup leaves the launch: the last of its group to leave clears the
// barrier counters for the next launch (every other workgroup of the group
// has passed its last barrier by then, so no wait can meet the clear).
__device__ inline void group_exit() {
    __syncthreads();
    const int G = grp.size;
    if (G > 1 && threadIdx.x == 0) {
        uint32_t* left = grp.barrier + 32;
        if (__hip_atomic_fetch_add(left, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 ==
            static_cast<uint32_t>(G)) {
            for (int l = 0; l < kLeaves; ++l)
                __hip_atomic_store(grp.barrier + 96 + 64 * l, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(left, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ------------------------------------------------------------ typed views

template <class VT> struct Values;
template <> struct Values<double> {
    const Instance& I;
    __device__ auto diag(int j) const { return gptr(static_cast<const double*>(I.diag64[j])); }
    __device__ auto dinv(int j) const { return gptr(static_cast<const double*>(I.dinv64[j])); }
    __device__ auto off(int j) const { return gptr(static_cast<const double*>(I.off64[j])); }
};
template <> struct Values<float> {
    const Instance& I;
    __device__ auto diag(int j) const { return gptr(static_cast<const float*>(I.diag32[j])); }
    __device__ auto dinv(int j) const { return gptr(static_cast<const float*>(I.dinv32[j])); }
    __device__ auto off(int j) const { return gptr(static_cast<const float*>(I.off32[j])); }
};

// The level-0 input of a V-cycle, evaluated where it is read:
//   kind 0: a;  kind 1: a + c1 (b - c2 c)  (p update);  kind 2: a - c1 b  (s)
// in FP64 and rounded to VT, exactly as the stand-alone update would.
template <class VT> struct Source {
    const VT* a;
    const VT* b;
    const VT* c;
    double c1, c2;
    int kind;
    template <bool kF = false> __device__ VT at(int k) const {
        if (kind == 0) return fetch<kF>(a + k);
        if (kind == 1)
            return static_cast<VT>(static_cast<double>(fetch<kF>(a + k)) +
                                   c1 * (static_cast<double>(fetch<kF>(b + k)) - c2 * static_cast<double>(fetch<kF>(c + k))));
        return static_cast<VT>(static_cast<double>(fetch<kF>(a + k)) - c1 * static_cast<double>(fetch<kF>(b + k)));
    }
};

// Two values one gather returns.
template <class VT> struct Pair {
    VT first, second;
};

// Row i of an ELL level: fold(off_s, gather(s, c_s)) over the row's slots
// in order, skipping padding (c_s < 0). All column and value loads are
// issued first, then every slot's gather (a padding slot gathers at row i,
// a valid address, and is discarded), then the folds, so a row costs two
// dependent memory round trips instead of two per slot. Fold order and
// every rounding are those of the plain slot loop.
constexpr int kMaxWidth = 7;   // widest icosphere level (hexagons + aggregation: 6, 7)

template <class CP, class OP, class Gather, class Fold>
__device__ inline void ell_row(CP col_in, OP off_in, int n, int W, int i, const Gather& gather, const Fold& fold) {
    using T = elem_t<OP>;
    const auto col = (const __attribute__((address_space(1))) int32_t*)col_in;
    const auto off = (const __attribute__((address_space(1))) T*)off_in;
    if (W <= kMaxWidth) {
        int c[kMaxWidth];
        T o[kMaxWidth];
#pragma unroll
        for (int s = 0; s < kMaxWidth; ++s)
            if (s < W) {
                c[s] = col[s * n + i];
                o[s] = off[s * n + i];
            }
        decltype(gather(0, 0)) v[kMaxWidth];
#pragma unroll
        for (int s = 0; s < kMaxWidth; ++s)
            if (s < W) v[s] = gather(s, c[s] >= 0 ? c[s] : i);
#pragma unroll
        for (int s = 0; s < kMaxWidth; ++s)
            if (s < W && c[s] >= 0) fold(o[s], v[s]);
    } else {
        for (int s = 0; s < W; ++s) {
            const int cs = col[s * n + i];
            if (cs >= 0) fold(off[s * n + i], gather(s, cs));
        }
    }
}

// The same for R rows of a thread at once (rows of different tiles): all
// rows' column and value loads, then all their gathers, then each row's
// folds in slot order. fold(r, off, value) and gather(r, slot, column).
template <int R, class CP, class OP, class Gather, class Fold>
__device__ inline void ell_rows(CP col_in, OP off_in, int n, int W, const int (&i)[R], const Gather& gather,
                                const Fold& fold) {
    using T = elem_t<OP>;
    const auto col = (const __attribute__((address_space(1))) int32_t*)col_in;
    const auto off = (const __attribute__((address_space(1))) T*)off_in;
    if (W <= kMaxWidth) {
        int c[R][kMaxWidth];
        T o[R][kMaxWidth];
#pragma unroll
        for (int s = 0; s < kMaxWidth; ++s)
            if (s < W) {
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    c[r][s] = col[s * n + i[r]];
                    o[r][s] = off[s * n + i[r]];
                }
            }
        decltype(gather(0, 0, 0)) v[R][kMaxWidth];
#pragma unroll
        for (int s = 0; s < kMaxWidth; ++s)
            if (s < W) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[r][s] = gather(r, s, c[r][s] >= 0 ? c[r][s] : i[r]);
            }
#pragma unroll
        for (int s = 0; s < kMaxWidth; ++s)
            if (s < W) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (c[r][s] >= 0) fold(r, o[r][s], v[r][s]);
            }
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
            for (int s = 0; s < W; ++s) {
                const int cs = col[s * n + i[r]];
                if (cs >= 0) fold(r, off[s * n + i[r]], gather(r, s, cs));
            }
    }
}

// Row of tile tl and this thread; rows past n read row 0 (a valid address)
// and are discarded.
__device__ inline int tile_row(int tl, int n, bool& ok) {
    const int i = tl * kTile + static_cast<int>(threadIdx.x);
    ok = i < n;
    return ok ? i : 0;
}

// ------------------------------------------------------------ V-cycle phases

// One restriction tile of down_phase (below) for a plain input w (no
// Source update, no A w), with every global load issued first and without
// a branch: the tile's row id and coarse-row range, then the row's
// columns, values and own inputs beside the aggregate bounds, then the
// neighbour gathers (slots past the width reload the last slot, padding
// rows reload row 0, and the results are discarded by selects), so the
// compiler keeps them all in flight instead of waiting at every join. The
// operations, their order and every rounding are down_phase's. kPre: the
// level has a two-sweep iterate.
template <class VT, bool kF, bool kPre>
__device__ __forceinline__ void down_tile_early(const Instance& I, const Values<VT>& V, int j, int tl, const VT* w,
                                                VT* keep, VT* keep2, VT* coarse_rhs, double* part, Lds& lds,
                                                const VT* pre) {
    const LevelView& L = I.lv[j];
    const int n = L.n, W = L.width;
    const VT om = static_cast<VT>(I.omega[j]);
    const auto diag = V.diag(j);
    const auto dinv = V.dinv(j);
    const auto off = V.off(j);
    const auto col = gptr(L.col);
    const int row_id = ld(L.down_rows, tl * kTile + static_cast<int>(threadIdx.x));
    const int c0 = ld(L.tile_start, tl), c1 = ld(L.tile_start, tl + 1);
    const bool ok = row_id >= 0;
    const int i = ok ? row_id : 0;
    const int base = ld(L.agg_start, c0);
    const int c = c0 + static_cast<int>(threadIdx.x);
    const int cq = min(c, L.n_next - 1);
    const int qa = ld(L.agg_start, cq), qb = ld(L.agg_start, cq + 1);
    const int q0 = c < L.n_next ? qa : 0, q1 = c < L.n_next ? qb : 0;
    const VT wi = fetch<kF>(w + i);
    const VT pre_i = kPre ? fetch<kF>(pre + i) : VT(0);
    const VT dinv_i = ld(dinv, i), diag_i = ld(diag, i);
    int cs[kMaxWidth];
    VT o[kMaxWidth];
#pragma unroll
    for (int s = 0; s < kMaxWidth; ++s) {
        const int e = min(s, W - 1) * n + i;
        const int cv = col[e];
        cs[s] = s < W ? cv : -1;
        o[s] = off[e];
    }
    VT gv[kMaxWidth];
#pragma unroll
    for (int s = 0; s < kMaxWidth; ++s) {
        const int cc = cs[s] >= 0 ? cs[s] : i;
        if constexpr (kPre)
            gv[s] = fetch<kF>(pre + cc);
        else
            gv[s] = om * ld(dinv, cc) * fetch<kF>(w + cc);
    }
    double sq[1][1];
    sq[0][0] = ok ? static_cast<double>(wi) * static_cast<double>(wi) : 0.0;
    const VT xi = kPre ? pre_i : om * dinv_i * wi;
    VT acc = wi - diag_i * xi;
#pragma unroll
    for (int s = 0; s < kMaxWidth; ++s)
        if (cs[s] >= 0) acc -= o[s] * gv[s];
    if (ok) {
        if (keep != nullptr) publish(&keep[i], wi);
        if (keep2 != nullptr) publish(&keep2[i], wi);
        lds.tile[0][threadIdx.x] = static_cast<double>(acc);
    }
    if (part != nullptr) {
        const int tls[1] = {tl};
        store_partials_n<1, 1>(part, tls, sq, lds);
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < c1 - c0) {
        double sum = 0.0;
        for (int q = q0 - base; q < q1 - base; ++q) sum += lds.tile[0][q];
        publish(&coarse_rhs[c0 + threadIdx.x], static_cast<VT>(sum));
    }
    __syncthreads();
}

// Pre-smoothing, residual and restriction of level j in one sweep over the
// level's restriction tiles. The pre-smoothed iterate is one Jacobi sweep
// from zero (x = om D^-1 w, evaluated where it is read), or, with `pre`
// set, the two-sweep iterate jacobi_phase stored there. On level 0 the input
// comes from a Source; each row stores its value in `keep` and `keep2` (if
// set) and adds its square to the tile partial (if `part` is set).
template <class VT, int kW = 0>
__device__ __forceinline__ void down_phase(const Instance& I, const Values<VT>& V, int j, const Source<VT>& src, VT* keep,
                           VT* coarse_rhs, double* part, int g, int G, Lds& lds, const VT* pre = nullptr,
                           VT* keep2 = nullptr, VT* au = nullptr) {
    constexpr bool kF = kW >= 1;   // lean barriers, coherent loads
    const LevelView& L = I.lv[j];
    const int n = L.n, W = L.width;
    const VT om = static_cast<VT>(I.omega[j]);
    const auto diag = V.diag(j);
    const auto dinv = V.dinv(j);
    const auto off = V.off(j);
    if (kW == 1 && src.kind == 0 && au == nullptr && W >= 1 && W <= kMaxWidth && L.n_next >= 1) {
        // the full register file: each tile's loads all issued up front
        for (int tl = g; tl < L.down_tiles; tl += G) {
            if (pre != nullptr)
                down_tile_early<VT, kF, true>(I, V, j, tl, src.a, keep, keep2, coarse_rhs, part, lds, pre);
            else
                down_tile_early<VT, kF, false>(I, V, j, tl, src.a, keep, keep2, coarse_rhs, part, lds, pre);
        }
        return;
    }
    tile_loop<sizeof(VT) == 4 || kW == 2>(L.down_tiles, g, G, [&]<int R>(const int (&tl)[R]) {
        // the tile's rows come from its padded row list (one load); the
        // aggregate bounds the restriction needs load alongside
        int c0[R], c1[R], base[R], i[R];
        bool ok[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            i[r] = ld(L.down_rows, tl[r] * kTile + threadIdx.x);
            c0[r] = ld(L.tile_start, tl[r]);
            c1[r] = ld(L.tile_start, tl[r] + 1);
        }
        int q0[R], q1[R];   // this thread's coarse row's fine-row range, for the restriction
#pragma unroll
        for (int r = 0; r < R; ++r) {
            ok[r] = i[r] >= 0;
            i[r] = ok[r] ? i[r] : 0;
            base[r] = ld(L.agg_start, c0[r]);
            const int c = c0[r] + static_cast<int>(threadIdx.x);
            q0[r] = c < L.n_next ? ld(L.agg_start, c) : 0;
            q1[r] = c < L.n_next ? ld(L.agg_start, c + 1) : 0;
        }
        VT wi[R], acc[R], aw[R];
        double sq[R][1];
#pragma unroll
        for (int r = 0; r < R; ++r) wi[r] = src.template at<kF>(i[r]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sq[r][0] = ok[r] ? static_cast<double>(wi[r]) * static_cast<double>(wi[r]) : 0.0;
            const VT xi = pre != nullptr ? fetch<kF>(pre + i[r]) : om * ld(dinv, i[r]) * wi[r];
            acc[r] = wi[r] - ld(diag, i[r]) * xi;
            aw[r] = ld(diag, i[r]) * wi[r];
        }
        if (au != nullptr) {
            // (A w)_i beside the residual: the neighbours' inputs and their
            // Jacobi iterates from the same gathers
            ell_rows<R>(
                L.col, off, n, W, i,
                [&](int, int, int c) {
                    const VT wc = src.template at<kF>(c);
                    return Pair<VT>{wc, om * ld(dinv, c) * wc};
                },
                [&](int r, VT o, Pair<VT> v) {
                    acc[r] -= o * v.second;
                    aw[r] += o * v.first;
                });
        } else {
            ell_rows<R>(
                L.col, off, n, W, i,
                [&](int, int, int c) { return pre != nullptr ? fetch<kF>(pre + c) : om * ld(dinv, c) * src.template at<kF>(c); },
                [&](int r, VT o, VT v) { acc[r] -= o * v; });
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (ok[r]) {
                if (keep != nullptr) publish(&keep[i[r]], wi[r]);
                if (keep2 != nullptr) publish(&keep2[i[r]], wi[r]);
                if (au != nullptr) publish(&au[i[r]], aw[r]);
                lds.tile[r][threadIdx.x] = static_cast<double>(acc[r]);
            }
        if (part != nullptr) store_partials_n<1, R>(part, tl, sq, lds);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (static_cast<int>(threadIdx.x) < c1[r] - c0[r]) {
                double sum = 0.0;
                for (int q = q0[r] - base[r]; q < q1[r] - base[r]; ++q) sum += lds.tile[r][q];
                publish(&coarse_rhs[c0[r] + threadIdx.x], static_cast<VT>(sum));
            }
        __syncthreads();
    });
}

// First of two pre-smoothing sweeps on level j, over natural 256-row tiles:
// with w the level's input, x1 = om D^-1 w (evaluated where it is read),
// r1 = w - A x1 and x2 = x1 + om D^-1 r1 into `pre`, which down_phase and
// the prolongation then read as the pre-smoothed iterate. Each row stores
// w in `keep` and `keep2` (if set) and adds w^2 to its tile partial (if
// `part` is set). With `r1_out` set (the split V-cycle of the mixed form),
// r1, computed in VT, is stored there rounded to FP32 instead of x2.
template <class VT, int kW = 0>
__device__ __forceinline__ void jacobi_phase(const Instance& I, const Values<VT>& V, int j, const Source<VT>& src,
                                             VT* keep, VT* keep2, VT* pre, double* part, int g, int G, Lds& lds,
                                             VT* au = nullptr, float* r1_out = nullptr) {
    constexpr bool kF = kW >= 1;   // lean barriers, coherent loads
    const LevelView& L = I.lv[j];
    const int n = L.n, W = L.width;
    const VT om = static_cast<VT>(I.omega[j]);
    const auto diag = V.diag(j);
    const auto dinv = V.dinv(j);
    const auto off = V.off(j);
    const int tiles = (n + kTile - 1) / kTile;
    // LDS-staged neighbour gathers (FP64 sweeps one tile at a time, plain
    // inputs): each row's input and D^-1 entry go to LDS once per tile, a
    // neighbour inside the tile (about 85 % of them on icosphere levels,
    // whose cell order keeps neighbours close) is read from LDS and only the
    // others from memory, so the sweep's coherent gathers drop to the tile
    // boundary's. The values, and so every operation, are the ones the
    // gathers from memory return.
    if constexpr (sizeof(VT) == 8 && kW != 2) {
        if (src.kind == 0 && W >= 1 && W <= kMaxWidth) {
            double* sw = lds.tile[0];
            double* sd = lds.tile[1];
            const auto col = gptr(L.col);
            for (int tile = g; tile < tiles; tile += G) {
                bool okr;
                const int ir = tile_row(tile, n, okr);
                const int base = tile * kTile;
                // the row's columns and values go out with its own inputs,
                // before the staging barrier (no branch: slots past the
                // width reload the last slot, marked padding)
                const int Wc = W <= kMaxWidth ? W : kMaxWidth;
                int cs[kMaxWidth];
                VT o[kMaxWidth];
#pragma unroll
                for (int q = 0; q < kMaxWidth; ++q) {
                    const int e = min(q, Wc - 1) * n + ir;
                    const int cv = col[e];
                    cs[q] = q < Wc ? cv : -1;
                    o[q] = off[e];
                }
                const VT wr = fetch<kF>(src.a + ir);
                const VT dr = ld(dinv, ir), gr = ld(diag, ir);
                sw[threadIdx.x] = static_cast<double>(wr);
                sd[threadIdx.x] = static_cast<double>(dr);
                __syncthreads();
                // the neighbour's input and D^-1 entry: LDS inside the tile
                Pair<VT> v[kMaxWidth];
#pragma unroll
                for (int q = 0; q < kMaxWidth; ++q) {
                    const int c = cs[q] >= 0 ? cs[q] : ir;
                    const unsigned k = static_cast<unsigned>(c - base);
                    if (k < static_cast<unsigned>(kTile))
                        v[q] = Pair<VT>{static_cast<VT>(sw[k]), static_cast<VT>(sd[k])};
                    else
                        v[q] = Pair<VT>{fetch<kF>(src.a + c), ld(dinv, c)};
                }
                double sq[1] = {okr ? static_cast<double>(wr) * static_cast<double>(wr) : 0.0};
                VT acc = wr - gr * (om * dr * wr);
                VT aw = gr * wr;
#pragma unroll
                for (int q = 0; q < kMaxWidth; ++q)
                    if (cs[q] >= 0) {
                        acc -= o[q] * (om * v[q].second * v[q].first);
                        if (au != nullptr) aw += o[q] * v[q].first;
                    }
                if (okr) {
                    if (r1_out != nullptr) {
                        publish(&r1_out[ir], static_cast<float>(acc));
                    } else {
                        const VT x1 = om * dr * wr;
                        publish(&pre[ir], static_cast<VT>(x1 + om * dr * acc));
                    }
                    if (keep != nullptr) publish(&keep[ir], wr);
                    if (keep2 != nullptr) publish(&keep2[ir], wr);
                    if (au != nullptr) publish(&au[ir], aw);
                }
                if (part != nullptr) store_partials<1>(part, tile, sq, lds);
                __syncthreads();
            }
            return;
        }
    }
    tile_loop<sizeof(VT) == 4 || kW == 2>(tiles, g, G, [&]<int R>(const int (&tl)[R]) {
        int i[R];
        bool ok[R];
        VT wi[R], acc[R], aw[R];
        double sq[R][1];
#pragma unroll
        for (int r = 0; r < R; ++r) i[r] = tile_row(tl[r], n, ok[r]);
#pragma unroll
        for (int r = 0; r < R; ++r) wi[r] = src.template at<kF>(i[r]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            sq[r][0] = ok[r] ? static_cast<double>(wi[r]) * static_cast<double>(wi[r]) : 0.0;
            acc[r] = wi[r] - ld(diag, i[r]) * (om * ld(dinv, i[r]) * wi[r]);
            aw[r] = ld(diag, i[r]) * wi[r];
        }
        if (au != nullptr) {
            // (A w)_i beside the residual, from the same gathers
            ell_rows<R>(
                L.col, off, n, W, i,
                [&](int, int, int c) {
                    const VT wc = src.template at<kF>(c);
                    return Pair<VT>{wc, om * ld(dinv, c) * wc};
                },
                [&](int r, VT o, Pair<VT> v) {
                    acc[r] -= o * v.second;
                    aw[r] += o * v.first;
                });
        } else {
            ell_rows<R>(
                L.col, off, n, W, i, [&](int, int, int c) { return om * ld(dinv, c) * src.template at<kF>(c); },
                [&](int r, VT o, VT v) { acc[r] -= o * v; });
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (ok[r]) {
                if (r1_out != nullptr) {
                    publish(&r1_out[i[r]], static_cast<float>(acc[r]));
                } else {
                    const VT x1 = om * ld(dinv, i[r]) * wi[r];
                    publish(&pre[i[r]], static_cast<VT>(x1 + om * ld(dinv, i[r]) * acc[r]));
                }
                if (keep != nullptr) publish(&keep[i[r]], wi[r]);
                if (keep2 != nullptr) publish(&keep2[i[r]], wi[r]);
                if (au != nullptr) publish(&au[i[r]], aw[r]);
            }
        if (part != nullptr) store_partials_n<1, R>(part, tl, sq, lds);
    });
}

// Rows of the coarsest level's solution A_c^-1 rhs into lds.cvec (indexed
// by coarse row), in FP64: the `count` rows list[0..count) (entries < 0 are
// skipped), or rows first .. first + count - 1 when list is null. One wave
// per row: lanes split the columns (row-major reads of the symmetric
// inverse, coalesced), each lane adding its columns in ascending order from
// 0, and wave_sum adds the lanes in its fixed order, so every workgroup computing a row
// gets the same bits. A wave takes kCoarseRows rows at a time.
//
// Two steps: coarse_rows_issue loads the first batch's row ids and, with
// kEarly, the first kCoarseChunks columns of each of those rows for this
// lane (independent of the right-hand side, so they are in flight while the
// caller issues its own loads and the right-hand side is staged);
// coarse_rows_finish stages the right-hand side in LDS and computes the
// rows. coarse_rows_lds is the two in a row.
constexpr int kCoarseRows = 4;     // rows in flight per wave
constexpr int kCoarseChunks = 4;   // columns per lane and row loaded at a time
struct CoarseFirst {
    int a[kCoarseRows];                        // the wave's first-round rows (-1: none)
    double m[kCoarseRows][kCoarseChunks];      // with kEarly: their first columns for this lane
};

// kEarly: list is set and count >= 1; the ids load without a branch (a
// position past the list reloads its last entry, discarded)
template <bool kEarly>
__device__ __forceinline__ void coarse_row_ids(const int32_t* list, int first, int count, int q0, int (&a)[kCoarseRows]) {
#pragma unroll
    for (int u = 0; u < kCoarseRows; ++u) {
        const int q = q0 + u;
        if constexpr (kEarly) {
            const int id = ld(list, min(q, count - 1));
            a[u] = q < count ? id : -1;
        } else {
            a[u] = q < count ? (list != nullptr ? ld(list, q) : first + q) : -1;
        }
    }
}

template <bool kEarly>
__device__ __forceinline__ CoarseFirst coarse_rows_issue(const Instance& I, const int32_t* list, int first, int count) {
    CoarseFirst cf;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    coarse_row_ids<kEarly>(list, first, count, wave * kCoarseRows, cf.a);
    if constexpr (kEarly) {
        const int n = I.lv[I.depth - 1].n;
        const auto inv = gptr(static_cast<const double*>(I.cinv));
#pragma unroll
        for (int u = 0; u < kCoarseRows; ++u) {
            const auto row = inv + static_cast<size_t>(cf.a[u] >= 0 ? cf.a[u] : 0) * n;
#pragma unroll
            for (int c = 0; c < kCoarseChunks; ++c) cf.m[u][c] = row[min(lane + 64 * c, n - 1)];
        }
    }
    return cf;
}

template <class VT, bool kF, bool kEarly>
__device__ __forceinline__ void coarse_rows_finish(const Instance& I, const VT* rhs, const int32_t* list, int first,
                                                   int count, Lds& lds, CoarseFirst& cf) {
    const int n = I.lv[I.depth - 1].n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < n; k += blockDim.x) lds.crhs[k] = static_cast<double>(fetch<kF>(&rhs[k]));
    __syncthreads();
    const auto inv = gptr(static_cast<const double*>(I.cinv));
    for (int q0 = wave * kCoarseRows; q0 < count; q0 += 4 * kCoarseRows) {
        const bool first_batch = q0 == wave * kCoarseRows;
        if (!first_batch) coarse_row_ids<kEarly>(list, first, count, q0, cf.a);
        double acc[kCoarseRows];
#pragma unroll
        for (int u = 0; u < kCoarseRows; ++u) acc[u] = 0.0;
        // kCoarseChunks columns per lane and row at a time, every load issued
        // before the first product (an index past the row or a skipped row
        // loads a valid element and is discarded); each lane still adds its
        // columns in ascending order
        int k0 = lane;
        if (kEarly && first_batch) {
            // the first columns, loaded by coarse_rows_issue
#pragma unroll
            for (int u = 0; u < kCoarseRows; ++u)
#pragma unroll
                for (int c = 0; c < kCoarseChunks; ++c) {
                    const int k = k0 + 64 * c;
                    if (cf.a[u] >= 0 && k < n) acc[u] += cf.m[u][c] * lds.crhs[k];
                }
            k0 += 64 * kCoarseChunks;
        }
        for (; k0 < n; k0 += 64 * kCoarseChunks) {
            double m[kCoarseRows][kCoarseChunks];
#pragma unroll
            for (int u = 0; u < kCoarseRows; ++u) {
                const auto row = inv + static_cast<size_t>(cf.a[u] >= 0 ? cf.a[u] : 0) * n;
#pragma unroll
                for (int c = 0; c < kCoarseChunks; ++c) m[u][c] = row[min(k0 + 64 * c, n - 1)];
            }
#pragma unroll
            for (int u = 0; u < kCoarseRows; ++u)
#pragma unroll
                for (int c = 0; c < kCoarseChunks; ++c) {
                    const int k = k0 + 64 * c;
                    if (cf.a[u] >= 0 && k < n) acc[u] += m[u][c] * lds.crhs[k];
                }
        }
#pragma unroll
        for (int u = 0; u < kCoarseRows; ++u) {
            const double total = wave_sum(acc[u]);
            if (lane == 0 && cf.a[u] >= 0) lds.cvec[cf.a[u]] = total;
        }
    }
    __syncthreads();
}

template <class VT, bool kF = false, bool kEarly = false>
__device__ __forceinline__ void coarse_rows_lds(const Instance& I, const VT* rhs, const int32_t* list, int first, int count,
                                                Lds& lds) {
    CoarseFirst cf = coarse_rows_issue<kEarly>(I, list, first, count);
    coarse_rows_finish<VT, kF, kEarly>(I, rhs, list, first, count, lds, cf);
}

// Level-0 output of the split V-cycle of the mixed form: its FP32 part
// works on the residual r1 = w - A x1 of the FP64 first sweep
// x1 = om D^-1 w, so the V-cycle's output is x1 + y, with x1 recomputed
// from the FP64 input w in the FP64 operations of that sweep and y the FP32
// part's level-0 result. out == nullptr: not split.
struct SplitOut {
    const double* w;     // the V-cycle's FP64 input
    double* out;         // x1 + y
    const double* dot;   // FP64 vector the output is dotted with (tile partials), or null
};

template <bool kF>
__device__ __forceinline__ double split_value(const Instance& I, const SplitOut& so, int i, float y) {
    const double om = I.omega[0];
    return om * ld(gptr(static_cast<const double*>(I.dinv64[0])), i) * fetch<kF>(&so.w[i]) + static_cast<double>(y);
}

// Prolongation of the coarse correction fused with post-smoothing on level
// j. The correction comes from `corr`, or, on the level above the coarsest,
// from the coarse solve each workgroup recomputes into LDS. The pre-smoothed
// iterate is om D^-1 rhs, or the two-sweep iterate in `pre` when set.
template <class VT, int kW = 0>
__device__ __forceinline__ void up_phase(const Instance& I, const Values<VT>& V, int j, const VT* rhs, const VT* corr, VT* out, int g,
                         int G, Lds& lds, double* dot_part = nullptr, const VT* dot_with = nullptr,
                         const VT* pre = nullptr, SplitOut so = SplitOut{}) {
    constexpr bool kF = kW >= 1;   // lean barriers, coherent loads
    const LevelView& L = I.lv[j];
    const int n = L.n, W = L.width;
    const VT om = static_cast<VT>(I.omega[j]);
    const auto diag = V.diag(j);
    const auto dinv = V.dinv(j);
    const auto off = V.off(j);
    const bool fused = j + 2 == I.depth;
    const int tiles = (n + kTile - 1) / kTile;
    if (fused) {
        // the level above the coarsest: each tile first solves the coarse rows
        // it reads into LDS, so its tiles go one at a time
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            double q[1] = {0.0};
            coarse_rows_lds<VT, kF>(I, static_cast<const VT*>(I.lvl_b[j + 1]), L.need + tile * L.need_stride, 0,
                                    L.need_stride, lds);
            if (i < n) {
                const VT rhs_i = fetch<kF>(&rhs[i]);
                const VT x2 = (pre != nullptr ? fetch<kF>(pre + i) : om * ld(dinv, i) * rhs_i) +
                              static_cast<VT>(lds.cvec[ld(L.agg, i)]);
                VT acc = rhs_i - ld(diag, i) * x2;
                // the neighbour's aggregate comes from the slot's own map, so the
                // correction read does not wait for the column
                ell_row(
                    L.col, off, n, W, i,
                    [&](int s, int c) {
                        const int a = ld(L.agg_col, s * n + i);
                        return (pre != nullptr ? fetch<kF>(pre + c) : om * ld(dinv, c) * fetch<kF>(&rhs[c])) +
                               static_cast<VT>(lds.cvec[a >= 0 ? a : 0]);
                    },
                    [&](VT o, VT v) { acc -= o * v; });
                const VT value = x2 + om * ld(dinv, i) * acc;
                if constexpr (sizeof(VT) == 4) {
                    if (so.out != nullptr) {
                        const double v = split_value<kF>(I, so, i, value);
                        publish(&so.out[i], v);
                        q[0] = so.dot != nullptr ? v * fetch<kF>(&so.dot[i]) : 0.0;
                    }
                }
                if (so.out == nullptr) {
                    publish(&out[i], value);
                    q[0] = dot_with != nullptr ? static_cast<double>(value) * static_cast<double>(fetch<kF>(&dot_with[i])) : 0.0;
                }
            }
            if (dot_part != nullptr) store_partials<1>(dot_part, tile, q, lds);
        }
        return;
    }
    tile_loop<sizeof(VT) == 4 || kW == 2>(tiles, g, G, [&]<int R>(const int (&tl)[R]) {
        int i[R];
        bool ok[R];
        VT x2[R], acc[R];
#pragma unroll
        for (int r = 0; r < R; ++r) i[r] = tile_row(tl[r], n, ok[r]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const VT rhs_i = fetch<kF>(&rhs[i[r]]);
            x2[r] = (pre != nullptr ? fetch<kF>(pre + i[r]) : om * ld(dinv, i[r]) * rhs_i) + fetch<kF>(&corr[ld(L.agg, i[r])]);
            acc[r] = rhs_i - ld(diag, i[r]) * x2[r];
        }
        ell_rows<R>(
            L.col, off, n, W, i,
            [&](int r, int s, int c) {
                const int a = ld(L.agg_col, s * n + i[r]);
                return (pre != nullptr ? fetch<kF>(pre + c) : om * ld(dinv, c) * fetch<kF>(&rhs[c])) +
                       fetch<kF>(&corr[a >= 0 ? a : 0]);
            },
            [&](int r, VT o, VT v) { acc[r] -= o * v; });
        double q[R][1];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            q[r][0] = 0.0;
            if (ok[r]) {
                const VT value = x2[r] + om * ld(dinv, i[r]) * acc[r];
                if constexpr (sizeof(VT) == 4) {
                    if (so.out != nullptr) {
                        const double v = split_value<kF>(I, so, i[r], value);
                        publish(&so.out[i[r]], v);
                        if (so.dot != nullptr) q[r][0] = v * fetch<kF>(&so.dot[i[r]]);
                    }
                }
                if (so.out == nullptr) {
                    publish(&out[i[r]], value);
                    if (dot_with != nullptr)
                        q[r][0] = static_cast<double>(value) * static_cast<double>(fetch<kF>(&dot_with[i[r]]));
                }
            }
        }
        if (dot_part != nullptr) store_partials_n<1, R>(dot_part, tl, q, lds);
    });
}

// One tile of merged_up_phase (below) with every global load it makes
// issued before the coarse solve: the coarse rows' ids and their inverse
// columns, this thread's level-0 row (columns, values, aggregate positions,
// its own inputs and its neighbours' gathers) and its level-(j+1) row (the
// same). Only LDS values (coarse solution, level-(j+1) results) are waited
// for between the steps, so the tile costs three dependent memory round
// trips (ids -> columns -> gathers) instead of six. The loads are
// branch-free (slots past the width reload the last slot, rows past the
// level reload row 0, and the results are discarded by selects), so the
// compiler keeps them all in flight instead of waiting at every join.
// Operations, their order and every rounding are those of the loop in
// merged_up_phase. kPre / kPreM: levels j / j + 1 have two-sweep iterates.
template <class VT, bool kF, bool kPre, bool kPreM>
__device__ __forceinline__ void merged_up_tile_early(const Instance& I, const Values<VT>& V, int j, int tile,
                                                     const VT* rhs, VT* out, Lds& lds, double* dot_part,
                                                     const VT* dot_with, const VT* pre, const VT* pre_m,
                                                     const SplitOut& so) {
    const LevelView& L = I.lv[j];
    const LevelView& M = I.lv[j + 1];
    const int n = L.n, W = L.width, nm = M.n, Wm = M.width;
    const VT om = static_cast<VT>(I.omega[j]), omm = static_cast<VT>(I.omega[j + 1]);
    const auto diag = V.diag(j);
    const auto dinv = V.dinv(j);
    const auto off = V.off(j);
    const auto diag_m = V.diag(j + 1);
    const auto dinv_m = V.dinv(j + 1);
    const auto off_m = V.off(j + 1);
    const VT* rhs_m = static_cast<const VT*>(I.lvl_b[j + 1]);
    const auto col = gptr(L.col);
    const auto col_m = gptr(M.col);
    double* mid = lds.tile[0];
    const bool split = sizeof(VT) == 4 && so.out != nullptr;
    // the arrays of the output step, or a valid array of the same type
    // (loaded and discarded) where there is none
    const double* split_w = split ? so.w : static_cast<const double*>(I.dinv64[0]);
    const double* split_dot = split && so.dot != nullptr ? so.dot : static_cast<const double*>(I.dinv64[0]);
    const VT* dot_in = !split && dot_with != nullptr ? dot_with : rhs;
    // first memory round trip: ids, static row data and own inputs
    const int mq = threadIdx.x;
    const int m_id = ld(L.mid, tile * L.mid_stride + min(mq, L.mid_stride - 1));
    const int m = mq < L.mid_stride ? m_id : -1;
    const int32_t* clist = L.mid_coarse + tile * L.mid_coarse_stride;
    CoarseFirst cf = coarse_rows_issue<true>(I, clist, 0, L.mid_coarse_stride);
    const int i0 = tile * kTile + static_cast<int>(threadIdx.x);
    const bool ok = i0 < n;
    const int i = ok ? i0 : 0;
    int c[kMaxWidth], al[kMaxWidth];
    VT o[kMaxWidth];
#pragma unroll
    for (int s = 0; s < kMaxWidth; ++s) {
        const int e = min(s, W - 1) * n + i;
        c[s] = s < W ? col[e] : -1;
        o[s] = off[e];
        al[s] = ld(L.agg_col_loc, e);
    }
    const VT rhs_i = fetch<kF>(&rhs[i]);
    const VT pre_i = kPre ? fetch<kF>(pre + i) : VT(0);
    const VT dinv_i = ld(dinv, i), diag_i = ld(diag, i);
    const int agl = ld(L.agg_loc, i);
    const double sd = ld(gptr(static_cast<const double*>(I.dinv64[0])), i);
    const double sw = fetch<kF>(&split_w[i]);
    const double sdot = fetch<kF>(&split_dot[i]);
    const double dv = static_cast<double>(fetch<kF>(&dot_in[i]));
    // second round trip: the level-(j+1) row (its id came in the first) and
    // the level-0 gathers
    const int mm = m >= 0 ? m : 0;
    int cm[kMaxWidth], am[kMaxWidth];
    VT omv[kMaxWidth];
#pragma unroll
    for (int s = 0; s < kMaxWidth; ++s) {
        const int e = min(s, Wm - 1) * nm + mm;
        const int cv = col_m[e];
        cm[s] = s < Wm ? cv : -1;
        omv[s] = off_m[e];
        am[s] = ld(M.agg_col, e);
    }
    const VT rhs_mm = fetch<kF>(&rhs_m[mm]);
    const VT pre_mm = kPreM ? fetch<kF>(pre_m + mm) : VT(0);
    const VT dinv_mm = ld(dinv_m, mm), diag_mm = ld(diag_m, mm);
    const int agg_mm = ld(M.agg, mm);
    VT gv[kMaxWidth];
#pragma unroll
    for (int s = 0; s < kMaxWidth; ++s) {
        const int cc = c[s] >= 0 ? c[s] : i;
        if constexpr (kPre)
            gv[s] = fetch<kF>(pre + cc);
        else
            gv[s] = om * ld(dinv, cc) * fetch<kF>(&rhs[cc]);
    }
    // third round trip: the level-(j+1) gathers
    VT gm[kMaxWidth];
#pragma unroll
    for (int s = 0; s < kMaxWidth; ++s) {
        const int cc = cm[s] >= 0 ? cm[s] : mm;
        if constexpr (kPreM)
            gm[s] = fetch<kF>(pre_m + cc);
        else
            gm[s] = omm * ld(dinv_m, cc) * fetch<kF>(&rhs_m[cc]);
    }
    // the coarse rows into lds.cvec
    coarse_rows_finish<VT, kF, true>(I, static_cast<const VT*>(I.lvl_b[j + 2]), clist, 0, L.mid_coarse_stride, lds, cf);
    if (m >= 0) {
        const VT x2 = (kPreM ? pre_mm : omm * dinv_mm * rhs_mm) + static_cast<VT>(lds.cvec[agg_mm]);
        VT acc = rhs_mm - diag_mm * x2;
#pragma unroll
        for (int s = 0; s < kMaxWidth; ++s)
            if (cm[s] >= 0) acc -= omv[s] * (gm[s] + static_cast<VT>(lds.cvec[am[s] >= 0 ? am[s] : 0]));
        mid[mq] = static_cast<double>(static_cast<VT>(x2 + omm * dinv_mm * acc));
    }
    __syncthreads();
    double q[1] = {0.0};
    if (ok) {
        const VT x2 = (kPre ? pre_i : om * dinv_i * rhs_i) + static_cast<VT>(mid[agl]);
        VT acc = rhs_i - diag_i * x2;
#pragma unroll
        for (int s = 0; s < kMaxWidth; ++s)
            if (c[s] >= 0) acc -= o[s] * (gv[s] + static_cast<VT>(mid[al[s] >= 0 ? al[s] : 0]));
        const VT value = x2 + om * dinv_i * acc;
        if (split) {
            // split_value's operations on the loaded inputs
            const double v = I.omega[0] * sd * sw + static_cast<double>(value);
            publish(&so.out[i], v);
            q[0] = so.dot != nullptr ? v * sdot : 0.0;
        } else {
            publish(&out[i], value);
            q[0] = dot_with != nullptr ? static_cast<double>(value) * dv : 0.0;
        }
    }
    if (dot_part != nullptr) store_partials<1>(dot_part, tile, q, lds);
    __syncthreads();
}

// Prolongation from the coarsest level through level j + 1 and on to level
// j (two levels above the coarsest) in one sweep, for groups with about a
// workgroup per tile of level j: each tile solves the coarsest rows its
// level-(j+1) rows read (coarse_rows_lds), post-smooths those level-(j+1)
// rows into LDS with exactly up_phase's operations, then prolongs and
// post-smooths its own rows reading them from LDS. Every value is the one
// the two separate phases compute; the barrier between them is gone.
//
// Called out of line: inlined, its live ranges would push the solve's
// registers into scratch.
template <class VT, int kW>
__device__ __noinline__ void merged_up_phase(const Instance* instance, int j, const VT* rhs, VT* out, int g, int G,
                                             Lds& lds, double* dot_part, const VT* dot_with, int mask,
                                             SplitOut so = SplitOut{}) {
    const Instance& I = *uniform_ptr(instance);
    const Values<VT> V{I};
    j = __builtin_amdgcn_readfirstlane(j);
    g = __builtin_amdgcn_readfirstlane(g);
    G = __builtin_amdgcn_readfirstlane(G);
    mask = __builtin_amdgcn_readfirstlane(mask);
    rhs = uniform_ptr(rhs);
    out = uniform_ptr(out);
    so = SplitOut{uniform_ptr(so.w), uniform_ptr(so.out), uniform_ptr(so.dot)};
    // the two-sweep pre-smoothed iterates of level j and j + 1, where used
    const VT* pre = (mask >> j) & 1 ? uniform_ptr(static_cast<const VT*>(I.lvl_x[j])) : nullptr;
    const VT* pre_m = (mask >> (j + 1)) & 1 ? uniform_ptr(static_cast<const VT*>(I.lvl_x[j + 1])) : nullptr;
    const LevelView& L = I.lv[j];
    const LevelView& M = I.lv[j + 1];
    const int n = L.n, W = L.width, nm = M.n, Wm = M.width;
    const VT om = static_cast<VT>(I.omega[j]), omm = static_cast<VT>(I.omega[j + 1]);
    const auto diag = V.diag(j);
    const auto dinv = V.dinv(j);
    const auto off = V.off(j);
    const auto diag_m = V.diag(j + 1);
    const auto dinv_m = V.dinv(j + 1);
    const auto off_m = V.off(j + 1);
    const VT* rhs_m = static_cast<const VT*>(I.lvl_b[j + 1]);
    constexpr bool kF = kW >= 1;
    double* mid = lds.tile[0];
    const int tiles = (n + kTile - 1) / kTile;
    // With the register budget of one or two workgroups per CU, a tile
    // issues every global load of the phase before the coarse solve
    // (merged_up_tile_early): three dependent memory round trips instead of
    // six, the same operations in the same order.
    if (kW <= 2 && W >= 1 && W <= kMaxWidth && Wm >= 1 && Wm <= kMaxWidth && L.mid_stride >= 1 &&
        L.mid_coarse_stride >= 1) {
        for (int tile = g; tile < tiles; tile += G) {
            if (pre != nullptr && pre_m != nullptr)
                merged_up_tile_early<VT, kF, true, true>(I, V, j, tile, rhs, out, lds, dot_part, dot_with, pre, pre_m, so);
            else if (pre != nullptr)
                merged_up_tile_early<VT, kF, true, false>(I, V, j, tile, rhs, out, lds, dot_part, dot_with, pre, pre_m, so);
            else if (pre_m != nullptr)
                merged_up_tile_early<VT, kF, false, true>(I, V, j, tile, rhs, out, lds, dot_part, dot_with, pre, pre_m, so);
            else
                merged_up_tile_early<VT, kF, false, false>(I, V, j, tile, rhs, out, lds, dot_part, dot_with, pre, pre_m, so);
        }
        return;
    }
    for (int tile = g; tile < tiles; tile += G) {
        // this thread's mid row id (a tile reads at most kTile of them) goes
        // out before the coarse solve
        const int mq = threadIdx.x;
        const int m = mq < L.mid_stride ? ld(L.mid, tile * L.mid_stride + mq) : -1;
        coarse_rows_lds<VT, kF>(I, static_cast<const VT*>(I.lvl_b[j + 2]), L.mid_coarse + tile * L.mid_coarse_stride, 0,
                                L.mid_coarse_stride, lds);
        {
            if (m >= 0) {
                const VT rhs_mm = fetch<kF>(&rhs_m[m]);
                const VT x2 = (pre_m != nullptr ? fetch<kF>(pre_m + m) : omm * ld(dinv_m, m) * rhs_mm) +
                              static_cast<VT>(lds.cvec[ld(M.agg, m)]);
                VT acc = rhs_mm - ld(diag_m, m) * x2;
                ell_row(
                    M.col, off_m, nm, Wm, m,
                    [&](int s, int c) {
                        const int a = ld(M.agg_col, s * nm + m);
                        return (pre_m != nullptr ? fetch<kF>(pre_m + c) : omm * ld(dinv_m, c) * fetch<kF>(&rhs_m[c])) +
                               static_cast<VT>(lds.cvec[a >= 0 ? a : 0]);
                    },
                    [&](VT o, VT v) { acc -= o * v; });
                mid[mq] = static_cast<double>(static_cast<VT>(x2 + omm * ld(dinv_m, m) * acc));
            }
        }
        __syncthreads();
        const int i = tile * kTile + threadIdx.x;
        double q[1] = {0.0};
        if (i < n) {
            const VT rhs_i = fetch<kF>(&rhs[i]);
            const VT x2 = (pre != nullptr ? fetch<kF>(pre + i) : om * ld(dinv, i) * rhs_i) + static_cast<VT>(mid[ld(L.agg_loc, i)]);
            VT acc = rhs_i - ld(diag, i) * x2;
            ell_row(
                L.col, off, n, W, i,
                [&](int s, int c) {
                    const int a = ld(L.agg_col_loc, s * n + i);
                    return (pre != nullptr ? fetch<kF>(pre + c) : om * ld(dinv, c) * fetch<kF>(&rhs[c])) +
                           static_cast<VT>(mid[a >= 0 ? a : 0]);
                },
                [&](VT o, VT v) { acc -= o * v; });
            const VT value = x2 + om * ld(dinv, i) * acc;
            if constexpr (sizeof(VT) == 4) {
                if (so.out != nullptr) {
                    const double v = split_value<kF>(I, so, i, value);
                    publish(&so.out[i], v);
                    q[0] = so.dot != nullptr ? v * fetch<kF>(&so.dot[i]) : 0.0;
                }
            }
            if (so.out == nullptr) {
                publish(&out[i], value);
                q[0] = dot_with != nullptr ? static_cast<double>(value) * static_cast<double>(fetch<kF>(&dot_with[i])) : 0.0;
            }
        }
        if (dot_part != nullptr) store_partials<1>(dot_part, tile, q, lds);
        __syncthreads();
    }
}

// First sweep of a V-cycle: level 0's down phase, or, when level 0 takes
// two pre-sweeps (bit 0 of `mask`), its first Jacobi sweep (input
// evaluated from `src`, stored into `keep` and `keep2`, squares into
// `part`); on a one-level hierarchy the input is only materialised. Ends
// with a barrier.
template <class VT, int kW = 0>
__device__ __forceinline__ void vcycle_first(const Instance& I, const Values<VT>& V, const Source<VT>& src, VT* keep, double* part,
                             int g, int G, Lds& lds, int& timeouts, int mask = 0, VT* keep2 = nullptr,
                             VT* au = nullptr, bool split = false) {
    constexpr bool kF = kW >= 1;   // lean barriers, coherent loads
    if (split && I.depth > 1) {
        // the split V-cycle's FP64 first sweep: r1 = w - A om D^-1 w into
        // level 0's FP32 iterate buffer, the input of the FP32 part
        jacobi_phase<VT, kW>(I, V, 0, src, keep, keep2, nullptr, part, g, G, lds, au,
                             static_cast<float*>(I.lvl_x[0]));
    } else if (I.depth == 1) {
        for (int tile = g; tile < I.tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            double sq[1] = {0.0};
            if (i < I.n) {
                const VT wi = src.template at<kF>(i);
                if (keep != nullptr) publish(&keep[i], wi);
                if (keep2 != nullptr) publish(&keep2[i], wi);
                sq[0] = static_cast<double>(wi) * static_cast<double>(wi);
            }
            if (part != nullptr) store_partials<1>(part, tile, sq, lds);
        }
    } else if (mask & 1) {
        jacobi_phase<VT, kW>(I, V, 0, src, keep, keep2, static_cast<VT*>(I.lvl_x[0]), part, g, G, lds, au);
    } else {
        down_phase<VT, kW>(I, V, 0, src, keep, static_cast<VT*>(I.lvl_b[1]), part, g, G, lds, nullptr, keep2, au);
    }
    group_sync<kF>(timeouts);
}

// Tiles the first sweep's partials cover.
__device__ inline int first_sweep_tiles(const Instance& I, int mask) {
    return I.depth == 1 || (mask & 1) ? I.tiles : I.lv[0].down_tiles;
}

// The rest of the V-cycle after vcycle_first; w0 is the materialised level-0
// input, out0 the level-0 output. Levels whose bit is set in `mask` take a
// second pre-smoothing sweep: a Jacobi phase (level 0's ran in
// vcycle_first), then the restriction sweep from its iterate. Ends with a
// barrier.
//
// Split (the mixed form, VT = float, `so.out` set): vcycle_first ran the
// FP64 first sweep and left r1 = w - A x1 in level 0's FP32 iterate buffer;
// this FP32 V-cycle then runs on r1 with one pre-sweep from zero on level 0
// (the second of its two) and writes x1 + y to so.out in FP64.
//
// keep0 (set when the previous chained step already ran level 0's first
// Jacobi sweep, so vcycle_first did not run): the level-0 down sweep stores
// its input w0 there, as the first sweep would have.
template <class VT, int kW>
__device__ __forceinline__ void vcycle_rest(const Instance& I, const Values<VT>& V, const VT* w0, VT* out0, int g, int G, Lds& lds,
                            int& timeouts, double* dot_part = nullptr, const VT* dot_with = nullptr, int mask = 0,
                            SplitOut so = SplitOut{}, VT* keep0 = nullptr) {
    constexpr bool kF = kW >= 1;
    const int J = I.depth;
    if (J == 1) {
        for (int tile = g; tile < I.tiles; tile += G) {
            const int first = tile * kTile, rows = min(kTile, I.n - first);
            coarse_rows_lds<VT, kF>(I, w0, nullptr, first, rows, lds);
            const int i = first + threadIdx.x;
            double q[1] = {0.0};
            if (i < I.n) {
                const VT value = static_cast<VT>(lds.cvec[i]);
                publish(&out0[i], value);
                if (dot_with != nullptr) q[0] = static_cast<double>(value) * static_cast<double>(fetch<kF>(&dot_with[i]));
            }
            if (dot_part != nullptr) store_partials<1>(dot_part, tile, q, lds);
        }
        group_sync<kF>(timeouts);
        return;
    }
    const bool split = so.out != nullptr;
    if (split) {
        // level 0 of the FP32 part: input r1, one pre-sweep from zero
        w0 = static_cast<const VT*>(I.lvl_x[0]);
        mask &= ~1;
    }
    const Source<VT> none{nullptr, nullptr, nullptr, 0.0, 0.0, 0};
    const auto pre_of = [&](int j) { return (mask >> j) & 1 ? static_cast<VT*>(I.lvl_x[j]) : nullptr; };
    if (split || (mask & 1)) {
        Source<VT> src = none;
        src.a = w0;
        down_phase<VT, kW>(I, V, 0, src, keep0, static_cast<VT*>(I.lvl_b[1]), nullptr, g, G, lds, pre_of(0));
        group_sync<kF>(timeouts);
    }
    for (int j = 1; j + 1 < J; ++j) {
        Source<VT> src = none;
        src.a = static_cast<const VT*>(I.lvl_b[j]);
        if ((mask >> j) & 1) {
            jacobi_phase<VT, kW>(I, V, j, src, nullptr, nullptr, pre_of(j), nullptr, g, G, lds);
            group_sync<kF>(timeouts);
        }
        down_phase<VT, kW>(I, V, j, src, nullptr, static_cast<VT*>(I.lvl_b[j + 1]), nullptr, g, G, lds, pre_of(j));
        group_sync<kF>(timeouts);
    }
    // with about a workgroup per tile, the last two prolongations are one
    // sweep (merged_up_phase); with fewer workgroups each tile would redo
    // the coarse work, so they stay two
    const int jm = J - 3;
    const bool merge = jm >= 0 && I.lv[jm].mid != nullptr && 2 * G >= (I.lv[jm].n + kTile - 1) / kTile;
    for (int j = J - 2; j >= 0; --j) {
        const VT* rhs = j == 0 ? w0 : static_cast<const VT*>(I.lvl_b[j]);
        VT* out = j == 0 ? out0 : static_cast<VT*>(I.lvl_e[j]);
        const SplitOut level_so = j == 0 ? so : SplitOut{};
        if (merge && j == jm + 1) continue;
        if (merge && j == jm)
            merged_up_phase<VT, kW>(&I, j, rhs, out, g, G, lds, j == 0 ? dot_part : nullptr, j == 0 ? dot_with : nullptr,
                                    mask, level_so);
        else
            up_phase<VT, kW>(I, V, j, rhs, static_cast<const VT*>(I.lvl_e[j + 1]), out, g, G, lds,
                             j == 0 ? dot_part : nullptr, j == 0 ? dot_with : nullptr, pre_of(j), level_so);
        group_sync<kF>(timeouts);
    }
}

// The FP32 part of the split V-cycle (mixed form) out of line: its phases
// are the FP32 instances of the V-cycle's, which would otherwise be inlined
// beside the FP64 ones into the solve and add to its registers. Returns the
// barrier timeout state.
template <int kW>
__device__ __noinline__ int vcycle_rest_split(const Instance* instance, const double* w0, double* out0, int g, int G,
                                              Lds& lds, int timeouts, double* dot_part, const double* dot_with,
                                              int mask) {
    const Instance& I = *uniform_ptr(instance);
    g = __builtin_amdgcn_readfirstlane(g);
    G = __builtin_amdgcn_readfirstlane(G);
    mask = __builtin_amdgcn_readfirstlane(mask);
    const SplitOut so{uniform_ptr(w0), uniform_ptr(out0), uniform_ptr(dot_with)};
    vcycle_rest<float, kW>(I, Values<float>{I}, nullptr, nullptr, g, G, lds, timeouts, uniform_ptr(dot_part), nullptr,
                           mask, so);
    return timeouts;
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

template <int kW> __device__ void coarse_inverse(const Instance& I, double* tri, Lds& lds) {
    const int j = I.depth - 1;
    const LevelView& L = I.lv[j];
    const int n = L.n, W = L.width;
    const int packed = n * (n + 1) / 2;
    const double* diag = I.diag64[j];
    const double* off = I.off64[j];
    for (int e = threadIdx.x; e < packed; e += blockDim.x) tri[e] = 0.0;
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
        tri[tri_index(r, r)] = ld(diag, r);
        for (int s = 0; s < W; ++s) {
            const int c = L.col[s * n + r];
            if (c >= 0 && c < r) tri[tri_index(r, c)] = off[s * n + r];
        }
    }
    __syncthreads();
    // Entries are strided over the threads (entry e = t + 256 m), so every
    // thread's loads and stores are independent and consecutive threads
    // touch consecutive entries; the row of entry e is recovered from
    // e = r (r + 1) / 2 + c.
    for (int k = 0; k < n; ++k) {
        for (int c = threadIdx.x; c < n; c += blockDim.x) lds.colk[c] = tri[tri_index(c, k)];
        __syncthreads();
        const double dinv = 1.0 / lds.colk[k];
        for (int e = threadIdx.x; e < packed; e += blockDim.x) {
            int r = static_cast<int>((sqrtf(8.0f * static_cast<float>(e) + 1.0f) - 1.0f) * 0.5f);
            if (r * (r + 1) / 2 > e) --r;
            if ((r + 1) * (r + 2) / 2 <= e) ++r;
            const int c = e - r * (r + 1) / 2;
            double value;
            if (r == k)
                value = c == k ? -dinv : lds.colk[c] * dinv;
            else if (c == k)
                value = lds.colk[r] * dinv;
            else
                value = tri[e] - lds.colk[r] * lds.colk[c] * dinv;
            tri[e] = value;
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int col = e / n, row = e - col * n;
        I.cinv[e] = -tri[tri_index(row, col)];
    }
}

// The same sweep for a coarsest level of at most kBlockedCoarse rows, by
// 16 x 16 blocks on the matrix cores (one workgroup). The matrix is padded
// with identity rows to N = 16 nb and its lower block triangle kept in
// I.tri, each block in the MFMA accumulator order (lane l holds rows
// l / 16 + 4 v, v = 0..3, of column l % 16, at [l * 4 + v]). Sweeping the
// 16 pivots of block K is one block step: with S = SWEEP(A_KK) = -A_KK^-1
// (its 16 scalar sweeps on the LDS copy, each entry computed from its
// lower-triangle formula so S is exactly symmetric),
//   D_i = A_iK S,   A_ij += D_i A_Kj  (i, j != K),   A_iK = -D_i,   A_KK = S,
// the products by v_mfma_f64_16x16x4 (four per 16 x 16 x 16 product, FP64
// throughout). After all nb steps I.tri holds -A^-1; the result is written
// negated to I.cinv from the lower triangle, mirrored, so it is exactly
// symmetric.
__device__ inline int block_index(int bi, int bj) { return bi * (bi + 1) / 2 + bj; }
// Position of entry (r, c) of a 16 x 16 block in accumulator order.
__device__ inline int acc_pos(int r, int c) { return ((r & 3) * kMB + c) * 4 + (r >> 2); }

template <int kW> __device__ __noinline__ void coarse_inverse_blocked(const Instance* instance, Lds& lds) {
    const Instance& I = *uniform_ptr(instance);
    const int j = I.depth - 1;
    const LevelView& L = I.lv[j];
    const int n = L.n, W = L.width;
    const int nb = (n + kMB - 1) / kMB;
    const int blocks = nb * (nb + 1) / 2;
    const int t = static_cast<int>(threadIdx.x);
    const int lane = t & 63, wave = t >> 6;
    double* M = I.tri;
    // entry (r, c), r >= c, of the stored lower triangle
    const auto at = [](int r, int c) {
        return block_index(r / kMB, c / kMB) * (kMB * kMB) + acc_pos(r % kMB, c % kMB);
    };
    for (int e = t; e < blocks * kMB * kMB; e += kTile) M[e] = 0.0;
    __syncthreads();
    for (int r = t; r < nb * kMB; r += kTile) {
        if (r >= n) {
            M[at(r, r)] = 1.0;   // identity padding
            continue;
        }
        M[at(r, r)] = I.diag64[j][r];
        for (int s = 0; s < W; ++s) {
            const int c = L.col[s * n + r];
            if (c >= 0 && c < r) M[at(r, c)] = I.off64[j][s * n + r];
        }
    }
    __syncthreads();
    const int er = t % kMB, ec = t / kMB;   // this thread's entry of a 16 x 16 block
    for (int K = 0; K < nb; ++K) {
        // 1. block column K as it is before the step (column-major blocks),
        // and the pivot block, symmetric from its lower half, row-major
        for (int i = 0; i < nb; ++i) {
            const int r = i * kMB + er, c = K * kMB + ec;   // entry (r, c) of A
            lds.colT[i][ec * kMB + er] = r >= c ? M[at(r, c)] : M[at(c, r)];
        }
        {
            const int hi = er > ec ? er : ec, lo = er > ec ? ec : er;
            lds.piv[0][er * kMB + ec] = M[at(K * kMB + hi, K * kMB + lo)];
        }
        __syncthreads();
        // 2. S = SWEEP(A_KK): 16 scalar sweeps; entries (r, c) and (c, r)
        // evaluate the same lower-triangle expression
        int cur = 0;
        {
            const int r = er > ec ? er : ec, c = er > ec ? ec : er;
            for (int k = 0; k < kMB; ++k) {
                const double* Pc = lds.piv[cur];
                const double pivot = Pc[k * kMB + k];
                const double rk = Pc[r * kMB + k], ck = Pc[k * kMB + c], rc = Pc[r * kMB + c];
                const double dinv = 1.0 / pivot;
                double value;
                if (r == k && c == k)
                    value = -dinv;
                else if (c == k)
                    value = rk * dinv;
                else if (r == k)
                    value = ck * dinv;
                else
                    value = fma(-(rk * dinv), ck, rc);
                lds.piv[cur ^ 1][er * kMB + ec] = value;
                __syncthreads();
                cur ^= 1;
            }
        }
        const double* S = lds.piv[cur];
        // 3. D_i = A_iK S for i != K (wave w takes i = w, w + 4, ...): kept in
        // LDS (column-major) for the update, -D_i stored as the new A_iK
        for (int i = wave; i < nb; i += 4) {
            if (i == K) continue;
            v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                const int k = 4 * kc + lane / 16;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(lds.colT[i][k * kMB + lane % 16], S[k * kMB + lane % 16],
                                                            acc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = lane / 16 + 4 * v, c = lane % 16;   // entry (r, c) of D_i
                lds.dT[i][c * kMB + r] = acc[v];
                if (i > K)
                    M[block_index(i, K) * (kMB * kMB) + lane * 4 + v] = -acc[v];
                else
                    M[block_index(K, i) * (kMB * kMB) + acc_pos(c, r)] = -acc[v];
            }
        }
        // the new pivot block: S
        M[block_index(K, K) * (kMB * kMB) + t] = S[((t / 4) / 16 + 4 * (t % 4)) * kMB + (t / 4) % 16];
        __syncthreads();
        // 4. A_ij += D_i A_Kj over the stored blocks off row and column K;
        // A_Kj[k][c] = A_jK[c][k] = colT[j][k * 16 + c]
        for (int b = wave; b < blocks; b += 4) {
            int bi = static_cast<int>((sqrtf(8.0f * static_cast<float>(b) + 1.0f) - 1.0f) * 0.5f);
            if (bi * (bi + 1) / 2 > b) --bi;
            if ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
            const int bj = b - bi * (bi + 1) / 2;
            if (bi == K || bj == K) continue;
            double* blk = M + b * (kMB * kMB) + lane * 4;
            v4d acc = {blk[0], blk[1], blk[2], blk[3]};
#pragma unroll
            for (int kc = 0; kc < 4; ++kc) {
                const int k = 4 * kc + lane / 16;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(lds.dT[bi][k * kMB + lane % 16],
                                                            lds.colT[bj][k * kMB + lane % 16], acc, 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v) blk[v] = acc[v];
        }
        __syncthreads();
    }
    // I.cinv = -M, each stored lower entry to both of its places
    for (int e = t; e < blocks * kMB * kMB; e += kTile) {
        const int b = e / (kMB * kMB), q = e % (kMB * kMB), l = q / 4, v = q % 4;
        int bi = static_cast<int>((sqrtf(8.0f * static_cast<float>(b) + 1.0f) - 1.0f) * 0.5f);
        if (bi * (bi + 1) / 2 > b) --bi;
        if ((bi + 1) * (bi + 2) / 2 <= b) ++bi;
        const int bj = b - bi * (bi + 1) / 2;
        const int r = bi * kMB + l / 16 + 4 * v, c = bj * kMB + l % 16;
        if (r < n && c < n && r >= c) {
            const double value = -M[e];
            I.cinv[c * n + r] = value;
            I.cinv[r * n + c] = value;
        }
    }
}

template <int kW>
__device__ __forceinline__ void set_matrix(const Instance& I, const Launch& P, int g, int G, Lds& lds, int& timeouts,
                                           int& parity) {
    const int J = I.depth;
    {
        // Level-0 values into ELL: assembled from the heat inputs
        // (diag = C A + dt sum_k G_k, off_k = -dt G_k, G_k = K_k w_k, in CSR
        // order), or copied from CSR-ordered off-diagonals.
        const LevelView& L = I.lv[0];
        const int n = L.n, W = L.width;
        for (int tile = g; tile < I.tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            if (i < n) {
                const int k0 = L.row[i], deg = L.row[i + 1] - k0;
                double sum = 0.0;
                for (int s = 0; s < W; ++s) {
                    double value = 0.0;
                    if (s < deg) {
                        if (P.op & kOpHeatAssemble) {
                            const double gk = ld(P.entry_conductance, k0 + s) * ld(I.weight, k0 + s);
                            sum += gk;
                            value = -P.dt * gk;
                        } else {
                            value = ld(P.in, k0 + s);
                        }
                    }
                    I.off64[0][s * n + i] = value;
                }
                if (P.op & kOpHeatAssemble) {
                    const double ca = ld(P.capacity_area, i);
                    I.diag64[0][i] = ca + P.dt * sum;
                    if (P.capacity_keep != nullptr) P.capacity_keep[i] = ca;
                }
            }
        }
        group_sync(timeouts);
    }
    for (int j = 0; j < J; ++j) {
        const LevelView& L = I.lv[j];
        const int n = L.n, W = L.width;
        double* part = I.partial + parity * I.partial_slots * kPartials;
        parity ^= 1;
        const int tiles = (n + kTile - 1) / kTile;
        // inverse diagonal, FP32 copies, Gershgorin ratio sum |off| / diag
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            double ratio = 0.0;
            if (i < n) {
                const double d = I.diag64[j][i];
                const double di = 1.0 / d;
                I.dinv64[j][i] = di;
                double s = 0.0;
                for (int q = 0; q < W; ++q) {
                    const double o = I.off64[j][q * n + i];
                    s += fabs(o);
                    if (I.off32[j] != nullptr) I.off32[j][q * n + i] = static_cast<float>(o);
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
            const LevelView& C = I.lv[j + 1];
            const int nc = C.n, slots = C.width * C.n;
            const int work_tiles = (nc + slots + kTile - 1) / kTile;
            for (int tile = g; tile < work_tiles; tile += G) {
                const int e = tile * kTile + threadIdx.x;
                // sums in list order; operands fetched sixteen at a time (a
                // coarse diagonal gathers ~50 fine values), codes first, then
                // all values, so a chunk costs two round trips
                if (e < nc) {
                    double s = 0.0;
                    const int q1 = L.gd_start[e + 1];
                    for (int q = L.gd_start[e]; q < q1; q += 16) {
                        int code[16];
                        double v[16];
#pragma unroll
                        for (int u = 0; u < 16; ++u) code[u] = q + u < q1 ? L.gd[q + u] : 0;
#pragma unroll
                        for (int u = 0; u < 16; ++u)
                            v[u] = q + u < q1 ? (code[u] < 0 ? I.diag64[j][-code[u] - 1] : I.off64[j][code[u]]) : 0.0;
#pragma unroll
                        for (int u = 0; u < 16; ++u)
                            if (q + u < q1) s += v[u];
                    }
                    I.diag64[j + 1][e] = s;
                } else if (e < nc + slots) {
                    const int p = e - nc;
                    double s = 0.0;
                    const int q1 = L.go_start[p + 1];
                    for (int q = L.go_start[p]; q < q1; q += 4) {
                        double v[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) v[u] = q + u < q1 ? I.off64[j][L.go[q + u]] : 0.0;
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (q + u < q1) s += v[u];
                    }
                    I.off64[j + 1][p] = s;
                }
            }
        }
        group_sync(timeouts);
        const double rho = group_max<false>(part, tiles, lds);   // after the fenced group_sync above
        // Chebyshev degree-1 weight on [a, b] = [max(1 - rho, (1 + rho) / 4), 1 + rho]
        const double hi = 1.0 + rho;
        const double lo = fmax(1.0 - rho, 0.25 * hi);
        if (g == 0 && threadIdx.x == 0) I.omega[j] = 2.0 / (lo + hi);
    }
    if (g == 0) {
        if (I.lv[J - 1].n <= kBlockedCoarse)
            coarse_inverse_blocked<kW>(&I, lds);
        else
            coarse_inverse<kW>(I, I.tri_lds ? dynamic_lds() : I.tri, lds);
    }
    group_sync(timeouts);
}

// set_matrix out of line (its register-blocked inverse would otherwise add
// to the solver's register pressure), with the barrier state passed by
// value so that it stays in registers on both sides.
struct SyncState {
    int timeouts, parity;
};
template <int kW>
__device__ __noinline__ SyncState set_matrix_call(const Instance& I, const Launch& P, int g, int G, Lds& lds,
                                                  SyncState st) {
    set_matrix<kW>(I, P, g, G, lds, st.timeouts, st.parity);
    return st;
}

// ------------------------------------------------------------ BiCGSTAB

// FP64 true residual row of x over level 0 (ELL): b_i - (diag_i x_i + sum).
template <bool kF = false>
__device__ inline double residual_row(const Instance& I, const double* b, const double* x, int i) {
    const LevelView& L = I.lv[0];
    const int n = L.n, W = L.width;
    double ax = I.diag64[0][i] * fetch<kF>(&x[i]);
    ell_row(
        L.col, I.off64[0], n, W, i, [&](int, int c) { return fetch<kF>(&x[c]); },
        [&](double o, double v) { ax += o * v; });
    return fetch<kF>(&b[i]) - ax;
}

// Levels that pre-smooth with two Jacobi sweeps in a solve: the system's
// mask, or, when it is negative (automatic), every level whose Jacobi
// weight is damped below 1 (Gershgorin ratio above 0.6: the smoother then
// reduces the error slowly), plus level 0 in the FP64 form, where one
// V-cycle must carry a half-step to the tolerance on its own.
__device__ inline int pre_sweep_mask(const Instance& I) {
    if (I.pre2 >= 0) return I.pre2;
    int mask = I.precision == 0 && I.depth > 1 ? 1 : 0;
    for (int j = 0; j + 1 < I.depth; ++j)
        if (I.omega[j] != 1.0) mask |= 1 << j;
    return mask;
}

// A chained step's right-hand side prepared by the previous solve: set when
// the phase that left that solve's x final also assembled the next b, r
// and r0, with the totals (b, b) and (r, r) the next solve starts from.
struct Prepared {
    bool ready = false;
    double totals[2] = {0.0, 0.0};
    // the next solve's first level-0 Jacobi sweep is done too: its two-sweep
    // iterate is in level 0's iterate buffer and A w in u (plain form)
    bool first_sweep = false;
    // (mixed form) the search-direction buffer that holds the next solve's
    // first direction: the one the preparing solve's current direction is not
    int p_next = 0;
};

// kForm: 0 the plain form, 1 the mixed form's split V-cycle, 2 either, by
// the instance's precision at run time.
template <class VT, int kW, int kForm = 0>
//
// Chained steps in the full-register-file program (kW == 1): with
// next_forcing set, every phase that leaves x final (the half-step exit and
// the FP64 true residual) also assembles the next step's right-hand side
// b' = C A x + dt A S' into the other (b, r, r0) pair (rp selects the pair in
// use), with its tile partials beside the phase's own, summed with them;
// `prepared` then holds those totals and the next solve starts from them
// instead of an assembly phase (the same values and partial sums, so the
// same bits), zeroing its x in its first SpMV sweep.
__device__ __forceinline__ SolveResult solve(const Instance& I, const Launch& P, const double* t_old, const double* forcing,
                             const double* next_forcing, double* t_new, int g, int G, Lds& lds, int& timeouts,
                             int& parity, int& xi, int& rp, Prepared& prepared) {
    constexpr bool kPrep = kW == 1;
    constexpr bool kF = kW >= 1;   // lean barriers, coherent loads
    // the solve's device clock (only the group's first workgroup reports it)
    const uint64_t t_begin = g == 0 ? static_cast<uint64_t>(wall_clock64()) : 0;
    const Values<VT> V{I};
    const LevelView& L0 = I.lv[0];
    const int n = I.n, W = L0.width, tiles = I.tiles;
    const bool swap = kPrep && prepared.ready;
    bool use_prepared_sweep = swap && prepared.first_sweep;
    if (swap) rp ^= 1;
    const bool alt = kPrep && rp != 0;
    VT* r = static_cast<VT*>(alt ? I.r_alt : I.r);
    VT* r0 = static_cast<VT*>(alt ? I.r0_alt : I.r0);
    double* b = alt ? I.b_alt : I.b;
    VT* v = static_cast<VT*>(I.v);
    VT* ph = static_cast<VT*>(I.ph);
    VT* s = static_cast<VT*>(I.s);
    VT* sh = static_cast<VT*>(I.sh);
    VT* t = static_cast<VT*>(I.t);
    constexpr int KQ = kPrep ? 4 : 2;   // partials of a phase that leaves x final
    // row i's share of the next step's right-hand side from its final x:
    // q[2], q[3] get (b', b') and (r', r') as the assembly phase sums them
    const auto prepare_row = [&](int i, double x_final, double (&q)[KQ]) -> double {
        if constexpr (kPrep) {
            double* bn = alt ? I.b : I.b_alt;
            VT* rn = static_cast<VT*>(alt ? I.r : I.r_alt);
            VT* rn0 = static_cast<VT*>(alt ? I.r0 : I.r0_alt);
            const double bi = ld(P.capacity_area, i) * x_final + P.dt * ld(I.area, i) * ld(next_forcing, i);
            const VT rv = static_cast<VT>(bi);
            publish(&bn[i], bi);
            publish(&rn[i], rv);
            publish(&rn0[i], rv);
            q[2] = bi * bi;
            q[3] = static_cast<double>(rv) * static_cast<double>(rv);
            return static_cast<double>(rv);
        }
        return 0.0;
    };
    Prepared prep;   // set by the last phase that prepared the next step
    // totals of a phase that leaves x final: its slots 0 and 1 into tot, and
    // with next_forcing set, the prepared right-hand side's (b', b') and
    // (r', r') (slots 2 and 3) into prep, from one read of the partials
    const auto final_totals = [&](const double* part_in, double (&tot)[2], bool swept = false) {
        if (kPrep && next_forcing != nullptr) {
            double t4[4];
            group_totals<4, kF>(part_in, tiles, t4, lds);
            tot[0] = t4[0];
            tot[1] = t4[1];
            prep.ready = true;
            prep.totals[0] = t4[2];
            prep.totals[1] = t4[3];
            prep.first_sweep = swept;
        } else {
            group_totals<2, kF>(part_in, tiles, tot, lds);
        }
    };
    const auto diag = V.diag(0);
    const auto off = V.off(0);
    const auto next_partials = [&]() {
        double* part = I.partial + parity * I.partial_slots * kPartials;
        parity ^= 1;
        return part;
    };
    // acc[r] = (A in)_i[r] for R rows at once
    const auto spmv_rows = [&]<int R>(const VT* in, const int (&i)[R], VT (&acc)[R]) {
#pragma unroll
        for (int r = 0; r < R; ++r) acc[r] = ld(diag, i[r]) * fetch<kF>(&in[i[r]]);
        ell_rows<R>(
            L0.col, off, n, W, i, [&](int, int, int c) { return fetch<kF>(&in[c]); },
            [&](int r, VT o, VT v) { acc[r] += o * v; });
    };

    // the mixed form: FP64 vectors and first level-0 sweep, the rest of the
    // V-cycle in FP32 (vcycle_rest_split)
    const bool split = sizeof(VT) == 8 && I.depth > 1 && (kForm == 1 || (kForm == 2 && I.precision == 1));
    const int mask = pre_sweep_mask(I) | (split ? 1 : 0);   // levels with two pre-smoothing sweeps

    // b (assembled or given), x = 0, r = r0 = b; or, after a prepared
    // step, b, r and r0 are in place and x is zeroed in the first SpMV sweep.
    double* part = nullptr;
    double tot2[2];
    bool zero_x = swap;
    // A heat step whose V-cycle starts with a level-0 Jacobi sweep (in the
    // full-register-file program; the batch and server programs measured
    // slower with it) assembles its right-hand side in that sweep's phase: each row forms its b and
    // recomputes its neighbours' from T, C A, A and S (what the assembly
    // would store there, operation for operation), then runs the sweep on
    // w = r = b, as the previous step's last phase does in a chain. x is
    // left to the first x update (known to be zero until then).
    const bool assemble_swept = kW == 1 && !swap && sizeof(VT) == 8 && (P.op & kOpHeatRhs) != 0 && I.depth > 1 &&
                                ((mask & 1) != 0 || split) && W >= 1 && W <= kMaxWidth && P.max_iterations > 0;
    if (swap) {
        tot2[0] = prepared.totals[0];
        tot2[1] = prepared.totals[1];
    } else if (assemble_swept) {
        part = next_partials();
        const double om0 = I.omega[0];
        const auto dinv0 = gptr(static_cast<const double*>(I.dinv64[0]));
        const auto diag0 = gptr(static_cast<const double*>(I.diag64[0]));
        const auto off0 = gptr(static_cast<const double*>(I.off64[0]));
        const auto col0 = gptr(L0.col);
        for (int tile = g; tile < tiles; tile += G) {
            const int i0 = tile * kTile + static_cast<int>(threadIdx.x);
            const bool ok = i0 < n;
            const int i = ok ? i0 : 0;
            int c[kMaxWidth];
            double o[kMaxWidth];
#pragma unroll
            for (int k = 0; k < kMaxWidth; ++k) {
                const int e = min(k, W - 1) * n + i;
                const int cv = col0[e];
                c[k] = k < W ? cv : -1;
                o[k] = off0[e];
            }
            const double t_i = fetch<kF>(&t_old[i]);
            const double ca_i = ld(P.capacity_area, i), area_i = ld(I.area, i), f_i = ld(forcing, i);
            const double diag_i = diag0[i], dinv_i = dinv0[i];
            double tc[kMaxWidth], cac[kMaxWidth], areac[kMaxWidth], fc[kMaxWidth], dc[kMaxWidth];
#pragma unroll
            for (int k = 0; k < kMaxWidth; ++k) {
                const int cc = c[k] >= 0 ? c[k] : i;
                tc[k] = fetch<kF>(&t_old[cc]);
                cac[k] = ld(P.capacity_area, cc);
                areac[k] = ld(I.area, cc);
                fc[k] = ld(forcing, cc);
                dc[k] = dinv0[cc];
            }
            double q[2] = {0.0, 0.0};
            if (ok) {
                const double bi = ca_i * t_i + P.dt * area_i * f_i;
                const VT rv = static_cast<VT>(bi);
                publish(&b[i], bi);
                publish(&r[i], rv);
                publish(&r0[i], rv);
                q[0] = bi * bi;
                q[1] = static_cast<double>(rv) * static_cast<double>(rv);
                const double w = static_cast<double>(rv);
                double acc = w - diag_i * (om0 * dinv_i * w);
                double aw = diag_i * w;
#pragma unroll
                for (int k = 0; k < kMaxWidth; ++k)
                    if (c[k] >= 0) {
                        const double wc = static_cast<double>(static_cast<VT>(cac[k] * tc[k] + P.dt * areac[k] * fc[k]));
                        acc -= o[k] * (om0 * dc[k] * wc);
                        aw += o[k] * wc;
                    }
                if (split) {
                    // the split V-cycle's first sweep: r1 in FP32, w kept as
                    // the first search direction
                    publish(&static_cast<float*>(I.lvl_x[0])[i], static_cast<float>(acc));
                    publish(&static_cast<double*>(I.p[1])[i], w);
                } else {
                    const double x1 = om0 * dinv_i * w;
                    publish(&static_cast<double*>(I.lvl_x[0])[i], x1 + om0 * dinv_i * acc);
                }
                publish(&static_cast<double*>(I.u)[i], aw);
            }
            store_partials<2>(part, tile, q, lds);
        }
        group_sync<kF>(timeouts);
        group_totals<2, kF>(part, tiles, tot2, lds);
        zero_x = true;
        use_prepared_sweep = true;
    } else {
        part = next_partials();
        {
            double* x = I.x[xi];
            tile_loop(tiles, g, G, [&]<int R>(const int (&tl)[R]) {
                int i[R];
                bool ok[R];
                double bi[R], q[R][2];
#pragma unroll
                for (int k = 0; k < R; ++k) i[k] = tile_row(tl[k], n, ok[k]);
#pragma unroll
                for (int k = 0; k < R; ++k)
                    bi[k] = (P.op & kOpHeatRhs) ? ld(P.capacity_area, i[k]) * fetch<kF>(&t_old[i[k]]) + P.dt * ld(I.area, i[k]) * ld(forcing, i[k])
                                                : ld(P.in, i[k]);
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    q[k][0] = q[k][1] = 0.0;
                    if (ok[k]) {
                        publish(&b[i[k]], bi[k]);
                        publish(&x[i[k]], 0.0);
                        const VT ri = static_cast<VT>(bi[k]);
                        publish(&r[i[k]], ri);
                        publish(&r0[i[k]], ri);
                        q[k][0] = bi[k] * bi[k];
                        q[k][1] = static_cast<double>(ri) * static_cast<double>(ri);
                    }
                }
                store_partials_n<2, R>(part, tl, q, lds);
            });
        }
        group_sync<kF>(timeouts);
        group_totals<2, kF>(part, tiles, tot2, lds);
    }
    const int prepared_p_next = prepared.p_next;
    prepared = Prepared{};
    const double b_norm = sqrt(tot2[0]);
    double rho1 = tot2[1];
    double r_norm = b_norm;
    int iterations = 0, restarts = 0, v_cycles = 0;
    int status = 0;
    const auto vcycle_rest_any = [&](const VT* w, VT* out, double* dp, const VT* dw) {
        if constexpr (kForm != 0 && sizeof(VT) == 8) {
            if (split) {
                if constexpr (kW == 1) {
                    // the whole register file: the FP32 part inlined (a
                    // separate out-of-line solve instance, run_heat_solve)
                    vcycle_rest<float, kW>(I, Values<float>{I}, nullptr, nullptr, g, G, lds, timeouts, dp, nullptr, mask,
                                           SplitOut{w, out, dw});
                } else {
                    timeouts = vcycle_rest_split<kW>(&I, w, out, g, G, lds, timeouts, dp, dw, mask);
                }
                return;
            }
        }
        vcycle_rest<VT, kW>(I, V, w, out, g, G, lds, timeouts, dp, dw, mask);
    };
    // heat steps solve the assembled heat matrix, which is symmetric
    const bool symmetric = (P.op & kOpHeatRhs) != 0;
    bool shadow_stale = false;    // r0 must be set from r at the next pass's start
    bool x_zero = true;           // x is known to be 0 (no x update since the start)
    const double target = P.tolerance * b_norm;
    bool first = true;            // first pass: r = b exactly
    bool have_residual = false;   // the true residual of x is in r, r_norm, rho1
    // current search-direction buffer; a prepared split step's first
    // direction was kept by the previous step in the buffer that was not its
    // own direction (the assembled form keeps it in p[1], the default)
    int pi = swap && use_prepared_sweep && split ? (prepared_p_next ^ 1) : 0;
    if (zero_x && !(b_norm > 0.0 && rho1 != 0.0 && P.max_iterations > 0)) [[unlikely]] {
        // no SpMV sweep will run before x is read: zero it here
        double* x = I.x[xi];
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            if (i < n) publish(&x[i], 0.0);
        }
        group_sync<kF>(timeouts);
        zero_x = false;
    }

    if (b_norm > 0.0) {
        for (;;) {
            if (!first) {
                if (!have_residual) [[unlikely]] {
                    // FP64 true residual r = b - A x; restart from x if needed.
                    part = next_partials();
                    const double* x = I.x[xi];
                    for (int tile = g; tile < tiles; tile += G) {
                        const int i = tile * kTile + threadIdx.x;
                        double q[KQ] = {};
                        if (i < n) {
                            const double ri = residual_row<kF>(I, b, x, i);
                            const VT rv = static_cast<VT>(ri);
                            publish(&r[i], rv);
                            publish(&r0[i], rv);
                            q[0] = ri * ri;
                            q[1] = static_cast<double>(rv) * static_cast<double>(rv);
                            if (kPrep && next_forcing != nullptr) prepare_row(i, fetch<kF>(&x[i]), q);
                        }
                        store_partials<KQ>(part, tile, q, lds);
                    }
                    group_sync<kF>(timeouts);
                    final_totals(part, tot2);
                    r_norm = sqrt(tot2[0]);
                    rho1 = tot2[1];
                }
                have_residual = false;
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
            bool first_iteration = true;
            while (iterations < P.max_iterations) {
                if (rho1 == 0.0 || omega == 0.0) break;
                const double beta = rho1 / rho * (alpha / omega);
                rho = rho1;
                // V-cycle on p = r + beta (p - omega v); p = r on a fresh start,
                // whose sweep also refreshes the shadow residual after a restart
                Source<VT> src{r, nullptr, nullptr, 0.0, 0.0, 0};
                if (!first_iteration) src = Source<VT>{r, static_cast<const VT*>(I.p[pi]), v, beta, omega, 1};
                VT* p_new = static_cast<VT*>(I.p[pi ^ 1]);
                pi ^= 1;
                // A symmetric (heat systems): (r0, A ph) = (A r0, ph), with A r0
                // from the first sweep's gathers and the product from the
                // V-cycle's last sweep, so alpha is known when the V-cycle
                // ends and the SpMV joins the true-residual phase below. FP64
                // inner vectors only: A r0 (about |C A| |b|) overflows FP32.
                const bool fuse = sizeof(VT) == 8 && first_iteration && symmetric && I.depth > 1;
                VT* u = static_cast<VT*>(I.u);
                // the previous chained step ran this V-cycle's first sweep on
                // w = r (its iterate and A w are in place): start at the down
                // sweep, which stores w into p
                const bool prepared_sweep = use_prepared_sweep && first_iteration && fuse && !shadow_stale;
                use_prepared_sweep = false;
                if (!prepared_sweep)
                    vcycle_first<VT, kW>(I, V, src, p_new, nullptr, g, G, lds, timeouts, mask,
                                         first_iteration && shadow_stale ? r0 : nullptr, fuse ? u : nullptr, split);
                shadow_stale = false;
                double* dot = fuse ? next_partials() : nullptr;
                if (prepared_sweep) {
                    if (split) {
                        // the split V-cycle after its first sweep (r1 and A w
                        // in place, w kept in p): its FP32 part on r1, output
                        // x1 + y with x1 from w = r
                        vcycle_rest_any(r, ph, dot, u);
                    } else {
                        vcycle_rest<VT, kW>(I, V, r, ph, g, G, lds, timeouts, dot, u, mask, SplitOut{}, p_new);
                    }
                } else
                    vcycle_rest_any(p_new, ph, dot, fuse ? u : nullptr);
                ++v_cycles;
                // fused: alpha = rho / (A r0, ph) from the V-cycle's partials,
                // read inside the next phase once its loads are in flight
                bool alpha_ready = !fuse;
                if (fuse) {
                } else [[unlikely]] {
                // v = A ph, (r0, v)
                part = next_partials();
                tile_loop<sizeof(VT) == 4 || kW == 2>(tiles, g, G, [&]<int R>(const int (&tl)[R]) {
                    int i[R];
                    bool ok[R];
                    VT acc[R], rz[R];
                    double q[R][1];
#pragma unroll
                    for (int k = 0; k < R; ++k) i[k] = tile_row(tl[k], n, ok[k]);
#pragma unroll
                    for (int k = 0; k < R; ++k) rz[k] = fetch<kF>(&r0[i[k]]);
                    spmv_rows(ph, i, acc);
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        q[k][0] = 0.0;
                        if (ok[k]) {
                            publish(&v[i[k]], acc[k]);
                            q[k][0] = static_cast<double>(rz[k]) * static_cast<double>(acc[k]);
                            if (kPrep && zero_x) publish(&I.x[xi][i[k]], 0.0);
                        }
                    }
                    store_partials_n<1, R>(part, tl, q, lds);
                });
                zero_x = false;
                group_sync<kF>(timeouts);
                double r0v[1];
                group_totals<1, kF>(part, tiles, r0v, lds);
                if (r0v[0] == 0.0) break;
                alpha = rho / r0v[0];
                }
                bool alpha_in_x = false;   // x already holds x + alpha ph
                if (first_iteration) [[likely]] {
                    // r is the true residual of x here, so s = r - alpha v is
                    // b - A (x + alpha ph) up to rounding: compute it as that
                    // FP64 true residual with x' = x + alpha ph, in the phase
                    // that a half-step exit would run anyway, and test it
                    part = next_partials();
                    const double* x = I.x[xi];
                    double* xn = I.x[xi ^ 1];
                    // Chained steps, plain form: this phase also runs the next
                    // solve's first level-0 Jacobi sweep on w' = r' = b' (what
                    // jacobi_phase would compute there, operation for
                    // operation): each row recomputes its neighbours' b'_c from
                    // x_c and ph_c, which it gathers anyway, and C A, A and S'
                    // at c, so the next solve starts at its down sweep.
                    // (the split V-cycle's first sweep is the same sweep, with
                    // its residual r1 = w' - A x1 stored in FP32 instead of the
                    // two-sweep iterate, and w' kept as the next search
                    // direction, which the down sweep cannot store from r1)
                    const bool sweep_next = kPrep && next_forcing != nullptr && (kForm == 0 || split) &&
                                            (mask & 1) != 0 && I.depth > 1 && sizeof(VT) == 8;
                    const double om0 = I.omega[0];
                    const auto dinv0 = gptr(static_cast<const double*>(I.dinv64[0]));
                    const auto diag0 = gptr(static_cast<const double*>(I.diag64[0]));
                    const auto off0 = gptr(static_cast<const double*>(I.off64[0]));
                    const auto col0 = gptr(L0.col);
                    // the next step's forcing, or a valid array of the same size
                    // (loaded and discarded) when there is none
                    const double* nforce = kPrep && next_forcing != nullptr ? next_forcing : I.area;
                    // Every load of a row is issued first and without a branch
                    // (slots past the width reload the last slot, rows past n
                    // reload row 0), then alpha is taken from the previous
                    // phase's partials while they are in flight; the
                    // operations and their order are those of the fused
                    // phase's definition: x' = x + alpha ph, s = r = b - A x',
                    // v = A ph and, in a chain, the next b, r, r0 and the next
                    // solve's first level-0 Jacobi sweep.
                    const auto take_alpha = [&]() {
                        if (!alpha_ready) {
                            double uph[1];
                            group_totals<1, kF>(dot, tiles, uph, lds);
                            alpha_ready = true;
                            if (uph[0] == 0.0) return false;
                            alpha = rho / uph[0];
                        }
                        return true;
                    };
                    bool stop = false;
                    if (W < 1 || W > kMaxWidth) [[unlikely]] {
                        // rows wider than the unrolled slots: alpha first,
                        // then the row loop
                        stop = !take_alpha();
                        if (!stop) {
                        for (int tile = g; tile < tiles; tile += G) {
                            const int i = tile * kTile + threadIdx.x;
                            double q[KQ] = {};
                            if (i < n) {
                                const double ph_i = static_cast<double>(fetch<kF>(&ph[i]));
                                const double x_new = (x_zero ? 0.0 : fetch<kF>(&x[i])) + alpha * ph_i;
                                publish(&xn[i], x_new);
                                double w_next = 0.0;
                                if (kPrep && next_forcing != nullptr) w_next = prepare_row(i, x_new, q);
                                double ax = I.diag64[0][i] * x_new;
                                double av = I.diag64[0][i] * ph_i;   // (A ph)_i when the SpMV phase was skipped
                                // the next sweep's residual and A w' at row i
                                double acc_next = 0.0, aw_next = 0.0;
                                if (sweep_next) {
                                    acc_next = w_next - I.diag64[0][i] * (om0 * ld(dinv0, i) * w_next);
                                    aw_next = I.diag64[0][i] * w_next;
                                }
                                struct Gathered {
                                    double x, ph, ca, area, forcing, dinv;
                                };
                                ell_row(
                                    L0.col, I.off64[0], n, W, i,
                                    [&](int, int c) {
                                        Gathered w{x_zero ? 0.0 : fetch<kF>(&x[c]), static_cast<double>(fetch<kF>(&ph[c])),
                                                   0.0, 0.0, 0.0, 0.0};
                                        if (sweep_next) {
                                            w.ca = ld(P.capacity_area, c);
                                            w.area = ld(I.area, c);
                                            w.forcing = ld(next_forcing, c);
                                            w.dinv = ld(dinv0, c);
                                        }
                                        return w;
                                    },
                                    [&](double o, const Gathered& w) {
                                        const double xc = w.x + alpha * w.ph;
                                        ax += o * xc;
                                        av += o * w.ph;
                                        if (sweep_next) {
                                            // b'_c as prepare_row forms it at row c
                                            const double wc = w.ca * xc + P.dt * w.area * w.forcing;
                                            acc_next -= o * (om0 * w.dinv * wc);
                                            aw_next += o * wc;
                                        }
                                    });
                                const double si = fetch<kF>(&b[i]) - ax;
                                const VT sv = static_cast<VT>(si);
                                publish(&s[i], sv);
                                publish(&r[i], sv);
                                if (fuse) publish(&v[i], static_cast<VT>(av));
                                if (sweep_next) {
                                    if (split) {
                                        publish(&static_cast<float*>(I.lvl_x[0])[i], static_cast<float>(acc_next));
                                        publish(&static_cast<double*>(I.p[pi ^ 1])[i], w_next);
                                    } else {
                                        const double x1 = om0 * ld(dinv0, i) * w_next;
                                        publish(&static_cast<double*>(I.lvl_x[0])[i], x1 + om0 * ld(dinv0, i) * acc_next);
                                    }
                                    publish(&static_cast<double*>(I.u)[i], aw_next);
                                }
                                q[0] = si * si;
                                q[1] = static_cast<double>(sv) * static_cast<double>(sv);
                            }
                            store_partials<KQ>(part, tile, q, lds);
                        }
                        }
                    }
                    for (int tile = g; tile < tiles && !stop && W >= 1 && W <= kMaxWidth; tile += G) {
                        const int i0 = tile * kTile + static_cast<int>(threadIdx.x);
                        const bool ok = i0 < n;
                        const int i = ok ? i0 : 0;
                        int c[kMaxWidth];
                        double o[kMaxWidth];
#pragma unroll
                        for (int k = 0; k < kMaxWidth; ++k) {
                            const int e = min(k, W - 1) * n + i;
                            const int cv = col0[e];
                            c[k] = k < W ? cv : -1;
                            o[k] = off0[e];
                        }
                        const double ph_i = static_cast<double>(fetch<kF>(&ph[i]));
                        const double x_i = fetch<kF>(&x[i]);
                        const double b_i = fetch<kF>(&b[i]);
                        const double diag_i = diag0[i], dinv_i = dinv0[i];
                        double ca_i = 0.0, area_i = 0.0, nf_i = 0.0;
                        if constexpr (kPrep) {
                            ca_i = ld(P.capacity_area, i);
                            area_i = ld(I.area, i);
                            nf_i = ld(nforce, i);
                        }
                        double gx[kMaxWidth], gph[kMaxWidth], gca[kMaxWidth], garea[kMaxWidth], gf[kMaxWidth],
                            gd[kMaxWidth];
#pragma unroll
                        for (int k = 0; k < kMaxWidth; ++k) {
                            const int cc = c[k] >= 0 ? c[k] : i;
                            gx[k] = fetch<kF>(&x[cc]);
                            gph[k] = static_cast<double>(fetch<kF>(&ph[cc]));
                            if constexpr (kPrep) {
                                gca[k] = ld(P.capacity_area, cc);
                                garea[k] = ld(I.area, cc);
                                gf[k] = ld(nforce, cc);
                                gd[k] = dinv0[cc];
                            }
                        }
                        if (tile == g && !take_alpha()) [[unlikely]] {
                            stop = true;
                            break;
                        }
                        double q[KQ] = {};
                        if (ok) {
                            const double x_new = (x_zero ? 0.0 : x_i) + alpha * ph_i;
                            publish(&xn[i], x_new);
                            double w_next = 0.0;
                            if constexpr (kPrep) {
                                if (next_forcing != nullptr) {
                                    // prepare_row's operations on the loaded inputs
                                    double* bn = alt ? I.b : I.b_alt;
                                    VT* rn = static_cast<VT*>(alt ? I.r : I.r_alt);
                                    VT* rn0 = static_cast<VT*>(alt ? I.r0 : I.r0_alt);
                                    const double bi = ca_i * x_new + P.dt * area_i * nf_i;
                                    const VT rv = static_cast<VT>(bi);
                                    publish(&bn[i], bi);
                                    publish(&rn[i], rv);
                                    publish(&rn0[i], rv);
                                    q[2] = bi * bi;
                                    q[3] = static_cast<double>(rv) * static_cast<double>(rv);
                                    w_next = static_cast<double>(rv);
                                }
                            }
                            double ax = diag_i * x_new;
                            double av = diag_i * ph_i;   // (A ph)_i when the SpMV phase was skipped
                            // the next sweep's residual and A w' at row i
                            double acc_next = 0.0, aw_next = 0.0;
                            if (sweep_next) {
                                acc_next = w_next - diag_i * (om0 * dinv_i * w_next);
                                aw_next = diag_i * w_next;
                            }
#pragma unroll
                            for (int k = 0; k < kMaxWidth; ++k)
                                if (c[k] >= 0) {
                                    const double xc = (x_zero ? 0.0 : gx[k]) + alpha * gph[k];
                                    ax += o[k] * xc;
                                    av += o[k] * gph[k];
                                    if constexpr (kPrep) {
                                        if (sweep_next) {
                                            // b'_c as the next right-hand side forms it at row c
                                            const double wc = gca[k] * xc + P.dt * garea[k] * gf[k];
                                            acc_next -= o[k] * (om0 * gd[k] * wc);
                                            aw_next += o[k] * wc;
                                        }
                                    }
                                }
                            const double si = b_i - ax;
                            const VT sv = static_cast<VT>(si);
                            publish(&s[i], sv);
                            publish(&r[i], sv);
                            if (fuse) publish(&v[i], static_cast<VT>(av));
                            if (sweep_next) {
                                if (split) {
                                    publish(&static_cast<float*>(I.lvl_x[0])[i], static_cast<float>(acc_next));
                                    publish(&static_cast<double*>(I.p[pi ^ 1])[i], w_next);
                                } else {
                                    const double x1 = om0 * dinv_i * w_next;
                                    publish(&static_cast<double*>(I.lvl_x[0])[i], x1 + om0 * dinv_i * acc_next);
                                }
                                publish(&static_cast<double*>(I.u)[i], aw_next);
                            }
                            q[0] = si * si;
                            q[1] = static_cast<double>(sv) * static_cast<double>(sv);
                        }
                        store_partials<KQ>(part, tile, q, lds);
                    }
                    if (!alpha_ready && !take_alpha()) stop = true;
                    if (stop) [[unlikely]] {
                        // breakdown: this phase's partial buffer was not used;
                        // give it back, so the next phase writes the buffer
                        // that no workgroup can still be reading
                        parity ^= 1;
                        break;
                    }
                    group_sync<kF>(timeouts);
                    xi ^= 1;
                    x_zero = false;
                    zero_x = false;   // x[xi] now holds x + alpha ph
                    final_totals(part, tot2, sweep_next);
                    prep.p_next = pi ^ 1;   // where this phase kept the next first direction
                    ++iterations;
                    if (sqrt(tot2[0]) <= inner_target) [[likely]] {
                        // r holds the true residual; a restart takes it as its
                        // shadow residual too
                        r_norm = sqrt(tot2[0]);
                        rho1 = tot2[1];
                        have_residual = true;
                        shadow_stale = true;
                        break;
                    }
                    alpha_in_x = true;
                    vcycle_first<VT, kW>(I, V, Source<VT>{s, nullptr, nullptr, 0.0, 0.0, 0}, nullptr, nullptr, g, G, lds,
                                         timeouts, mask, nullptr, nullptr, split);
                } else [[unlikely]] {
                    // s = r - alpha v inside the next V-cycle's first sweep, (s, s)
                    part = next_partials();
                    vcycle_first<VT, kW>(I, V, Source<VT>{r, v, nullptr, alpha, 0.0, 2}, s, part, g, G, lds, timeouts,
                                         mask, nullptr, nullptr, split);
                    double ss[1];
                    group_totals<1, kF>(part, first_sweep_tiles(I, mask), ss, lds);
                    ++iterations;
                    if (sqrt(ss[0]) <= inner_target) {
                        // x' = x + alpha ph fused with the FP64 true residual of x'
                        part = next_partials();
                        const double* x = I.x[xi];
                        double* xn = I.x[xi ^ 1];
                        for (int tile = g; tile < tiles; tile += G) {
                            const int i = tile * kTile + threadIdx.x;
                            double q[KQ] = {};
                            if (i < n) {
                                const double x_new = fetch<kF>(&x[i]) + alpha * static_cast<double>(fetch<kF>(&ph[i]));
                                publish(&xn[i], x_new);
                                if (kPrep && next_forcing != nullptr) prepare_row(i, x_new, q);
                                double ax = I.diag64[0][i] * x_new;
                                ell_row(
                                    L0.col, I.off64[0], n, W, i,
                                    [&](int, int c) { return fetch<kF>(&x[c]) + alpha * static_cast<double>(fetch<kF>(&ph[c])); },
                                    [&](double o, double w) { ax += o * w; });
                                const double ri = fetch<kF>(&b[i]) - ax;
                                const VT rv = static_cast<VT>(ri);
                                publish(&r[i], rv);
                                publish(&r0[i], rv);
                                q[0] = ri * ri;
                                q[1] = static_cast<double>(rv) * static_cast<double>(rv);
                            }
                            store_partials<KQ>(part, tile, q, lds);
                        }
                        group_sync<kF>(timeouts);
                        xi ^= 1;
                        final_totals(part, tot2);
                        r_norm = sqrt(tot2[0]);
                        rho1 = tot2[1];
                        have_residual = true;
                        break;
                    }
                }
                first_iteration = false;
                vcycle_rest_any(s, sh, nullptr, nullptr);
                ++v_cycles;
                // t = A sh, (t, t), (t, s)
                part = next_partials();
                tile_loop<sizeof(VT) == 4 || kW == 2>(tiles, g, G, [&]<int R>(const int (&tl)[R]) {
                    int i[R];
                    bool ok[R];
                    VT acc[R], sv[R];
                    double q[R][2];
#pragma unroll
                    for (int k = 0; k < R; ++k) i[k] = tile_row(tl[k], n, ok[k]);
#pragma unroll
                    for (int k = 0; k < R; ++k) sv[k] = fetch<kF>(&s[i[k]]);
                    spmv_rows(sh, i, acc);
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        q[k][0] = q[k][1] = 0.0;
                        if (ok[k]) {
                            publish(&t[i[k]], acc[k]);
                            const double td = static_cast<double>(acc[k]);
                            q[k][0] = td * td;
                            q[k][1] = td * static_cast<double>(sv[k]);
                        }
                    }
                    store_partials_n<2, R>(part, tl, q, lds);
                });
                group_sync<kF>(timeouts);
                double tt[2];
                group_totals<2, kF>(part, tiles, tt, lds);
                omega = tt[0] > 0.0 ? tt[1] / tt[0] : 0.0;
                // x += alpha ph + omega sh (omega sh when x holds x + alpha ph
                // already), r = s - omega t, (r, r), (r0, r)
                part = next_partials();
                prep = Prepared{};
                {
                    double* x = I.x[xi];
                    tile_loop(tiles, g, G, [&]<int R>(const int (&tl)[R]) {
                        int i[R];
                        bool ok[R];
                        double xv[R], pv[R], shv[R], sv[R], tv[R], rz[R], q[R][2];
#pragma unroll
                        for (int k = 0; k < R; ++k) i[k] = tile_row(tl[k], n, ok[k]);
#pragma unroll
                        for (int k = 0; k < R; ++k) {
                            xv[k] = fetch<kF>(&x[i[k]]);
                            pv[k] = alpha_in_x ? 0.0 : static_cast<double>(fetch<kF>(&ph[i[k]]));
                            shv[k] = static_cast<double>(fetch<kF>(&sh[i[k]]));
                            sv[k] = static_cast<double>(fetch<kF>(&s[i[k]]));
                            tv[k] = static_cast<double>(fetch<kF>(&t[i[k]]));
                            rz[k] = static_cast<double>(fetch<kF>(&r0[i[k]]));
                        }
#pragma unroll
                        for (int k = 0; k < R; ++k) {
                            q[k][0] = q[k][1] = 0.0;
                            if (ok[k]) {
                                publish(&x[i[k]], alpha_in_x ? xv[k] + omega * shv[k] : xv[k] + (alpha * pv[k] + omega * shv[k]));
                                const double ri = sv[k] - omega * tv[k];
                                const VT rv = static_cast<VT>(ri);
                                publish(&r[i[k]], rv);
                                q[k][0] = static_cast<double>(rv) * static_cast<double>(rv);
                                q[k][1] = rz[k] * static_cast<double>(rv);
                            }
                        }
                        store_partials_n<2, R>(part, tl, q, lds);
                    });
                }
                group_sync<kF>(timeouts);
                double rr[2];
                group_totals<2, kF>(part, tiles, rr, lds);
                rho1 = rr[1];
                if (sqrt(rr[0]) <= inner_target) break;
            }
        }
    }
    if (zero_x) [[unlikely]] {
        // a prepared step that broke down before any x update: x still
        // holds the previous step's T
        double* x = I.x[xi];
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            if (i < n) publish(&x[i], 0.0);
        }
        group_sync<kF>(timeouts);
        zero_x = false;
    }
    if (t_new != nullptr) {
        const double* x = I.x[xi];
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            if (i < n) publish(&t_new[i], fetch<kF>(&x[i]));
        }
    }
    prepared = prep;
    SolveResult out{};
    out.iterations = iterations;
    out.restarts = restarts;
    out.status = status;
    out.relative_residual = b_norm > 0.0 ? r_norm / b_norm : 0.0;
    out.b_norm = b_norm;
    out.v_cycles = v_cycles;
    out.t_begin = t_begin;
    out.t_end = g == 0 ? static_cast<uint64_t>(wall_clock64()) : 0;
    return out;
}

// ------------------------------------------------------------ PCG

// Preconditioned conjugate gradients on the system's FP64 matrix, the
// reference's Newton-system solve (heat_transport.cpp:328-429): x = 0,
// r = b (held rows already zeroed by the caller; the matrix has identity
// rows there), z = M r, d = z, rho = (r, z), target = tol ||b||; then while
// it < max, ||r|| > target and rho > 0:
//   q = A d, alpha = rho / (d, q), x += alpha d, r -= alpha q,
//   z = M r, rho' = (r, z), beta = rho' / rho, d = z + beta d.
// The operator's diagonal may differ from the V-cycle's matrix (a
// preconditioner regularised against a null mode, pressure_redistribution.
// cpp:48-58,150-178). M is the V-cycle (symmetric: the same Jacobi weights before and after the
// coarse correction), restricted to the free rows when rows are held (held
// rows are decoupled identity rows with b = 0: their direction is 0, as the
// reference's Gauss-Seidel sweep leaves it). d = z + beta d is evaluated inside the next SpMV
// (double-buffered d), and (r, z) inside the V-cycle's last sweep, so an
// iteration is six barrier-separated phases. x goes to P.out.
template <int kW>
__device__ __forceinline__ SolveResult pcg(const Instance& I, const Launch& P, int g, int G, Lds& lds, int& timeouts,
                                           int& parity) {
    constexpr bool kF = kW >= 1;   // lean barriers, coherent loads
    const uint64_t t_begin = g == 0 ? static_cast<uint64_t>(wall_clock64()) : 0;
    const Values<double> V{I};
    const LevelView& L0 = I.lv[0];
    const int n = I.n, W = L0.width, tiles = I.tiles;
    double* r = static_cast<double*>(I.r);
    double* z = static_cast<double*>(I.ph);
    double* q = static_cast<double*>(I.v);
    double* kept = static_cast<double*>(I.s);
    double* x = I.x[0];
    double* d[2] = {static_cast<double*>(I.p[0]), static_cast<double*>(I.p[1])};
    const double* diag = P.operator_diag != nullptr ? P.operator_diag : I.diag64[0];
    const double* off = I.off64[0];
    const auto next_partials = [&]() {
        double* part = I.partial + parity * I.partial_slots * kPartials;
        parity ^= 1;
        return part;
    };
    // r = b, x = 0, (r, r)
    double* part = next_partials();
    for (int tile = g; tile < tiles; tile += G) {
        const int i = tile * kTile + threadIdx.x;
        double v[1] = {0.0};
        if (i < n) {
            const double bi = ld(P.in, i);
            publish(&r[i], bi);
            publish(&x[i], 0.0);
            v[0] = bi * bi;
        }
        store_partials<1>(part, tile, v, lds);
    }
    group_sync<kF>(timeouts);
    double tot[1];
    group_totals<1, kF>(part, tiles, tot, lds);
    const double b_norm = sqrt(tot[0]);
    const double target = P.tolerance * b_norm;
    double r_norm = b_norm;
    // z = M r with (r, z)
    const auto precondition = [&]() {
        const Source<double> src{r, nullptr, nullptr, 0.0, 0.0, 0};
        vcycle_first<double, kW>(I, V, src, kept, nullptr, g, G, lds, timeouts);
        double* dot = next_partials();
        vcycle_rest<double, kW>(I, V, r, z, g, G, lds, timeouts, dot, r);
        double rz[1];
        group_totals<1, kF>(dot, tiles, rz, lds);   // one partial per level-0 tile
        return rz[0];
    };
    double rho = b_norm > 0.0 ? precondition() : 0.0;
    int iterations = 0, di = 0, v_cycles = b_norm > 0.0 ? 1 : 0;
    double beta = 0.0;
    bool first = true;
    while (iterations < P.max_iterations && r_norm > target && rho > 0.0) {
        // d = z + beta d (d = z at first), q = A d, (d, q)
        const Source<double> src = first ? Source<double>{z, nullptr, nullptr, 0.0, 0.0, 0}
                                         : Source<double>{z, d[di], d[di], beta, 0.0, 1};
        double* dn = d[di ^ 1];
        part = next_partials();
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            double v[1] = {0.0};
            if (i < n) {
                // held rows keep d = 0, so x, r and q stay exactly 0 there
                const double di_value = P.hold != nullptr && P.hold[i] != 0 ? 0.0 : src.template at<kF>(i);
                double acc = ld(diag, i) * di_value;
                ell_row(
                    L0.col, off, n, W, i, [&](int, int c) { return src.template at<kF>(c); },
                    [&](double o, double w) { acc += o * w; });
                publish(&dn[i], di_value);
                publish(&q[i], acc);
                v[0] = di_value * acc;
            }
            store_partials<1>(part, tile, v, lds);
        }
        group_sync<kF>(timeouts);
        group_totals<1, kF>(part, tiles, tot, lds);
        di ^= 1;
        first = false;
        const double alpha = rho / tot[0];
        // x += alpha d, r -= alpha q, (r, r)
        part = next_partials();
        for (int tile = g; tile < tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            double v[1] = {0.0};
            if (i < n) {
                const double ri = fetch<kF>(&r[i]) - alpha * fetch<kF>(&q[i]);
                publish(&x[i], fetch<kF>(&x[i]) + alpha * fetch<kF>(&dn[i]));
                publish(&r[i], ri);
                v[0] = ri * ri;
            }
            store_partials<1>(part, tile, v, lds);
        }
        group_sync<kF>(timeouts);
        group_totals<1, kF>(part, tiles, tot, lds);
        r_norm = sqrt(tot[0]);
        ++iterations;
        if (!(iterations < P.max_iterations && r_norm > target)) break;
        const double next_rho = precondition();
        ++v_cycles;
        beta = next_rho / rho;
        rho = next_rho;
    }
    for (int tile = g; tile < tiles; tile += G) {
        const int i = tile * kTile + threadIdx.x;
        if (i < n) publish(&P.out[i], fetch<kF>(&x[i]));
    }
    SolveResult out{};
    out.iterations = iterations;
    out.status = r_norm <= target ? 0 : kNotConverged;
    out.relative_residual = b_norm > 0.0 ? r_norm / b_norm : 0.0;
    out.b_norm = b_norm;
    out.v_cycles = v_cycles;
    out.t_begin = t_begin;
    out.t_end = g == 0 ? static_cast<uint64_t>(wall_clock64()) : 0;
    return out;
}

// PCG out of line (not the hot path; its registers stay out of the solver's).
template <int kW>
__device__ __noinline__ SolveResult pcg_call(const Instance* instance, const Launch& P, int g, int G, Lds& lds,
                                             SyncState& st) {
    const Instance& I = *uniform_ptr(instance);
    int timeouts = st.timeouts, parity = st.parity;
    const SolveResult res = pcg<kW>(I, P, __builtin_amdgcn_readfirstlane(g), __builtin_amdgcn_readfirstlane(G), lds,
                                timeouts, parity);
    st = SyncState{timeouts, parity};
    return res;
}

// ------------------------------------------------------------ the kernel

template <class VT, int kW, int kForm>
__device__ __forceinline__ void run_solve(const Instance& I, const Launch& P, int g, int G, Lds& lds, int& timeouts, int& parity) {
    int xi = 0, rp = 0;
    Prepared prepared;
    const int steps = (P.op & kOpChain) ? P.steps : 1;
    SolveResult last{};
    for (int step = 0; step < steps; ++step) {
        const double* t_old = P.t_old;
        const double* forcing = P.forcing;
        const double* next_forcing = nullptr;
        double* t_new = (P.op & kOpHeatRhs) ? P.t_new : P.out;
        if (P.op & kOpChain) {
            // step k reads the previous step's x (row by row, before this
            // step zeroes it) and writes the final T only at the end; in the
            // full-register-file program each step but the last prepares the
            // next one's right-hand side
            t_old = step == 0 ? P.t_old : I.x[xi];
            forcing = P.forcing + static_cast<size_t>(step) * I.n;
            if (kW == 1 && step + 1 < steps) next_forcing = forcing + I.n;
            t_new = step + 1 == steps ? P.t_new : nullptr;
        }
        last = solve<VT, kW, kForm>(I, P, t_old, forcing, next_forcing, t_new, g, G, lds, timeouts, parity, xi, rp,
                             prepared);
        if (g == 0 && threadIdx.x == 0 && P.step_results != nullptr) P.step_results[step] = last;
    }
    if (g == 0 && threadIdx.x == 0) {
        grp.result->iterations = last.iterations;
        grp.result->restarts = last.restarts;
        grp.result->status = last.status;
        grp.result->relative_residual = last.relative_residual;
        grp.result->b_norm = last.b_norm;
        grp.result->v_cycles = last.v_cycles;
        grp.result->t_begin = last.t_begin;
        grp.result->t_end = last.t_end;
    }
}

// The heat solve in the system's form: FP64 vectors in both, the mixed form
// with the split V-cycle. Where the split lives was chosen by measurement
// (the FP32 part's registers and its call affect the FP64 phases' spills):
// with the whole register file (kW == 1, single instances) the mixed form is
// its own instance of the solve, out of line, with the FP32 part inlined, and
// the plain solve compiles as if the mixed form did not exist; in the
// programs of two and three workgroups per CU (batches, servers) one solve
// picks the split at run time and calls the FP32 part out of line.
template <int kW>
__device__ __noinline__ SyncState run_split_solve_call(const Instance* instance, const Launch& P, int g, int G, Lds& lds,
                                                       SyncState st) {
    const Instance& I = *uniform_ptr(instance);
    g = __builtin_amdgcn_readfirstlane(g);
    G = __builtin_amdgcn_readfirstlane(G);
    run_solve<double, kW, 1>(I, P, g, G, lds, st.timeouts, st.parity);
    return st;
}

template <int kW>
__device__ __forceinline__ void run_heat_solve(const Instance& I, const Launch& P, int g, int G, Lds& lds, int& timeouts,
                                               int& parity) {
    if constexpr (kW == 1) {
        if (I.precision == 1) {
            const SyncState st = run_split_solve_call<kW>(&I, P, g, G, lds, SyncState{timeouts, parity});
            timeouts = st.timeouts;
            parity = st.parity;
        } else {
            run_solve<double, kW, 0>(I, P, g, G, lds, timeouts, parity);
        }
    } else {
        run_solve<double, kW, 2>(I, P, g, G, lds, timeouts, parity);
    }
}

// The solve out of line, for the three-workgroups-per-CU program: inlined
// there, the V-cycle variants of both precisions push the kernel's live
// state into scratch; a call keeps the solve's registers its own.
template <int kW>
__device__ __noinline__ SyncState run_solve_call(const Instance* instance, const Launch& P, int g, int G, Lds& lds,
                                                 SyncState st) {
    const Instance& I = *uniform_ptr(instance);
    g = __builtin_amdgcn_readfirstlane(g);
    G = __builtin_amdgcn_readfirstlane(G);
    run_heat_solve<kW>(I, P, g, G, lds, st.timeouts, st.parity);
    return st;
}

// Three workgroups per CU (12 waves): the register budget that keeps the
// solver's live state out of scratch (a tighter cap spills it) while
// leaving batches 768 co-resident workgroups.
//
// Two instances of the program: program_kernel_t<3> for batches and
// servers (three workgroups per CU), and program_kernel_t<1> for one instance whose
// group has at most one workgroup per CU, where occupancy buys nothing and
// the register file is worth more: one wave per SIMD lets the compiler keep
// the solve's state in registers (512 VGPRs) instead of scratch.
// The program's body. kMixedInline (single mixed systems in the
// full-register-file program): the solve is inlined in the split form
// instead of called out of line, so a launch does not save and restore the
// called solve's registers in scratch memory (about 30 us per separate
// step); a kernel of its own, so the plain program keeps its registers.
template <int kWaves, bool kMixedInline>
__device__ __forceinline__ void program_body(const Instance* instances, const int32_t* block_instance,
                                             const Launch* launches, const Launch& single, Lds& lds, Launch& params) {
    // the workgroup's instance, uniform: its descriptor fields load as scalar
    // loads into SGPRs instead of per-lane vector loads
    const int index = __builtin_amdgcn_readfirstlane(block_instance[blockIdx.x]);
    const Instance& I = *uniform_ptr(instances + index);
    if (threadIdx.x == 0) params = launches != nullptr ? launches[index] : single;
    set_group(I.group_first, I.group_size, I.barrier, I.result, I.trace, I.trace_count, I.trace_capacity);
    const Launch& P = params;
    const int g = static_cast<int>(blockIdx.x) - I.group_first;
    const int G = I.group_size;
    int timeouts = 0;
    int parity = 0;
    if (I.trace != nullptr && threadIdx.x == 0 && g == 0) {
        I.trace_count[0] = 1;
        I.trace[0] = static_cast<uint64_t>(wall_clock64());
    }
    if (P.op & kOpSetMatrix) {
        const SyncState st = set_matrix_call<kWaves>(I, P, g, G, lds, SyncState{timeouts, parity});
        timeouts = st.timeouts;
        parity = st.parity;
    }
    if (P.op & kOpPrecondition) {
        if (I.precision == 0) {
            const Source<double> src{P.in, nullptr, nullptr, 0.0, 0.0, 0};
            double* keep = static_cast<double*>(I.p[0]);
            vcycle_first<double>(I, Values<double>{I}, src, keep, nullptr, g, G, lds, timeouts);
            vcycle_rest<double, kWaves>(I, Values<double>{I}, keep, P.out, g, G, lds, timeouts);
        } else {
            float* pin = static_cast<float*>(I.p[0]);
            float* pout = static_cast<float*>(I.ph);
            for (int tile = g; tile < I.tiles; tile += G) {
                const int i = tile * kTile + threadIdx.x;
                if (i < I.n) pin[i] = static_cast<float>(ld(P.in, i));
            }
            group_sync(timeouts);
            const Source<float> src{pin, nullptr, nullptr, 0.0, 0.0, 0};
            vcycle_first<float>(I, Values<float>{I}, src, static_cast<float*>(I.p[1]), nullptr, g, G, lds, timeouts);
            vcycle_rest<float, kWaves>(I, Values<float>{I}, static_cast<float*>(I.p[1]), pout, g, G, lds, timeouts);
            for (int tile = g; tile < I.tiles; tile += G) {
                const int i = tile * kTile + threadIdx.x;
                if (i < I.n) P.out[i] = static_cast<double>(fetch<true>(&pout[i]));   // after a lean barrier
            }
        }
    }
    if (P.op & kOpApply) {
        const LevelView& L = I.lv[0];
        const int n = L.n, W = L.width;
        for (int tile = g; tile < I.tiles; tile += G) {
            const int i = tile * kTile + threadIdx.x;
            if (i < n) {
                double acc = I.diag64[0][i] * ld(P.in, i);
                ell_row(
                    L.col, I.off64[0], n, W, i, [&](int, int c) { return ld(P.in, c); },
                    [&](double o, double v) { acc += o * v; });
                P.out[i] = acc;
            }
        }
    }
    if (P.op & kOpSolve) {
        if constexpr (kWaves == 1 && kMixedInline) {
            run_solve<double, kWaves, 2>(I, P, g, G, lds, timeouts, parity);
        } else if constexpr (kWaves == 1) {
            run_heat_solve<kWaves>(I, P, g, G, lds, timeouts, parity);
        } else {
            const SyncState st = run_solve_call<kWaves>(&I, P, g, G, lds, SyncState{timeouts, parity});
            timeouts = st.timeouts;
            parity = st.parity;
        }
    }
    if (P.op & kOpPcg) {
        SyncState st{timeouts, parity};
        const SolveResult res = pcg_call<kWaves>(&I, P, g, G, lds, st);
        timeouts = st.timeouts;
        parity = st.parity;
        if (g == 0 && threadIdx.x == 0) {
            grp.result->iterations = res.iterations;
            grp.result->restarts = 0;
            grp.result->status = res.status;
            grp.result->relative_residual = res.relative_residual;
            grp.result->b_norm = res.b_norm;
            grp.result->v_cycles = res.v_cycles;
            grp.result->t_begin = res.t_begin;
            grp.result->t_end = res.t_end;
        }
    }
    if (timeouts != 0 && threadIdx.x == 0) atomicAdd(&grp.result->barrier_timeouts, timeouts);
    group_exit();
}

template <int kWaves>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(kWaves))) void program_kernel_t(
    const Instance* instances, const int32_t* block_instance, const Launch* launches, Launch single) {
    __shared__ Lds lds;
    __shared__ Launch params;
    program_body<kWaves, false>(instances, block_instance, launches, single, lds, params);
}

// One mixed-precision system in the full-register-file program.
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(1))) void program_kernel_mixed(
    const Instance* instances, const int32_t* block_instance, const Launch* launches, Launch single) {
    __shared__ Lds lds;
    __shared__ Launch params;
    program_body<1, true>(instances, block_instance, launches, single, lds, params);
}

template __global__ void program_kernel_t<3>(const Instance*, const int32_t*, const Launch*, Launch);
template __global__ void program_kernel_t<2>(const Instance*, const int32_t*, const Launch*, Launch);
template __global__ void program_kernel_t<1>(const Instance*, const int32_t*, const Launch*, Launch);

// ------------------------------------------------------------ the server

constexpr long long kMemberGrace = 1000000000;   // 10 s past the lifetime

__device__ inline uint64_t load_system(const uint64_t* p) {
    return __hip_atomic_load(const_cast<uint64_t*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Slot leader: wait for a request on the slot's queue and claim it (its
// ring position), or return -1 once the host asks to stop or the launch's
// lifetime is over while the slot is idle.
__device__ int64_t claim_request(const ServeParams& S, const ServeQueue& Q, long long start) {
    uint64_t head = __hip_atomic_load(Q.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int idle = 0;   // empty polls so far: the wait between polls doubles up to ~14 us
    for (;;) {
        const uint64_t tail =
            __hip_atomic_load(const_cast<uint64_t*>(Q.tail), __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        while (head < tail) {
            if (__hip_atomic_compare_exchange_strong(Q.head, &head, head + 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT))
                return static_cast<int64_t>(head);
        }
        if (__hip_atomic_load(const_cast<uint32_t*>(S.stop), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0)
            return -1;
        if (wall_clock64() - start > static_cast<long long>(S.lifetime)) return -1;
        // Every poll of the tail crosses the host link; idle leaders back off
        // so that hundreds of them do not load it while other slots solve.
        const int naps = 1 << (idle < 4 ? idle : 4);
        for (int k = 0; k < naps; ++k) __builtin_amdgcn_s_sleep(32);
        ++idle;
        head = __hip_atomic_load(Q.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One request on the slot's group, out of line so that the server loop's
// own state does not add to the solver's register pressure.
template <int kW>
__device__ __noinline__ int serve_one(const Instance* instance, const Launch& P, int g, int G, Lds& lds,
                                      int timeouts) {
    const Instance& I = *uniform_ptr(instance);
    g = __builtin_amdgcn_readfirstlane(g);
    G = __builtin_amdgcn_readfirstlane(G);
    int parity = 0;
    if (P.op & kOpSetMatrix) {
        const SyncState st = set_matrix_call<kW>(I, P, g, G, lds, SyncState{timeouts, parity});
        timeouts = st.timeouts;
        parity = st.parity;
    }
    run_heat_solve<kW>(I, P, g, G, lds, timeouts, parity);
    // every workgroup's rows of T are written before the completion
    group_sync(timeouts);
    return timeouts;
}

// The persistent server: workgroup w belongs to slot block_slot[w]. The
// slot's first workgroup claims requests from the slot's queue and hands
// each one to the others through the slot's mailbox; the slot then runs it
// exactly as a launch of that one request would (the same phases, barriers
// and reductions over the slot's workgroups), so results do not depend on
// what else is being served. Every wait is bounded by the host's stop flag
// or the launch's lifetime.
template <int kWaves>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(kWaves))) void server_kernel(ServeParams S) {
    __shared__ Lds lds;
    __shared__ Launch params;
    __shared__ int32_t state;
    __shared__ int64_t claimed;
    __shared__ long long start;          // thread 0: the launch's start on the wall clock
    __shared__ uint32_t generation;      // thread 0: mailbox generations seen
    if (threadIdx.x == 0) {
        start = wall_clock64();
        generation = 0;
    }
    const ServeSlot& slot = S.slots[S.block_slot[blockIdx.x]];
    const int g = static_cast<int>(blockIdx.x) - slot.first;
    const int G = slot.size;
    set_group(slot.first, slot.size, slot.barrier, slot.result, nullptr, nullptr, 0);
    int timeouts = 0;          // sticky: a broken slot stops waiting
    for (;;) {
        if (threadIdx.x == 0) {
            const ServeQueue& Q = S.queues[slot.queue];
            int64_t position = -1;
            if (g == 0) {
                position = claim_request(S, Q, start);
                if (G > 1) {
                    __hip_atomic_store(slot.mailbox + 32, position >= 0 ? static_cast<uint32_t>(position % Q.capacity) : 0u,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(slot.mailbox + 64, position >= 0 ? 0u : 1u, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(slot.mailbox, generation + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                }
            } else {
                bool alive = true;
                while (__hip_atomic_load(slot.mailbox, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == generation) {
                    __builtin_amdgcn_s_sleep(4);
                    if (wall_clock64() - start > static_cast<long long>(S.lifetime) + kMemberGrace) {
                        alive = false;
                        break;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                if (alive &&
                    __hip_atomic_load(slot.mailbox + 64, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
                    position = __hip_atomic_load(slot.mailbox + 32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            ++generation;
            claimed = position;
            if (position >= 0) {
                // the request, read word by word from pinned host memory
                ServeRequest r;
                const uint64_t* src = reinterpret_cast<const uint64_t*>(Q.ring + position % Q.capacity);
                uint64_t* dst = reinterpret_cast<uint64_t*>(&r);
                for (size_t k = 0; k < sizeof(ServeRequest) / 8; ++k) dst[k] = load_system(src + k);
                Launch p{};
                p.op = r.op;
                p.max_iterations = r.max_iterations;
                p.max_restarts = r.max_restarts;
                p.steps = 1;
                p.tolerance = r.tolerance;
                p.inner_relative = r.inner_relative;
                p.dt = r.dt;
                p.entry_conductance = r.entry_conductance;
                p.capacity_area = r.capacity_area;
                p.capacity_keep = r.capacity_keep;
                p.t_old = r.t_old;
                p.forcing = r.forcing;
                p.t_new = r.t_new;
                p.ticket = r.ticket;
                p.completion = r.completion;
                params = p;
                state = r.state;
            }
        }
        __syncthreads();
        if (claimed < 0) break;
        timeouts = serve_one<kWaves>(S.table + state, params, g, G, lds, timeouts);
        const Launch& P = params;
        if (g == 0 && threadIdx.x == 0) {
            const SolveResult& res = *slot.result;
            ServeCompletion c{};
            c.ticket = P.ticket;
            c.status = res.status;
            c.iterations = res.iterations;
            c.restarts = res.restarts;
            c.v_cycles = res.v_cycles;
            c.relative_residual = res.relative_residual;
            c.t_begin = res.t_begin;
            c.t_end = res.t_end;
            c.barrier_timeouts = timeouts;
            c.done = 0;
            ServeCompletion* out = S.completions + P.completion;
            const uint64_t* src = reinterpret_cast<const uint64_t*>(&c);
            uint64_t* dst = reinterpret_cast<uint64_t*>(out);
            constexpr size_t words = sizeof(ServeCompletion) / 8;
            for (size_t k = 0; k + 1 < words; ++k)
                __hip_atomic_store(dst + k, src[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            __hip_atomic_store(dst + words - 1, c.ticket + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
    }
    group_exit();
}

// The kernels of one occupancy: 1, 2 or 3 workgroups per CU.
const void* program_kernel(int waves) {
    if (waves == kMixedProgram) return reinterpret_cast<const void*>(&program_kernel_mixed);
    if (waves == 1) return reinterpret_cast<const void*>(&program_kernel_t<1>);
    if (waves == 2) return reinterpret_cast<const void*>(&program_kernel_t<2>);
    return reinterpret_cast<const void*>(&program_kernel_t<3>);
}

const void* server_kernel_of(int waves) {
    if (waves == 2) return reinterpret_cast<const void*>(&server_kernel<2>);
    return reinterpret_cast<const void*>(&server_kernel<3>);
}

}  // namespace

hipError_t launch_server(const ServeParams& params, int32_t blocks, hipStream_t stream, int waves) {
    ServeParams copy = params;
    void* args[] = {&copy};
    return hipLaunchCooperativeKernel(server_kernel_of(waves), dim3(blocks), dim3(kTile), args, 0, stream);
}

int server_blocks_per_cu(int waves) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, server_kernel_of(waves), kTile, 0) != hipSuccess)
        return 0;
    return blocks;
}

hipError_t launch_program(const Instance* instances, const int32_t* block_instance, const Launch* launches,
                          const Launch& single, int32_t blocks, size_t dynamic_lds, hipStream_t stream,
                          bool cooperative, int waves) {
    const void* kernel = program_kernel(waves);
    if (dynamic_lds > 0) {
        const hipError_t e =
            hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(dynamic_lds));
        if (e != hipSuccess) return e;
    }
    Launch copy = single;
    void* args[] = {const_cast<Instance**>(&instances), const_cast<int32_t**>(&block_instance),
                    const_cast<Launch**>(&launches), &copy};
    if (cooperative)
        return hipLaunchCooperativeKernel(kernel, dim3(blocks), dim3(kTile), args, static_cast<unsigned>(dynamic_lds),
                                          stream);
    return hipLaunchKernel(kernel, dim3(blocks), dim3(kTile), args, dynamic_lds, stream);
}

size_t program_static_lds() {
    hipFuncAttributes attributes{};
    if (hipFuncGetAttributes(&attributes, program_kernel(3)) != hipSuccess) return sizeof(Lds) + sizeof(Launch) + 256;
    return attributes.sharedSizeBytes;
}

int program_blocks_per_cu(size_t dynamic_lds, int waves) {
    const void* kernel = program_kernel(waves);
    int blocks = 0;
    if (dynamic_lds > 0 &&
        hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(dynamic_lds)) !=
            hipSuccess)
        return 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, kTile, dynamic_lds) == hipSuccess ? blocks : 0;
}

}  // namespace psim::dev
