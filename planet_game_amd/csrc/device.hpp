// Device-side data layout shared by the HIP sources of libpsim.
//
// HBM layout, per mesh level (psim_graph, shared by all systems on it):
//   level j of the multigrid hierarchy: int32 CSR (row, col) of the
//   off-diagonals, aggregate map, aggregate-major row order with its
//   restriction tiles, and the Galerkin gather lists of level j+1.
// Per system (psim_system, one heat-transport instance):
//   FP64 diagonal / inverse diagonal / off-diagonals of every level, FP32
//   copies for the mixed-precision form, the coarsest level's dense inverse
//   (FP64), the solver vectors, per-tile reduction partials, a barrier word
//   pair and a result record.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <vector>

namespace psim::dev {

constexpr int kTile = 256;        // rows per tile = threads per workgroup = the reference's block
constexpr int kMaxLevels = 8;
constexpr int kPartials = 4;      // reduction slots per tile
constexpr int kMaxCoarse = 512;   // dense coarsest level (the hierarchy stops at <= 400 rows)
// Coarsest levels up to this size are inverted by the block sweep on the
// matrix cores (16 pivots per step, working matrix in sweep_words(n) words
// of I.tri: the lower triangle of 16 x 16 blocks).
constexpr int kBlockedCoarse = 168;
constexpr int kSweepBlock = 16;      // pivots per block step of the coarsest inverse
// Words of I.tri the block sweep uses for a coarsest level of n <= kBlockedCoarse rows.
constexpr int sweep_words(int n) {
    return (n + kSweepBlock - 1) / kSweepBlock * ((n + kSweepBlock - 1) / kSweepBlock + 1) / 2 * kSweepBlock *
           kSweepBlock;
}
constexpr int32_t kNotConverged = 5;   // PSIM_NOT_CONVERGED

enum Op : int32_t {
    kOpSetMatrix = 1,     // Galerkin values, smoother weights, coarse inverse
    kOpHeatAssemble = 2,  // level-0 values from entry conductances, C A and dt (with kOpSetMatrix)
    kOpSolve = 4,         // BiCGSTAB with true-residual restarts on b
    kOpHeatRhs = 8,       // b = C A T_old + dt A S (with kOpSolve); T_new = x
    kOpPrecondition = 16, // out = V-cycle(in)
    kOpApply = 32,        // out = A in (FP64)
    kOpCsrValues = 64,    // level-0 off-diagonals given in CSR order in `in` (with kOpSetMatrix)
    kOpChain = 128,       // `steps` heat steps, each T feeding the next (with kOpSolve | kOpHeatRhs)
    kOpPcg = 256          // PCG on the FP64 matrix: b in `in`, x to `out` (plain systems)
};

// Structure of one multigrid level (per graph). Off-diagonals are stored
// slot-major ELL: entry s of row i at s * n + i, in the row's CSR order,
// padding marked by column -1.
struct LevelView {
    int32_t n = 0, nnz = 0, width = 0, n_next = 0, down_tiles = 0;
    const int32_t* row = nullptr;         // CSR offsets, n + 1 (CSR-ordered inputs)
    const int32_t* col = nullptr;         // ELL columns, width * n
    const int32_t* agg = nullptr;         // n (not on the coarsest level)
    const int32_t* agg_col = nullptr;     // width * n: aggregate of each slot's column (-1 padding)
    const int32_t* agg_start = nullptr;   // n_next + 1
    const int32_t* agg_rows = nullptr;    // n, aggregate-major
    const int32_t* tile_start = nullptr;  // down_tiles + 1 (coarse-row ranges)
    const int32_t* gd_start = nullptr;    // Galerkin lists of level j+1's diagonal
    const int32_t* gd = nullptr;          //   (ELL positions, or -(row+1) for a fine diagonal)
    const int32_t* go_start = nullptr;    // ... and of level j+1's ELL positions
    const int32_t* go = nullptr;
    const int32_t* down_rows = nullptr;   // down_tiles * 256: each restriction tile's rows (-1 padding)
    const int32_t* need = nullptr;        // level above the coarsest: coarse rows each 256-row tile
    int32_t need_stride = 0;              //   reads, tile t's at need[t * need_stride ..) (-1 padding)
    // two levels above the coarsest (merged prolongation): per tile, the
    // next level's rows and the coarsest rows it reads (-1 padding), and each
    // row's / ELL slot's aggregate as a position in its tile's mid list
    const int32_t* mid = nullptr;
    const int32_t* mid_coarse = nullptr;
    const int32_t* agg_loc = nullptr;
    const int32_t* agg_col_loc = nullptr;
    int32_t mid_stride = 0, mid_coarse_stride = 0;
};

struct SolveResult {
    int32_t iterations;
    int32_t restarts;
    int32_t status;
    int32_t barrier_timeouts;
    double relative_residual;
    double b_norm;
    int32_t v_cycles;
    int32_t pad;
    uint64_t t_begin, t_end;     // device wall clock at the solve's start and end
};

// Everything one instance's program touches. Lives in device memory; every
// workgroup of the instance's group reads it (uniform, scalar loads).
struct Instance {
    int32_t precision = 0;       // 0 = FP64 inner vectors, 1 = FP32
    int32_t depth = 0;
    int32_t n = 0, nnz = 0, tiles = 0;   // fine level; tiles of 256 rows
    int32_t partial_slots = 0;   // tiles per partial buffer (>= tiles, >= level-0 restriction tiles)
    int32_t group_first = 0, group_size = 1;   // workgroups [first, first+size)
    LevelView lv[kMaxLevels];
    // values (per system), off-diagonals in ELL
    double* diag64[kMaxLevels] = {};
    double* dinv64[kMaxLevels] = {};
    double* off64[kMaxLevels] = {};
    float* diag32[kMaxLevels] = {};
    float* dinv32[kMaxLevels] = {};
    float* off32[kMaxLevels] = {};
    double* omega = nullptr;     // smoother weight per level
    double* cinv = nullptr;      // dense inverse of the coarsest level, column k contiguous
    // graph data
    const double* area = nullptr;
    const double* weight = nullptr;   // CSR order
    // vectors
    double* b = nullptr;
    double* x[2] = {};           // FP64 iterate, double-buffered for fused updates
    void* r = nullptr;
    void* r0 = nullptr;
    // b, r, r0 of the other parity: in the full-register-file program a
    // chained step's last phase assembles the next step's right-hand side
    // there (the solve's `rp` selects the pair in use)
    double* b_alt = nullptr;
    void* r_alt = nullptr;
    void* r0_alt = nullptr;
    void* p[2] = {};             // search direction, double-buffered
    void* v = nullptr;
    void* ph = nullptr;
    void* s = nullptr;
    void* sh = nullptr;
    void* t = nullptr;
    void* u = nullptr;           // A r0 of a pass's first iteration (symmetric A: (r0, A ph) = (A r0, ph))
    void* lvl_b[kMaxLevels] = {};   // V-cycle right-hand sides, levels >= 1
    void* lvl_e[kMaxLevels] = {};   // V-cycle corrections, levels >= 1
    void* lvl_x[kMaxLevels] = {};   // pre-smoothed iterate after two Jacobi sweeps, levels < depth - 1
    // BiCGSTAB solves: bit j set = level j pre-smooths with two Jacobi
    // sweeps (each its own phase) instead of one; 0 = the symmetric V(1,1)
    // that precondition and PCG always use; -1 = chosen per solve from the
    // levels' smoother weights
    int32_t pre2 = -1;
    double* tri = nullptr;          // coarsest working matrix: block sweep, or packed when not in LDS
    int32_t tri_lds = 0;            // 1: set_matrix launches carry it in dynamic LDS
    double* partial = nullptr;      // 2 x partial_slots x kPartials (double-buffered)
    uint32_t* barrier = nullptr;    // group barrier words (kBarrierWordCount, zeroed)
    SolveResult* result = nullptr;
    uint64_t* trace = nullptr;      // phase timestamps (wall clock), or null
    uint32_t* trace_count = nullptr;
    int32_t trace_capacity = 0;
};

// What one launch asks of one instance (passed by value for a single
// instance, as a device array for a batch). Plain data: value-initialise
// with Launch{}.
struct Launch {
    int32_t op;
    int32_t max_iterations, max_restarts;
    int32_t steps;               // kOpChain: number of chained steps
    double tolerance, inner_relative;
    double dt;
    const double* entry_conductance;   // K per CSR entry
    const double* capacity_area;       // C A per cell
    double* capacity_keep;             // kOpHeatAssemble: where C A is kept for later steps, or null
    const double* t_old;
    const double* forcing;             // kOpChain: steps x n, step-major
    double* t_new;
    const double* in;   // b for kOpSolve, input of precondition / apply, CSR values
    double* out;        // output of precondition / apply
    SolveResult* step_results;         // kOpChain: one record per step, or null
    const uint8_t* hold;               // kOpPcg: rows held at 0 (identity rows), or null
    const double* operator_diag;       // kOpPcg: the operator's diagonal where it differs from the V-cycle's, or null
    uint64_t ticket;                   // server: echoed in the completion record
    int64_t completion;                // server: completion record index
};

// Launch the persistent program. Workgroup w serves instance
// block_instance[w]; its parameters are launches[that] when `launches` is
// non-null, else `single`. `waves`: the instance of the program compiled
// for 1, 2 or 3 workgroups per CU (the register budget of 512, 256 or 168
// VGPRs): 1 for a single instance whose group fits one workgroup per CU, 2
// or 3 for batches.
// kMixedProgram as `waves`: the full-register-file program with the
// mixed-precision solve inlined, for one mixed system.
constexpr int kMixedProgram = -1;
hipError_t launch_program(const Instance* instances, const int32_t* block_instance, const Launch* launches,
                          const Launch& single, int32_t blocks, size_t dynamic_lds, hipStream_t stream,
                          bool cooperative, int waves = 3);

// Co-resident program workgroups per compute unit with `dynamic_lds`
// bytes of dynamic LDS (0 if the kernel cannot launch with that much).
int program_blocks_per_cu(size_t dynamic_lds, int waves = 3);
// The program kernel's static LDS bytes.
size_t program_static_lds();

constexpr int kBarrierWordCount = 96 + 64 * 8;   // generation, 2 roots, 8 leaves x 2 (128-byte apart)

// ---- server: a persistent launch whose workgroups are partitioned into
// slots; each slot takes requests from one queue, one at a time.
//
// Host -> device: request rings and their `tail` counters live in pinned,
// device-visible host memory; the host fills an entry, then publishes it by
// storing tail + 1 (release). Slot leaders claim entries by incrementing
// the queue's `head` in device memory. Device -> host: a completion record
// per request in pinned memory, whose `done` word (ticket + 1) is stored
// last (system-scope release).
constexpr int kMaxQueues = 8;

struct ServeRequest {
    int32_t state;              // index into the server's instance table
    int32_t op;                 // kOpSolve | kOpHeatRhs, plus kOpSetMatrix | kOpHeatAssemble to assemble
    int32_t max_iterations, max_restarts;
    double dt, tolerance, inner_relative;
    const double* entry_conductance;
    const double* capacity_area;
    double* capacity_keep;
    const double* t_old;
    const double* forcing;
    double* t_new;
    uint64_t ticket;
    int32_t completion;         // index of the completion record to write
    int32_t pad;
};
static_assert(sizeof(ServeRequest) % 8 == 0, "requests are read as 64-bit words");

struct ServeCompletion {
    uint64_t ticket;
    int32_t status, iterations, restarts, v_cycles;
    double relative_residual;
    uint64_t t_begin, t_end;    // device wall clock
    int32_t barrier_timeouts, pad;
    uint64_t done;              // ticket + 1 once the record is complete
};
static_assert(sizeof(ServeCompletion) % 8 == 0, "completions are written as 64-bit words");

struct ServeQueue {
    const ServeRequest* ring;   // pinned host memory, `capacity` entries
    const uint64_t* tail;       // pinned host memory: entries published
    uint64_t* head;             // device memory: entries claimed
    int32_t capacity;
    int32_t pad;
};

struct ServeSlot {
    int32_t first, size, queue, pad;
    uint32_t* barrier;          // the slot's group barrier words (zeroed)
    uint32_t* mailbox;          // [0] generation, [32] ring position, [64] command (0 run, 1 exit)
    SolveResult* result;        // the slot's result record
};

struct ServeParams {
    const Instance* table;      // instance descriptors (static while serving)
    const ServeSlot* slots;
    const int32_t* block_slot;  // workgroup -> slot
    ServeQueue queues[kMaxQueues];
    ServeCompletion* completions;   // pinned host memory, indexed by ticket % capacity
    int32_t completion_capacity;
    int32_t pad;
    const uint32_t* stop;       // pinned host memory: nonzero ends the launch once slots are idle
    uint64_t lifetime;          // wall-clock ticks after its start that an idle slot gives up
};

// Launch the server (cooperative, `blocks` = sum of slot sizes), compiled
// for `waves` (2 or 3) workgroups per CU.
hipError_t launch_server(const ServeParams& params, int32_t blocks, hipStream_t stream, int waves = 3);
// Co-resident server workgroups per compute unit.
int server_blocks_per_cu(int waves = 3);

void launch_diffusion(int n, const int32_t* row, const int32_t* col, const double* weight, const double* area,
                      double conductance, const double* t, double* out, hipStream_t stream);

}  // namespace psim::dev

// Host-side context shared by the device entry points.
struct psim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> events;
    int compute_units = 0;
    int blocks_per_cu = 0;      // co-resident program workgroups per CU
    size_t lds_per_block = 0;   // largest dynamic LDS a program workgroup may use
    size_t lds_per_cu = 0;
    size_t static_lds = 0;      // the program kernel's static LDS
    double wall_clock_hz = 1.0e8;
    bool cooperative = false;   // cooperative launch supported
};

namespace psim {
// Throws DeviceFailure (PSIM_DEVICE_ERROR) or DeviceOutOfMemory with the
// HIP error text when `status` is not hipSuccess.
void hip_check(hipError_t status, const char* what);

// Host-pointer calls stage through temporary device buffers, freed when the
// call returns (after its stream has been synchronized).
struct Staging {
    std::vector<void*> blocks;
    Staging() = default;
    Staging(const Staging&) = delete;
    Staging& operator=(const Staging&) = delete;
    ~Staging() {
        for (void* p : blocks) (void)hipFree(p);
    }
    void* in(const void* host, size_t bytes) {
        void* d = out(bytes);
        hip_check(hipMemcpy(d, host, bytes, hipMemcpyHostToDevice), "staging copy");
        return d;
    }
    void* out(size_t bytes) {
        void* d = nullptr;
        hip_check(hipMalloc(&d, bytes), "staging allocation");
        blocks.push_back(d);
        return d;
    }
};
}  // namespace psim
