// Host-side state of the device entry points: graphs, systems and the
// helpers that launch the solver program for them (shared by the single-
// system C ABI and the batch engine).
#pragma once

#include "device.hpp"
#include "internal.hpp"

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

namespace psim {

// One device allocation carved into 256-byte aligned arrays. Sizes are
// recorded first (plan), then the block is allocated and the pointers set.
class Arena {
  public:
    template <class T> void plan(T** slot, size_t count) {
        slots_.push_back({reinterpret_cast<void**>(slot), bytes_});
        bytes_ += (std::max<size_t>(count, 1) * sizeof(T) + 255) / 256 * 256;
    }
    void allocate() {
        hip_check(hipMalloc(&block_, bytes_), "device allocation");
        for (auto& s : slots_) *s.slot = static_cast<char*>(block_) + s.offset;
    }
    void release() {
        if (block_ != nullptr) (void)hipFree(block_);
        block_ = nullptr;
    }
    size_t bytes() const { return bytes_; }
    void* block() const { return block_; }

  private:
    struct Slot {
        void** slot;
        size_t offset;
    };
    std::vector<Slot> slots_;
    size_t bytes_ = 0;
    void* block_ = nullptr;
};

// PSIM_HOST or PSIM_DEVICE, else invalid_argument.
void check_where(int where);

}  // namespace psim

struct psim_graph {
    psim_ctx* ctx = nullptr;
    psim::HostGraph graph;         // the caller's graph (area, CSR, weights) on the host
    psim::HostHierarchy host;
    int32_t n = 0, nnz = 0, depth = 0;
    psim::Arena arena;
    double* area = nullptr;
    double* weight = nullptr;
    int32_t* csr_col = nullptr;   // level-0 CSR columns (the diffusion sweep reads CSR order)
    psim::dev::LevelView views[psim::dev::kMaxLevels];
    // the Newton-system solver of psim_implicit_transport and its device
    // vectors, made on first use; calls on the graph hold newton_mutex
    psim_system* newton = nullptr;
    std::shared_ptr<void> transport_work;
    std::shared_ptr<void> tracer_work;   // stacked tracer systems by layer count
    std::shared_ptr<void> operator_work; // device vectors of psim_operator_solve
    std::shared_ptr<void> advection_work;   // psim_implicit_advection_diffusion's systems, under advection_mutex
    std::mutex advection_mutex;
    std::mutex newton_mutex;
    ~psim_graph() {
        if (newton != nullptr) psim_system_destroy(newton);
        arena.release();
    }
};

struct psim_system {
    psim_graph* graph = nullptr;
    int precision = PSIM_PLAIN_FP64;
    int32_t workgroups = 1;
    int32_t map_capacity = 1;
    size_t tri_bytes = 0;          // packed coarsest matrix
    psim::Arena arena;
    psim::dev::Instance host;      // the uploaded descriptor
    psim::dev::Instance* d_inst = nullptr;
    int32_t* d_map = nullptr;      // workgroup -> instance (all 0)
    psim::dev::SolveResult* d_result = nullptr;
    double* entry_conductance = nullptr;
    double* capacity_area = nullptr;
    double* stage_a = nullptr;     // host-call staging (n)
    double* stage_b = nullptr;
    double* stage_c = nullptr;
    double* csr_stage = nullptr;   // CSR-ordered off-diagonals for set_matrix
    bool matrix_set = false;
    bool heat_matrix = false;
    void* trace_block = nullptr;   // uint32 count + uint64 stamps
    ~psim_system() {
        arena.release();
        if (trace_block != nullptr) (void)hipFree(trace_block);
    }
};

namespace psim::sys {

// Launch the program for one system on its context's stream.
void launch(psim_system* s, const dev::Launch& p);
// Heat-assembly inputs: host arrays are staged into the system, device
// arrays are read in place (C A kept by the assembly).
void stage_heat_inputs(psim_system* s, const double* conductance_per_entry, const double* capacity_area, int where,
                       dev::Launch& p);
// The system's last result (synchronises; throws on a barrier timeout).
dev::SolveResult fetch_result(psim_system* s);
void fill_info(const dev::SolveResult& r, psim_solve_info* info, double clock_hz = 1.0e8);
// BiCGSTAB settings of a solve launch for the system's precision.
dev::Launch solve_params(const psim_system* s, int op, double tolerance, int32_t max_iterations);
void require_matrix(const psim_system* s);

}  // namespace psim::sys

namespace psim::noalloc {
// Device address and first clear of the allocation-free heat step's
// workspace array on `device` (current), once per device.
void init_device(int device, hipStream_t stream);
}  // namespace psim::noalloc
