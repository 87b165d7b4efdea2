// The allocation-free heat step: psim_heat_step's solve with the solver's
// workspace (every BiCGSTAB and V-cycle vector, the reduction partials, the
// group barrier words, the result record, the instance descriptor and the
// workgroup map) in one file-scope device array instead of the system's own
// buffers. The matrix and hierarchy still come from the caller's system,
// set up beforehand; the call itself makes no hipMalloc, hipFree or
// allocating library call, for callers that cannot allocate (code under
// stream capture, tight serving loops).
//
// The array is per device (each GPU holds its own copy until the process
// ends); its device address is taken once per device at context creation.
// Calls on one device share it: the last call's contents stay, and calls
// that overlap on one device (two streams, two threads) race.
#include "system.hpp"

#include <cstring>
#include <mutex>

using psim::guarded;
using psim::dev::Instance;
using psim::dev::kMaxLevels;
using psim::dev::kPartials;
using psim::dev::Launch;
using psim::dev::SolveResult;

namespace {

// 8 MiB: an L6 system (40,962 cells) in FP64 needs about 5.6 MB.
constexpr size_t kScratchBytes = size_t{8} << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Fixed offsets of the control words, the same for every system, so that
// the barrier protocol's state carries over from one call to the next.
constexpr size_t kBarrierOffset = 0;                // kBarrierWordCount uint32
constexpr size_t kResultOffset = 4096;              // SolveResult
constexpr size_t kDescriptorOffset = 8192;          // Instance
constexpr size_t kMapOffset = 16384;                // workgroup -> instance (all 0)
constexpr size_t kMaxMap = 4096;                    // workgroups
constexpr size_t kWorkOffset = kMapOffset + kMaxMap * sizeof(int32_t);
static_assert(sizeof(Instance) <= kMapOffset - kDescriptorOffset, "descriptor slot");
static_assert(psim::dev::kBarrierWordCount * sizeof(uint32_t) <= kResultOffset, "barrier slot");

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    unsigned char* base = nullptr;   // the array's device address (from hipGetSymbolAddress)
    Instance uploaded{};             // the descriptor now in the array
    bool have_descriptor = false;
};

std::mutex g_mutex;
DeviceScratch g_devices[kMaxDevices];

__global__ __launch_bounds__(256) void zero_kernel(uint4* p, size_t words) {
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t w = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; w < words; w += stride)
        p[w] = uint4{0u, 0u, 0u, 0u};
}

void zero(void* p, size_t bytes, hipStream_t stream) {
    const size_t words = bytes / sizeof(uint4);
    const unsigned blocks = static_cast<unsigned>(std::min<size_t>((words + 255) / 256, 1024));
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, stream, static_cast<uint4*>(p), words);
    psim::hip_check(hipGetLastError(), "scratch clear");
}

// Carve the workspace of system `s` out of the array at `base`: the system's
// descriptor with every workspace pointer moved into the array. Returns the
// bytes used (beyond kScratchBytes: the system does not fit).
size_t carve(const psim_system* s, unsigned char* base, Instance& I) {
    I = s->host;
    size_t at = kWorkOffset;
    const auto take = [&](size_t bytes) {
        unsigned char* p = base + at;
        at += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        return p;
    };
    const size_t n = static_cast<size_t>(I.n);
    const size_t vt = I.precision == 1 ? 4 : 8;
    I.barrier = reinterpret_cast<uint32_t*>(base + kBarrierOffset);
    I.result = reinterpret_cast<SolveResult*>(base + kResultOffset);
    I.trace = nullptr;
    I.trace_count = nullptr;
    I.trace_capacity = 0;
    I.partial = reinterpret_cast<double*>(take(2 * static_cast<size_t>(I.partial_slots) * kPartials * 8));
    I.b = reinterpret_cast<double*>(take(n * 8));
    I.b_alt = reinterpret_cast<double*>(take(n * 8));
    I.x[0] = reinterpret_cast<double*>(take(n * 8));
    I.x[1] = reinterpret_cast<double*>(take(n * 8));
    void** vecs[12] = {&I.r, &I.r0, &I.r_alt, &I.r0_alt, &I.p[0], &I.p[1], &I.v, &I.ph, &I.s, &I.sh, &I.t, &I.u};
    for (void** v : vecs) *v = take(n * 8);   // FP64 in both forms
    for (int j = 0; j < I.depth && j < kMaxLevels; ++j) {
        const size_t nj = static_cast<size_t>(I.lv[j].n);
        I.lvl_b[j] = j > 0 ? take(nj * vt) : nullptr;
        I.lvl_e[j] = j > 0 ? take(nj * vt) : nullptr;
        I.lvl_x[j] = j + 1 < I.depth ? take(nj * vt) : nullptr;
    }
    return at;
}

}  // namespace

namespace psim::noalloc {

// Once per device, with that device current: the array's device address,
// and a clear of the whole array (barrier words and result record start at
// zero). Called by psim_ctx_create.
void init_device(int device, hipStream_t stream) {
    if (device < 0 || device >= kMaxDevices) return;   // the entry point then reports it
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceScratch& d = g_devices[device];
    if (d.base != nullptr) return;
    void* ptr = nullptr;
    psim::hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "scratch address");
    zero(ptr, kScratchBytes, stream);
    psim::hip_check(hipStreamSynchronize(stream), "scratch clear");
    d.base = static_cast<unsigned char*>(ptr);
}

}  // namespace psim::noalloc

extern "C" {

int psim_heat_step_noalloc(psim_system* s, const double* t_old, const double* forcing, double dt, double tolerance,
                           int32_t max_iterations, double* t_new, psim_solve_info* info) {
    return guarded([&] {
        psim::sys::require_matrix(s);
        if (!s->heat_matrix) throw std::logic_error("psim_heat_set_matrix has not been called");
        if (t_old == nullptr || forcing == nullptr || t_new == nullptr) throw std::invalid_argument("null argument");
        if (!(dt > 0.0) || !std::isfinite(dt)) throw std::invalid_argument("time step must be finite and positive");
        if (!(tolerance >= 0.0) || max_iterations < 0) throw std::invalid_argument("invalid solver settings");
        psim_ctx* ctx = s->graph->ctx;
        if (ctx->device < 0 || ctx->device >= kMaxDevices) throw std::invalid_argument("device index out of range");
        if (static_cast<size_t>(s->workgroups) > kMaxMap)
            throw std::invalid_argument("allocation-free heat step: more than 4096 workgroups");
        hipStream_t stream = ctx->stream;
        std::lock_guard<std::mutex> lock(g_mutex);
        DeviceScratch& d = g_devices[ctx->device];
        if (d.base == nullptr) throw std::logic_error("allocation-free workspace not initialised on this device");
        Instance I{};
        const size_t need = carve(s, d.base, I);
        if (need > kScratchBytes)
            throw std::invalid_argument("allocation-free heat step: the system needs " + std::to_string(need) +
                                        " bytes of solver workspace, above the " + std::to_string(kScratchBytes) +
                                        "-byte maximum (FP64 systems up to 40,962 cells fit)");
        // the descriptor goes up only when it differs from the one in place
        // (pageable source: the copy completes before hipMemcpy returns)
        if (!d.have_descriptor || std::memcmp(&d.uploaded, &I, sizeof(Instance)) != 0) {
            psim::hip_check(hipMemcpyAsync(d.base + kDescriptorOffset, &I, sizeof(Instance), hipMemcpyHostToDevice,
                                           stream),
                            "descriptor upload");
            psim::hip_check(hipStreamSynchronize(stream), "descriptor upload");
            d.uploaded = I;
            d.have_descriptor = true;
        }
        Launch p = psim::sys::solve_params(s, psim::dev::kOpSolve | psim::dev::kOpHeatRhs, tolerance, max_iterations);
        p.dt = dt;
        p.capacity_area = s->capacity_area;
        p.t_old = t_old;
        p.forcing = forcing;
        p.t_new = t_new;
        const bool cooperative = s->workgroups > 1;
        if (cooperative && !ctx->cooperative) throw psim::DeviceFailure("cooperative launch unsupported on this device");
        const bool wide = cooperative && s->workgroups <= ctx->compute_units;
        psim::hip_check(psim::dev::launch_program(reinterpret_cast<const Instance*>(d.base + kDescriptorOffset),
                                                  reinterpret_cast<const int32_t*>(d.base + kMapOffset), nullptr, p,
                                                  s->workgroups, 0, stream, cooperative, wide ? 1 : 3),
                        "solver launch");
        if (info != nullptr) {
            SolveResult r{};
            psim::hip_check(hipMemcpyAsync(&r, d.base + kResultOffset, sizeof r, hipMemcpyDeviceToHost, stream),
                            "result copy");
            psim::hip_check(hipStreamSynchronize(stream), "solve");
            if (r.barrier_timeouts != 0) {
                // the barrier words and the result record start again from zero
                zero(d.base, kDescriptorOffset, stream);
                psim::hip_check(hipStreamSynchronize(stream), "barrier reset");
                throw psim::DeviceFailure("solver group barrier timed out");
            }
            psim::sys::fill_info(r, info, ctx->wall_clock_hz);
        }
    });
}

}  // extern "C"
