// Process-wide device context and status-to-exception mapping of the
// drop-in C++ layer over libpsim's C ABI (include/psim.h).
#pragma once

#include "../psim.h"

#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

namespace planetsim::gpu {

// Maps a libpsim status to the exception type the reference throws for the
// same condition (std::invalid_argument for argument mismatches,
// std::logic_error for invariants, std::runtime_error otherwise).
inline void check(int status) {
    if (status == PSIM_OK) return;
    const std::string message = psim_last_error();
    switch (status) {
        case PSIM_INVALID_ARGUMENT: throw std::invalid_argument(message);
        case PSIM_LOGIC_ERROR: throw std::logic_error(message);
        case PSIM_OUT_OF_MEMORY: throw std::bad_alloc();
        default: throw std::runtime_error(message);
    }
}

// The library's context on device `PLANETSIM_DEVICE` (default 0), created
// on first use and kept for the process. Throws std::runtime_error when no
// GPU is visible: there is no CPU fallback.
inline psim_ctx* context() {
    static std::once_flag once;
    static psim_ctx* ctx = nullptr;
    std::call_once(once, [] {
        int device = 0;
        if (const char* env = std::getenv("PLANETSIM_DEVICE")) device = std::atoi(env);
        check(psim_ctx_create(device, &ctx));
    });
    return ctx;
}

// Device memory owned by the drop-in layer (staging of host fields).
class DeviceBuffer {
  public:
    explicit DeviceBuffer(std::size_t bytes) { check(psim_device_alloc(context(), bytes, &ptr_)); }
    DeviceBuffer(DeviceBuffer&& other) noexcept : ptr_(other.ptr_) { other.ptr_ = nullptr; }
    DeviceBuffer& operator=(DeviceBuffer&& other) noexcept {
        if (this != &other) {
            release();
            ptr_ = other.ptr_;
            other.ptr_ = nullptr;
        }
        return *this;
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void* get() const noexcept { return ptr_; }

  private:
    void release() noexcept {
        if (ptr_ != nullptr) (void)psim_device_free(context(), ptr_);
        ptr_ = nullptr;
    }
    void* ptr_ = nullptr;
};

// worker_count is accepted for signature compatibility; device results do
// not depend on it, but 0 is still rejected as the reference rejects it
// (deterministic_executor.hpp:17-19).
inline void check_workers(std::size_t worker_count) {
    if (worker_count == 0U) throw std::invalid_argument("worker count must be positive");
}

}  // namespace planetsim::gpu
