// IncompleteLU with the reference's signatures (sim/core/math/
// incomplete_lu.hpp:14-43): ILU(0), or block ILU(0) on fixed row blocks,
// factored and solved on the device through libpsim (bit-identical to the
// reference's factors and solves).
#pragma once

#include "device.hpp"

#include <cstddef>
#include <cstdint>
#include <memory>
#include <span>
#include <stdexcept>
#include <vector>

namespace planetsim {

class IncompleteLU {
  public:
    IncompleteLU(std::span<const std::size_t> offset, std::span<const std::size_t> neighbour,
                 std::span<const double> diagonal, std::span<const double> off, double diagonal_scale = 1.0)
        : IncompleteLU(offset, neighbour, diagonal, off, std::vector<std::size_t>{0U, diagonal.size()}, diagonal_scale,
                       1U) {}

    IncompleteLU(std::span<const std::size_t> offset, std::span<const std::size_t> neighbour,
                 std::span<const double> diagonal, std::span<const double> off,
                 std::span<const std::size_t> block_start, double diagonal_scale, std::size_t worker_count) {
        if (offset.empty() || diagonal.size() + 1U != offset.size() || off.size() != neighbour.size() ||
            offset.back() != neighbour.size() || block_start.size() < 2U || block_start.front() != 0U ||
            block_start.back() != diagonal.size())
            throw std::invalid_argument("incomplete LU: the pattern does not match");
        static_cast<void>(worker_count);   // results do not depend on it
        static_assert(sizeof(std::size_t) == sizeof(std::uint64_t));
        psim_ilu* handle = nullptr;
        gpu::check(psim_ilu_create(gpu::context(), diagonal.size(), reinterpret_cast<const std::uint64_t*>(offset.data()),
                                   reinterpret_cast<const std::uint64_t*>(neighbour.data()),
                                   reinterpret_cast<const std::uint64_t*>(block_start.data()), block_start.size() - 1U,
                                   &handle));
        handle_.reset(handle);
        gpu::check(psim_ilu_factor(handle, diagonal.data(), off.data(), diagonal_scale, PSIM_HOST));
        rows_ = diagonal.size();
    }

    // x <- (LU)^-1 x.
    void solve(std::vector<double>& x) const {
        if (x.size() != rows_) throw std::invalid_argument("incomplete LU: the vector does not match");
        gpu::check(psim_ilu_solve(handle_.get(), x.data(), PSIM_HOST));
    }

  private:
    struct Free {
        void operator()(psim_ilu* f) const noexcept { psim_ilu_destroy(f); }
    };
    std::unique_ptr<psim_ilu, Free> handle_;
    std::size_t rows_ = 0;
};

}  // namespace planetsim
