// The cell-to-group maps of the agglomerated transport on the device
// (sim/planet/surface/surface_energy.cpp:1226-1247,1285-1300,1306-1338):
// the area-weighted mean a group's response is restricted by, the highest
// floor of a group, and the broadcast of group values back to cells.
//
// Layout: the members of each group in cell order (CSR by group: offset,
// member, and the member's area next to it, so a group's terms are read
// contiguously). One thread owns one group and folds its members in that
// order, so a mean is the reference's sum bit for bit (no FMA contraction:
// the library is built with -ffp-contract=off). The sweeps are bound by
// HBM gathers of the cell field: per group of m cells and per field,
// 12 m + 8 bytes of map and areas plus 8 m gathered and 8 written.
#include "system.hpp"

#include <cmath>
#include <limits>
#include <memory>

struct psim_groups {
    psim_ctx* ctx = nullptr;
    int32_t cells = 0, groups = 0;
    psim::Arena arena;
    int32_t* offset = nullptr;       // groups + 1
    int32_t* member = nullptr;       // cells, grouped, cell order within a group
    double* member_area = nullptr;   // cells, area of member[k]
    double* group_area = nullptr;    // groups
    int32_t* group_of = nullptr;     // cells
    ~psim_groups() { arena.release(); }
};

namespace psim::dev {
namespace {

constexpr int kGroupThreads = 256;

__global__ __launch_bounds__(kGroupThreads) void group_mean_kernel(int groups, int cells, const int32_t* offset,
                                                                   const int32_t* member, const double* member_area,
                                                                   const double* group_area, const double* in,
                                                                   double* out) {
    const int g = static_cast<int>(blockIdx.x) * kGroupThreads + static_cast<int>(threadIdx.x);
    if (g >= groups) return;
    const size_t field = blockIdx.y;
    in += field * static_cast<size_t>(cells);
    double sum = 0.0;
    for (int k = offset[g]; k < offset[g + 1]; ++k) sum += member_area[k] * in[member[k]];
    out[field * static_cast<size_t>(groups) + g] = sum / group_area[g];
}

__global__ __launch_bounds__(kGroupThreads) void group_max_kernel(int groups, int cells, const int32_t* offset,
                                                                  const int32_t* member, const double* in,
                                                                  double* out) {
    const int g = static_cast<int>(blockIdx.x) * kGroupThreads + static_cast<int>(threadIdx.x);
    if (g >= groups) return;
    const size_t field = blockIdx.y;
    in += field * static_cast<size_t>(cells);
    double high = -std::numeric_limits<double>::infinity();
    for (int k = offset[g]; k < offset[g + 1]; ++k) {
        const double v = in[member[k]];
        high = high < v ? v : high;
    }
    out[field * static_cast<size_t>(groups) + g] = high;
}

__global__ __launch_bounds__(kGroupThreads) void group_broadcast_kernel(int groups, int cells,
                                                                        const int32_t* group_of, const double* in,
                                                                        double* out) {
    const int c = static_cast<int>(blockIdx.x) * kGroupThreads + static_cast<int>(threadIdx.x);
    if (c >= cells) return;
    const size_t field = blockIdx.y;
    out[field * static_cast<size_t>(cells) + c] = in[field * static_cast<size_t>(groups) + group_of[c]];
}

dim3 grid_over(int count, uint32_t fields) {
    return dim3(static_cast<unsigned>((count + kGroupThreads - 1) / kGroupThreads), fields);
}

}  // namespace
}  // namespace psim::dev

namespace {

void check_fields(const psim_groups* g, uint32_t fields, const void* in, const void* out, int where) {
    if (g == nullptr || in == nullptr || out == nullptr) throw std::invalid_argument("null argument");
    if (fields == 0 || fields > 65535) throw std::invalid_argument("field count out of range");
    psim::check_where(where);
}

// Runs `launch(in, out)` on device pointers, staging host arrays.
template <class Launch>
void run(psim_groups* g, int where, const void* in, size_t in_bytes, void* out, size_t out_bytes, const char* what,
         const Launch& launch) {
    psim::hip_check(hipSetDevice(g->ctx->device), "hipSetDevice");
    hipStream_t stream = g->ctx->stream;
    psim::Staging st;
    const void* d_in = where == PSIM_HOST ? st.in(in, in_bytes) : in;
    void* d_out = where == PSIM_HOST ? st.out(out_bytes) : out;
    launch(static_cast<const double*>(d_in), static_cast<double*>(d_out), stream);
    psim::hip_check(hipGetLastError(), what);
    if (where == PSIM_HOST) {
        psim::hip_check(hipMemcpyAsync(out, d_out, out_bytes, hipMemcpyDeviceToHost, stream), what);
        psim::hip_check(hipStreamSynchronize(stream), what);
    }
}

}  // namespace

extern "C" {

int psim_groups_create(psim_ctx* ctx, uint64_t cells, uint64_t groups, const uint32_t* group_of_cell,
                       const double* cell_area, const double* group_area, psim_groups** out) {
    return psim::guarded([&] {
        if (ctx == nullptr || group_of_cell == nullptr || cell_area == nullptr || group_area == nullptr ||
            out == nullptr)
            throw std::invalid_argument("null argument");
        if (cells == 0 || groups == 0 || cells >= (1ULL << 31) || groups > cells)
            throw std::invalid_argument("cell or group count out of range");
        std::vector<int32_t> offset(groups + 1, 0), member(cells), group_of(cells);
        std::vector<double> member_area(cells);
        for (uint64_t c = 0; c < cells; ++c) {
            if (group_of_cell[c] >= groups) throw std::invalid_argument("group index out of range");
            ++offset[group_of_cell[c] + 1];
            group_of[c] = static_cast<int32_t>(group_of_cell[c]);
        }
        for (uint64_t q = 0; q < groups; ++q) {
            if (offset[q + 1] == 0) throw std::logic_error("a group has no cell");
            offset[q + 1] += offset[q];
        }
        std::vector<int32_t> next(offset.begin(), offset.end() - 1);
        for (uint64_t c = 0; c < cells; ++c) {
            const int32_t k = next[group_of_cell[c]]++;
            member[k] = static_cast<int32_t>(c);
            member_area[k] = cell_area[c];
        }
        psim::hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        auto g = std::make_unique<psim_groups>();
        g->ctx = ctx;
        g->cells = static_cast<int32_t>(cells);
        g->groups = static_cast<int32_t>(groups);
        g->arena.plan(&g->offset, groups + 1);
        g->arena.plan(&g->member, cells);
        g->arena.plan(&g->member_area, cells);
        g->arena.plan(&g->group_area, groups);
        g->arena.plan(&g->group_of, cells);
        g->arena.allocate();
        const auto put = [](void* dst, const void* src, size_t bytes) {
            psim::hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "group map upload");
        };
        put(g->offset, offset.data(), offset.size() * 4);
        put(g->member, member.data(), cells * 4);
        put(g->member_area, member_area.data(), cells * 8);
        put(g->group_area, group_area, groups * 8);
        put(g->group_of, group_of.data(), cells * 4);
        *out = g.release();
    });
}

void psim_groups_destroy(psim_groups* g) { delete g; }

int psim_groups_mean(psim_groups* g, uint32_t fields, const double* cell_values, double* group_values, int where) {
    return psim::guarded([&] {
        check_fields(g, fields, cell_values, group_values, where);
        run(g, where, cell_values, size_t{8} * fields * g->cells, group_values, size_t{8} * fields * g->groups,
            "group mean", [&](const double* in, double* out, hipStream_t stream) {
                hipLaunchKernelGGL(psim::dev::group_mean_kernel, psim::dev::grid_over(g->groups, fields),
                                   dim3(psim::dev::kGroupThreads), 0, stream, g->groups, g->cells, g->offset, g->member,
                                   g->member_area, g->group_area, in, out);
            });
    });
}

int psim_groups_max(psim_groups* g, uint32_t fields, const double* cell_values, double* group_values, int where) {
    return psim::guarded([&] {
        check_fields(g, fields, cell_values, group_values, where);
        run(g, where, cell_values, size_t{8} * fields * g->cells, group_values, size_t{8} * fields * g->groups,
            "group max", [&](const double* in, double* out, hipStream_t stream) {
                hipLaunchKernelGGL(psim::dev::group_max_kernel, psim::dev::grid_over(g->groups, fields),
                                   dim3(psim::dev::kGroupThreads), 0, stream, g->groups, g->cells, g->offset, g->member,
                                   in, out);
            });
    });
}

int psim_groups_broadcast(psim_groups* g, uint32_t fields, const double* group_values, double* cell_values,
                          int where) {
    return psim::guarded([&] {
        check_fields(g, fields, group_values, cell_values, where);
        run(g, where, group_values, size_t{8} * fields * g->groups, cell_values, size_t{8} * fields * g->cells,
            "group broadcast", [&](const double* in, double* out, hipStream_t stream) {
                hipLaunchKernelGGL(psim::dev::group_broadcast_kernel, psim::dev::grid_over(g->cells, fields),
                                   dim3(psim::dev::kGroupThreads), 0, stream, g->groups, g->cells, g->group_of, in,
                                   out);
            });
    });
}

}  // extern "C"
