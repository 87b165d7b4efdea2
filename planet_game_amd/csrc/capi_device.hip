// C ABI: device context, transport graphs, linear systems and the heat step.
#include "system.hpp"

#include <algorithm>
#include <cstring>
#include <memory>
#include <vector>

using psim::guarded;
using psim::dev::Instance;
using psim::dev::kMaxLevels;
using psim::dev::kPartials;
using psim::dev::kTile;
using psim::dev::Launch;
using psim::dev::LevelView;
using psim::dev::SolveResult;

namespace psim {

void hip_check(hipError_t status, const char* what) {
    if (status == hipSuccess) return;
    const std::string text = std::string(what) + ": " + hipGetErrorString(status);
    if (status == hipErrorOutOfMemory) throw DeviceOutOfMemory(text);
    throw DeviceFailure(text);
}

namespace {

void put(void* dst, const void* src, size_t bytes) {
    if (bytes != 0) hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "upload");
}

}  // namespace

void check_where(int where) {
    if (where != PSIM_HOST && where != PSIM_DEVICE) throw std::invalid_argument("unknown memory location");
}

}  // namespace psim

namespace {

// Plain streaming copy (16 B per lane, grid-stride): the HBM reference
// kernel whose bytes per launch are known exactly.
__global__ __launch_bounds__(256) void copy_words_kernel(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                         uint64_t words) {
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; w < words; w += stride)
        dst[w] = src[w];
}

}  // namespace

namespace psim::sys {

void launch(psim_system* s, const Launch& p) {
    psim_ctx* ctx = s->graph->ctx;
    const bool cooperative = s->workgroups > 1;
    if (cooperative && !ctx->cooperative) throw psim::DeviceFailure("cooperative launch unsupported on this device");
    const size_t lds = (p.op & psim::dev::kOpSetMatrix) && s->host.tri_lds ? s->tri_bytes : 0;
    // a group of more than one workgroup that fits one per CU runs the
    // full-register-file instance of the program
    const bool wide = cooperative && s->workgroups <= ctx->compute_units;
    // a mixed system's solve in the full-register-file program: the kernel
    // with that solve inlined
    const bool mixed_solve = wide && (p.op & psim::dev::kOpSolve) && s->host.precision == 1;
    psim::hip_check(psim::dev::launch_program(s->d_inst, s->d_map, nullptr, p, s->workgroups, lds, ctx->stream,
                                              cooperative, mixed_solve ? psim::dev::kMixedProgram : wide ? 1 : 3),
                    "solver launch");
}

// Heat-assembly inputs of a launch: host arrays are copied into the
// system's buffers first; device arrays are read in place by the assembly,
// which keeps C A in the system for the steps that follow.
void stage_heat_inputs(psim_system* s, const double* conductance_per_entry, const double* capacity_area, int where,
                       Launch& p) {
    if (where == PSIM_HOST) {
        hipStream_t stream = s->graph->ctx->stream;
        psim::hip_check(hipMemcpyAsync(s->entry_conductance, conductance_per_entry,
                                       static_cast<size_t>(s->graph->nnz) * 8, hipMemcpyHostToDevice, stream),
                        "copy");
        psim::hip_check(hipMemcpyAsync(s->capacity_area, capacity_area, static_cast<size_t>(s->graph->n) * 8,
                                       hipMemcpyHostToDevice, stream),
                        "copy");
        p.entry_conductance = s->entry_conductance;
        p.capacity_area = s->capacity_area;
        p.capacity_keep = nullptr;
    } else {
        p.entry_conductance = conductance_per_entry;
        p.capacity_area = capacity_area;
        p.capacity_keep = s->capacity_area;
    }
}

SolveResult fetch_result(psim_system* s) {
    SolveResult r{};
    psim_ctx* ctx = s->graph->ctx;
    psim::hip_check(hipMemcpyAsync(&r, s->d_result, sizeof r, hipMemcpyDeviceToHost, ctx->stream), "result copy");
    psim::hip_check(hipStreamSynchronize(ctx->stream), "solve");
    if (r.barrier_timeouts != 0) {
        // A barrier gave up waiting: the result is unusable and the words
        // must be cleared before the next launch.
        psim::hip_check(hipMemset(s->host.barrier, 0, psim::dev::kBarrierWordCount * sizeof(uint32_t)),
                        "barrier reset");
        psim::hip_check(hipMemset(s->d_result, 0, sizeof(SolveResult)), "result reset");
        throw psim::DeviceFailure("solver group barrier timed out");
    }
    return r;
}

void fill_info(const SolveResult& r, psim_solve_info* info, double clock_hz) {
    if (info == nullptr) return;
    info->device_ms = static_cast<double>(r.t_end - r.t_begin) / clock_hz * 1e3;
    info->iterations = r.iterations;
    info->restarts = r.restarts;
    info->status = r.status;
    info->v_cycles = r.v_cycles;
    info->relative_residual = r.relative_residual;
}

Launch solve_params(const psim_system* s, int op, double tolerance, int32_t max_iterations) {
    Launch p{};
    p.op = op;
    p.tolerance = tolerance;
    p.max_iterations = max_iterations;
    // both forms iterate FP64 vectors: no inner target below the tolerance
    p.max_restarts = 8;
    p.inner_relative = 0.0;
    return p;
}

void require_matrix(const psim_system* s) {
    if (s == nullptr) throw std::invalid_argument("null system");
    if (!s->matrix_set) throw std::logic_error("set_matrix has not been called");
}

}  // namespace psim::sys

using namespace psim::sys;

extern "C" {

int psim_ctx_create(int device, psim_ctx** out) {
    return guarded([&] {
        if (out == nullptr) throw std::invalid_argument("null output handle");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count == 0) throw psim::DeviceFailure("no GPU is visible");
        if (device < 0 || device >= count) throw std::invalid_argument("device index out of range");
        auto ctx = std::make_unique<psim_ctx>();
        ctx->device = device;
        psim::hip_check(hipSetDevice(device), "hipSetDevice");
        psim::hip_check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "stream create");
        hipDeviceProp_t prop{};
        psim::hip_check(hipGetDeviceProperties(&prop, device), "device properties");
        ctx->compute_units = prop.multiProcessorCount;
        ctx->cooperative = prop.cooperativeLaunch != 0;
        ctx->blocks_per_cu = psim::dev::program_blocks_per_cu(0);
        int clock_khz = 0;
        if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && clock_khz > 0)
            ctx->wall_clock_hz = 1.0e3 * clock_khz;
        ctx->static_lds = psim::dev::program_static_lds();
        ctx->lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
        ctx->lds_per_block =
            prop.sharedMemPerBlock > ctx->static_lds ? prop.sharedMemPerBlock - ctx->static_lds : 0;
        psim::noalloc::init_device(device, ctx->stream);
        *out = ctx.release();
    });
}

void psim_ctx_destroy(psim_ctx* ctx) {
    if (ctx == nullptr) return;
    for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
    if (ctx->stream != nullptr) {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
}

int psim_ctx_synchronize(psim_ctx* ctx) {
    return guarded([&] {
        if (ctx == nullptr) throw std::invalid_argument("null context");
        psim::hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
    });
}

int psim_ctx_info(const psim_ctx* ctx, psim_device_info* info) {
    return guarded([&] {
        if (ctx == nullptr || info == nullptr) throw std::invalid_argument("null argument");
        hipDeviceProp_t prop{};
        psim::hip_check(hipGetDeviceProperties(&prop, ctx->device), "device properties");
        info->compute_units = ctx->compute_units;
        info->program_blocks_per_cu = ctx->blocks_per_cu;
        info->cooperative = ctx->cooperative ? 1 : 0;
        info->reserved = 0;
        info->lds_per_block = static_cast<uint64_t>(prop.sharedMemPerBlock);
        info->lds_per_cu = static_cast<uint64_t>(prop.maxSharedMemoryPerMultiProcessor);
        info->hbm_bytes = static_cast<uint64_t>(prop.totalGlobalMem);
        info->wall_clock_hz = ctx->wall_clock_hz;
    });
}

int psim_ctx_events(psim_ctx* ctx, uint32_t count) {
    return guarded([&] {
        if (ctx == nullptr) throw std::invalid_argument("null context");
        if (count > (1U << 20)) throw std::invalid_argument("too many events");
        psim::hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
        for (hipEvent_t e : ctx->events) (void)hipEventDestroy(e);
        ctx->events.assign(count, nullptr);
        for (auto& e : ctx->events) psim::hip_check(hipEventCreate(&e), "event create");
    });
}

int psim_ctx_record(psim_ctx* ctx, uint32_t index) {
    return guarded([&] {
        if (ctx == nullptr || index >= ctx->events.size()) throw std::invalid_argument("no such event");
        psim::hip_check(hipEventRecord(ctx->events[index], ctx->stream), "event record");
    });
}

int psim_ctx_elapsed(psim_ctx* ctx, uint32_t from, uint32_t to, float* ms) {
    return guarded([&] {
        if (ctx == nullptr || ms == nullptr || from >= ctx->events.size() || to >= ctx->events.size())
            throw std::invalid_argument("no such event");
        psim::hip_check(hipEventSynchronize(ctx->events[to]), "event wait");
        psim::hip_check(hipEventElapsedTime(ms, ctx->events[from], ctx->events[to]), "event elapsed");
    });
}

int psim_device_alloc(psim_ctx* ctx, uint64_t bytes, void** out) {
    return guarded([&] {
        if (ctx == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        psim::hip_check(hipMalloc(out, std::max<uint64_t>(bytes, 1)), "device allocation");
    });
}

int psim_device_free(psim_ctx* ctx, void* ptr) {
    return guarded([&] {
        if (ctx == nullptr) throw std::invalid_argument("null context");
        psim::hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
        psim::hip_check(hipFree(ptr), "device free");
    });
}

int psim_copy_to_device(psim_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        if (ctx == nullptr) throw std::invalid_argument("null context");
        psim::hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream), "copy to device");
        psim::hip_check(hipStreamSynchronize(ctx->stream), "copy to device");
    });
}

int psim_copy_to_host(psim_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        if (ctx == nullptr) throw std::invalid_argument("null context");
        psim::hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream), "copy to host");
        psim::hip_check(hipStreamSynchronize(ctx->stream), "copy to host");
    });
}

int psim_device_copy(psim_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded([&] {
        if (ctx == nullptr || dst == nullptr || src == nullptr) throw std::invalid_argument("null argument");
        if (bytes % 16 != 0 || reinterpret_cast<uintptr_t>(dst) % 16 != 0 || reinterpret_cast<uintptr_t>(src) % 16 != 0)
            throw std::invalid_argument("device copy needs 16-byte aligned pointers and size");
        if (bytes == 0) return;
        const uint64_t words = bytes / 16;
        const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((words + 255) / 256, 8192));
        hipLaunchKernelGGL(copy_words_kernel, dim3(blocks), dim3(256), 0, ctx->stream,
                           static_cast<uint4*>(dst), static_cast<const uint4*>(src), words);
        psim::hip_check(hipGetLastError(), "device copy");
    });
}

int psim_graph_create(psim_ctx* ctx, uint64_t n, const double* area, const uint64_t* offset, const uint64_t* neighbour,
                      const double* weight, const double* weight_sum, psim_graph** out) {
    return guarded([&] {
        if (ctx == nullptr || out == nullptr || !area || !offset || !weight_sum)
            throw std::invalid_argument("null argument");
        if (n == 0 || n >= (1ULL << 30)) throw std::invalid_argument("graph size out of range");
        // a graph without entries (isolated nodes only) may pass no neighbour
        // or weight arrays
        const uint64_t nnz = psim::checked_graph_entries(n, offset);
        if (nnz > 0 && (!neighbour || !weight)) throw std::invalid_argument("null argument");
        psim::HostGraph g;
        g.n = static_cast<int32_t>(n);
        g.row.resize(n + 1);
        for (uint64_t i = 0; i <= n; ++i) g.row[i] = static_cast<int32_t>(offset[i]);
        g.col.resize(nnz);
        for (uint64_t k = 0; k < nnz; ++k) {
            if (neighbour[k] >= n) throw std::invalid_argument("graph neighbour out of range");
            g.col[k] = static_cast<int32_t>(neighbour[k]);
        }
        g.area.assign(area, area + n);
        g.weight.assign(weight, weight + nnz);
        g.weight_sum.assign(weight_sum, weight_sum + n);
        auto gr = std::make_unique<psim_graph>();
        gr->ctx = ctx;
        gr->host = psim::build_hierarchy(g);
        gr->graph = g;
        gr->n = g.n;
        gr->nnz = static_cast<int32_t>(nnz);
        gr->depth = static_cast<int32_t>(gr->host.levels.size());
        if (gr->depth > kMaxLevels) throw std::logic_error("hierarchy deeper than supported");
        const auto& lv = gr->host.levels;
        if (lv.back().n > psim::dev::kMaxCoarse) throw std::logic_error("coarsest level larger than supported");
        psim::hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        // plan
        auto& A = gr->arena;
        A.plan(&gr->area, n);
        A.plan(&gr->weight, nnz);
        A.plan(&gr->csr_col, nnz);
        int32_t* ptrs[kMaxLevels][17] = {};
        for (int j = 0; j < gr->depth; ++j) {
            const psim::HostLevel& L = lv[j];
            const bool last = j + 1 == gr->depth;
            A.plan(&ptrs[j][0], L.row.size());
            A.plan(&ptrs[j][1], L.col_ell.size());
            if (!last) {
                A.plan(&ptrs[j][2], L.aggregate.size());
                A.plan(&ptrs[j][3], L.agg_start.size());
                A.plan(&ptrs[j][4], L.agg_rows.size());
                A.plan(&ptrs[j][5], L.tile_start.size());
                A.plan(&ptrs[j][6], L.g_diag_start.size());
                A.plan(&ptrs[j][7], L.d_diag.size());
                A.plan(&ptrs[j][8], L.d_off_start.size());
                A.plan(&ptrs[j][9], L.d_off.size());
                A.plan(&ptrs[j][12], L.agg_ell.size());
                A.plan(&ptrs[j][11], L.down_rows.size());
            }
            if (!L.need_pad.empty()) A.plan(&ptrs[j][10], L.need_pad.size());
            if (!L.mid_pad.empty()) {
                A.plan(&ptrs[j][13], L.mid_pad.size());
                A.plan(&ptrs[j][14], L.mid_coarse_pad.size());
                A.plan(&ptrs[j][15], L.agg_loc.size());
                A.plan(&ptrs[j][16], L.agg_col_loc.size());
            }
        }
        A.allocate();
        psim::put(gr->area, area, n * 8);
        psim::put(gr->weight, weight, nnz * 8);
        psim::put(gr->csr_col, g.col.data(), nnz * 4);
        for (int j = 0; j < gr->depth; ++j) {
            const psim::HostLevel& L = lv[j];
            const bool last = j + 1 == gr->depth;
            LevelView& v = gr->views[j];
            v.n = L.n;
            v.nnz = L.nnz;
            v.width = L.width;
            psim::put(ptrs[j][0], L.row.data(), L.row.size() * 4);
            psim::put(ptrs[j][1], L.col_ell.data(), L.col_ell.size() * 4);
            v.row = ptrs[j][0];
            v.col = ptrs[j][1];
            if (!last) {
                const std::vector<int32_t>* src[8] = {&L.aggregate, &L.agg_start, &L.agg_rows, &L.tile_start,
                                                      &L.g_diag_start, &L.d_diag, &L.d_off_start, &L.d_off};
                for (int q = 0; q < 8; ++q) psim::put(ptrs[j][2 + q], src[q]->data(), src[q]->size() * 4);
                psim::put(ptrs[j][12], L.agg_ell.data(), L.agg_ell.size() * 4);
                psim::put(ptrs[j][11], L.down_rows.data(), L.down_rows.size() * 4);
                v.down_rows = ptrs[j][11];
                v.agg = ptrs[j][2];
                v.agg_col = ptrs[j][12];
                v.agg_start = ptrs[j][3];
                v.agg_rows = ptrs[j][4];
                v.tile_start = ptrs[j][5];
                v.down_tiles = static_cast<int32_t>(L.tile_start.size()) - 1;
                v.gd_start = ptrs[j][6];
                v.gd = ptrs[j][7];
                v.go_start = ptrs[j][8];
                v.go = ptrs[j][9];
                v.n_next = lv[j + 1].n;
            }
            if (!L.need_pad.empty()) {
                psim::put(ptrs[j][10], L.need_pad.data(), L.need_pad.size() * 4);
                v.need = ptrs[j][10];
                v.need_stride = L.need_stride;
            }
            if (!L.mid_pad.empty()) {
                psim::put(ptrs[j][13], L.mid_pad.data(), L.mid_pad.size() * 4);
                psim::put(ptrs[j][14], L.mid_coarse_pad.data(), L.mid_coarse_pad.size() * 4);
                psim::put(ptrs[j][15], L.agg_loc.data(), L.agg_loc.size() * 4);
                psim::put(ptrs[j][16], L.agg_col_loc.data(), L.agg_col_loc.size() * 4);
                v.mid = ptrs[j][13];
                v.mid_coarse = ptrs[j][14];
                v.agg_loc = ptrs[j][15];
                v.agg_col_loc = ptrs[j][16];
                v.mid_stride = L.mid_stride;
                v.mid_coarse_stride = L.mid_coarse_stride;
            }
        }
        *out = gr.release();
    });
}

void psim_graph_destroy(psim_graph* graph) { delete graph; }
uint64_t psim_graph_size(const psim_graph* g) { return g ? static_cast<uint64_t>(g->n) : 0; }
uint64_t psim_graph_entries(const psim_graph* g) { return g ? static_cast<uint64_t>(g->nnz) : 0; }
uint64_t psim_graph_depth(const psim_graph* g) { return g ? static_cast<uint64_t>(g->depth) : 0; }

int psim_diffusion_source(psim_graph* g, double conductance, const double* temperature, double* source, int where) {
    return guarded([&] {
        if (g == nullptr || temperature == nullptr || source == nullptr) throw std::invalid_argument("null argument");
        psim::check_where(where);
        hipStream_t stream = g->ctx->stream;
        const size_t bytes = static_cast<size_t>(g->n) * 8;
        double* t = const_cast<double*>(temperature);
        double* o = source;
        if (where == PSIM_HOST) {
            psim::hip_check(hipMalloc(&t, 2 * bytes), "staging allocation");
            o = t + g->n;
            psim::hip_check(hipMemcpy(t, temperature, bytes, hipMemcpyHostToDevice), "staging copy");
        }
        psim::dev::launch_diffusion(g->n, g->views[0].row, g->csr_col, g->weight, g->area, conductance, t, o, stream);
        if (where == PSIM_HOST) {
            const hipError_t e1 = hipMemcpyAsync(source, o, bytes, hipMemcpyDeviceToHost, stream);
            const hipError_t e2 = hipStreamSynchronize(stream);
            (void)hipFree(t);
            psim::hip_check(e1, "diffusion copy");
            psim::hip_check(e2, "diffusion_source");
        }
    });
}

int psim_system_create(psim_graph* g, int precision, psim_system** out) {
    return guarded([&] {
        if (g == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        if (precision != PSIM_PLAIN_FP64 && precision != PSIM_MIXED) throw std::invalid_argument("unknown precision");
        psim_ctx* ctx = g->ctx;
        psim::hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        auto s = std::make_unique<psim_system>();
        s->graph = g;
        s->precision = precision;
        const int32_t n = g->n, tiles = (n + kTile - 1) / kTile;
        const size_t nc_rows = static_cast<size_t>(g->views[g->depth - 1].n);
        s->tri_bytes = nc_rows * (nc_rows + 1) / 2 * sizeof(double);
        int per_cu = ctx->blocks_per_cu;
        if (nc_rows <= static_cast<size_t>(psim::dev::kBlockedCoarse)) {
            s->tri_bytes = 0;   // swept by blocks in global scratch (I.tri)
        } else if (s->tri_bytes <= ctx->lds_per_block) {
            // set_matrix launches carry the packed coarsest matrix in dynamic
            // LDS; co-residency then follows from the CU's LDS
            const size_t per_block = ctx->static_lds + s->tri_bytes;
            const int by_lds = static_cast<int>(ctx->lds_per_cu / per_block);
            if (by_lds > 0) {
                s->host.tri_lds = 1;
                per_cu = std::min(per_cu, by_lds);
            }
        }
        const int32_t cap = std::max(1, ctx->compute_units * per_cu);
        // one workgroup per tile of the widest level-0 sweep: the restriction
        // sweep packs whole aggregates into its tiles, so it has a few more
        // tiles than the natural 256-row split, and a workgroup with two of
        // them would hold up the phase
        const int32_t widest = g->depth > 1 ? std::max(tiles, g->views[0].down_tiles) : tiles;
        s->workgroups = std::min(widest, cap);
        Instance& I = s->host;
        I.precision = precision == PSIM_MIXED ? 1 : 0;
        I.depth = g->depth;
        I.n = n;
        I.nnz = g->nnz;
        I.tiles = tiles;
        I.group_first = 0;
        I.group_size = s->workgroups;
        for (int j = 0; j < g->depth; ++j) I.lv[j] = g->views[j];
        I.area = g->area;
        I.weight = g->weight;
        auto& A = s->arena;
        const bool mixed = precision == PSIM_MIXED;
        const size_t vt = mixed ? 4 : 8;
        for (int j = 0; j < g->depth; ++j) {
            const size_t nj = static_cast<size_t>(g->views[j].n);
            const size_t slots = nj * static_cast<size_t>(g->views[j].width);
            A.plan(&I.diag64[j], nj);
            A.plan(&I.dinv64[j], nj);
            A.plan(&I.off64[j], slots);
            if (mixed) {
                A.plan(&I.diag32[j], nj);
                A.plan(&I.dinv32[j], nj);
                A.plan(&I.off32[j], slots);
            }
            if (j > 0) {
                A.plan(reinterpret_cast<char**>(&I.lvl_b[j]), nj * vt);
                A.plan(reinterpret_cast<char**>(&I.lvl_e[j]), nj * vt);
            }
            if (j + 1 < g->depth) A.plan(reinterpret_cast<char**>(&I.lvl_x[j]), nj * vt);
        }
        const size_t nc = static_cast<size_t>(g->views[g->depth - 1].n);
        A.plan(&I.omega, kMaxLevels);
        A.plan(&I.cinv, nc * nc);
        A.plan(&I.b, n);
        A.plan(&I.b_alt, n);
        A.plan(&I.x[0], n);
        A.plan(&I.x[1], n);
        // the BiCGSTAB vectors are FP64 in both forms (the mixed form runs the
        // V-cycle's interior in FP32, on the level vectors above)
        void** vecs[12] = {&I.r, &I.r0, &I.r_alt, &I.r0_alt, &I.p[0], &I.p[1], &I.v, &I.ph, &I.s, &I.sh, &I.t, &I.u};
        for (void** v : vecs) A.plan(reinterpret_cast<char**>(v), static_cast<size_t>(n) * 8);
        I.partial_slots = std::max(tiles, g->depth > 1 ? g->views[0].down_tiles : 0);
        A.plan(&I.partial, 2 * static_cast<size_t>(I.partial_slots) * kPartials);
        A.plan(&s->csr_stage, static_cast<size_t>(g->nnz));
        A.plan(&I.barrier, psim::dev::kBarrierWordCount);
        if (!I.tri_lds && s->tri_bytes > 0)
            A.plan(&I.tri, nc_rows * (nc_rows + 1) / 2);
        else if (nc_rows <= static_cast<size_t>(psim::dev::kBlockedCoarse))
            A.plan(&I.tri, static_cast<size_t>(psim::dev::sweep_words(static_cast<int>(nc_rows))));
        A.plan(&I.result, 1);
        A.plan(&s->d_inst, 1);
        s->map_capacity = cap;
        A.plan(&s->d_map, static_cast<size_t>(cap));
        A.plan(&s->entry_conductance, static_cast<size_t>(g->nnz));
        A.plan(&s->capacity_area, n);
        A.plan(&s->stage_a, n);
        A.plan(&s->stage_b, n);
        A.plan(&s->stage_c, n);
        A.allocate();
        s->d_result = I.result;
        psim::hip_check(hipMemset(A.block(), 0, A.bytes()), "system clear");
        psim::put(s->d_inst, &I, sizeof(Instance));
        *out = s.release();
    });
}

void psim_system_destroy(psim_system* s) {
    if (s == nullptr) return;
    (void)hipStreamSynchronize(s->graph->ctx->stream);
    delete s;
}

int psim_system_set_workgroups(psim_system* s, int32_t workgroups) {
    return guarded([&] {
        if (s == nullptr) throw std::invalid_argument("null system");
        const psim_ctx* ctx = s->graph->ctx;
        const int32_t cap = s->map_capacity;
        if (workgroups < 1 || workgroups > cap) throw std::invalid_argument("workgroup count out of range");
        if (workgroups > 1 && !ctx->cooperative) throw std::invalid_argument("device cannot run a multi-workgroup group");
        psim::hip_check(hipStreamSynchronize(ctx->stream), "synchronize");
        // the workgroup map holds one entry per workgroup: grow it if needed
        if (workgroups > s->map_capacity) throw std::invalid_argument("workgroup count above the system's capacity");
        s->workgroups = workgroups;
        s->host.group_size = workgroups;
        psim::put(s->d_inst, &s->host, sizeof(Instance));
    });
}

int32_t psim_system_workgroups(const psim_system* s) { return s ? s->workgroups : 0; }

int psim_system_set_pre_sweeps(psim_system* s, int32_t level_mask) {
    return guarded([&] {
        if (s == nullptr) throw std::invalid_argument("null system");
        const int32_t smoothed = (1 << std::max(s->graph->depth - 1, 0)) - 1;
        if (level_mask < -1 || (level_mask >= 0 && (level_mask & ~smoothed) != 0))
            throw std::invalid_argument("pre-sweep mask names a level that is not smoothed");
        psim::hip_check(hipStreamSynchronize(s->graph->ctx->stream), "synchronize");
        s->host.pre2 = level_mask;
        psim::put(s->d_inst, &s->host, sizeof(Instance));
    });
}

int32_t psim_system_pre_sweeps(const psim_system* s) { return s ? s->host.pre2 : 0; }

int psim_system_trace(psim_system* s, int32_t capacity) {
    return guarded([&] {
        if (s == nullptr || capacity < 0 || capacity > (1 << 20)) throw std::invalid_argument("invalid trace capacity");
        psim::hip_check(hipStreamSynchronize(s->graph->ctx->stream), "synchronize");
        if (s->trace_block != nullptr) psim::hip_check(hipFree(s->trace_block), "trace free");
        s->trace_block = nullptr;
        s->host.trace = nullptr;
        s->host.trace_count = nullptr;
        s->host.trace_capacity = 0;
        if (capacity > 0) {
            psim::hip_check(hipMalloc(&s->trace_block, 256 + static_cast<size_t>(capacity) * 8), "trace allocation");
            psim::hip_check(hipMemset(s->trace_block, 0, 256), "trace clear");
            s->host.trace_count = static_cast<uint32_t*>(s->trace_block);
            s->host.trace = reinterpret_cast<uint64_t*>(static_cast<char*>(s->trace_block) + 256);
            s->host.trace_capacity = capacity;
        }
        psim::put(s->d_inst, &s->host, sizeof(Instance));
    });
}

int psim_system_trace_read(psim_system* s, uint64_t* ticks, int32_t capacity, int32_t* count) {
    return guarded([&] {
        if (s == nullptr || ticks == nullptr || count == nullptr) throw std::invalid_argument("null argument");
        if (s->trace_block == nullptr) throw std::logic_error("tracing is off");
        hipStream_t stream = s->graph->ctx->stream;
        uint32_t n = 0;
        psim::hip_check(hipMemcpyAsync(&n, s->host.trace_count, 4, hipMemcpyDeviceToHost, stream), "copy");
        psim::hip_check(hipStreamSynchronize(stream), "trace");
        const uint32_t m = std::min<uint32_t>({n, static_cast<uint32_t>(capacity),
                                               static_cast<uint32_t>(s->host.trace_capacity)});
        if (m > 0) psim::hip_check(hipMemcpy(ticks, s->host.trace, m * 8ULL, hipMemcpyDeviceToHost), "copy");
        *count = static_cast<int32_t>(n);
    });
}

int psim_system_level_values(psim_system* s, uint64_t level, double* diagonal, double* off) {
    return guarded([&] {
        require_matrix(s);
        if (level >= static_cast<uint64_t>(s->graph->depth)) throw std::invalid_argument("no such level");
        hipStream_t stream = s->graph->ctx->stream;
        const LevelView& v = s->graph->views[level];
        const psim::HostLevel& H = s->graph->host.levels[level];
        if (diagonal)
            psim::hip_check(hipMemcpyAsync(diagonal, s->host.diag64[level], static_cast<size_t>(v.n) * 8,
                                           hipMemcpyDeviceToHost, stream), "copy");
        if (off) {
            std::vector<double> ell(static_cast<size_t>(v.n) * static_cast<size_t>(v.width));
            psim::hip_check(hipMemcpyAsync(ell.data(), s->host.off64[level], ell.size() * 8, hipMemcpyDeviceToHost,
                                           stream), "copy");
            psim::hip_check(hipStreamSynchronize(stream), "level values");
            for (int32_t i = 0; i < H.n; ++i)
                for (int32_t k = H.row[i]; k < H.row[i + 1]; ++k) off[k] = ell[psim::ell_position(H, i, k)];
        }
        psim::hip_check(hipStreamSynchronize(stream), "level values");
    });
}

int psim_system_set_matrix(psim_system* s, const double* diagonal, uint64_t n, const double* off, uint64_t nnz,
                           int where) {
    return guarded([&] {
        if (s == nullptr || diagonal == nullptr || off == nullptr) throw std::invalid_argument("null argument");
        psim::check_where(where);
        if (n != static_cast<uint64_t>(s->graph->n) || nnz != static_cast<uint64_t>(s->graph->nnz))
            throw std::invalid_argument("multigrid matrix does not match the graph");
        hipStream_t stream = s->graph->ctx->stream;
        const hipMemcpyKind kind = where == PSIM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
        psim::hip_check(hipMemcpyAsync(s->host.diag64[0], diagonal, n * 8, kind, stream), "matrix copy");
        psim::hip_check(hipMemcpyAsync(s->csr_stage, off, nnz * 8, kind, stream), "matrix copy");
        Launch p{};
        p.op = psim::dev::kOpSetMatrix | psim::dev::kOpCsrValues;
        p.in = s->csr_stage;
        launch(s, p);
        if (where == PSIM_HOST) psim::hip_check(hipStreamSynchronize(stream), "set_matrix");
        s->matrix_set = true;
        s->heat_matrix = false;
    });
}

int psim_system_precondition(psim_system* s, const double* rhs, double* out, int where) {
    return guarded([&] {
        require_matrix(s);
        psim::check_where(where);
        if (rhs == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        hipStream_t stream = s->graph->ctx->stream;
        const size_t bytes = static_cast<size_t>(s->graph->n) * 8;
        Launch p{};
        p.op = psim::dev::kOpPrecondition;
        p.in = rhs;
        p.out = out;
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(s->stage_a, rhs, bytes, hipMemcpyHostToDevice, stream), "copy");
            p.in = s->stage_a;
            p.out = s->stage_b;
        }
        launch(s, p);
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(out, s->stage_b, bytes, hipMemcpyDeviceToHost, stream), "copy");
            psim::hip_check(hipStreamSynchronize(stream), "precondition");
        }
    });
}

int psim_system_apply(psim_system* s, const double* x, double* out, int where) {
    return guarded([&] {
        require_matrix(s);
        psim::check_where(where);
        if (x == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        hipStream_t stream = s->graph->ctx->stream;
        const size_t bytes = static_cast<size_t>(s->graph->n) * 8;
        Launch p{};
        p.op = psim::dev::kOpApply;
        p.in = x;
        p.out = out;
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(s->stage_a, x, bytes, hipMemcpyHostToDevice, stream), "copy");
            p.in = s->stage_a;
            p.out = s->stage_b;
        }
        launch(s, p);
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(out, s->stage_b, bytes, hipMemcpyDeviceToHost, stream), "copy");
            psim::hip_check(hipStreamSynchronize(stream), "apply");
        }
    });
}

int psim_system_solve(psim_system* s, const double* b, double* x, double tolerance, int32_t max_iterations,
                      psim_solve_info* info, int where) {
    return guarded([&] {
        require_matrix(s);
        psim::check_where(where);
        if (b == nullptr || x == nullptr) throw std::invalid_argument("null argument");
        if (!(tolerance >= 0.0) || max_iterations < 0) throw std::invalid_argument("invalid solver settings");
        hipStream_t stream = s->graph->ctx->stream;
        const size_t bytes = static_cast<size_t>(s->graph->n) * 8;
        Launch p = solve_params(s, psim::dev::kOpSolve, tolerance, max_iterations);
        p.in = b;
        p.out = x;
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(s->stage_a, b, bytes, hipMemcpyHostToDevice, stream), "copy");
            p.in = s->stage_a;
            p.out = s->stage_b;
        }
        launch(s, p);
        if (where == PSIM_HOST)
            psim::hip_check(hipMemcpyAsync(x, s->stage_b, bytes, hipMemcpyDeviceToHost, stream), "copy");
        if (where == PSIM_HOST || info != nullptr) fill_info(fetch_result(s), info, s->graph->ctx->wall_clock_hz);
    });
}

int psim_system_pcg(psim_system* s, const double* b, const uint8_t* hold, double* x, double tolerance,
                    int32_t max_iterations, psim_solve_info* info, int where) {
    return psim_system_pcg_operator(s, nullptr, b, hold, x, tolerance, max_iterations, info, where);
}

int psim_system_pcg_operator(psim_system* s, const double* operator_diagonal, const double* b, const uint8_t* hold,
                             double* x, double tolerance, int32_t max_iterations, psim_solve_info* info, int where) {
    return guarded([&] {
        require_matrix(s);
        psim::check_where(where);
        if (s->precision != PSIM_PLAIN_FP64) throw std::invalid_argument("PCG needs a plain FP64 system");
        if (b == nullptr || x == nullptr) throw std::invalid_argument("null argument");
        if (!(tolerance >= 0.0) || max_iterations < 0) throw std::invalid_argument("invalid solver settings");
        hipStream_t stream = s->graph->ctx->stream;
        const size_t bytes = static_cast<size_t>(s->graph->n) * 8;
        Launch p{};
        p.op = psim::dev::kOpPcg;
        p.tolerance = tolerance;
        p.max_iterations = max_iterations;
        p.in = b;
        p.out = x;
        p.hold = hold;
        p.operator_diag = operator_diagonal;
        psim::Staging staging;
        if (where == PSIM_HOST && operator_diagonal != nullptr)
            p.operator_diag = static_cast<const double*>(staging.in(operator_diagonal, bytes));
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(s->stage_a, b, bytes, hipMemcpyHostToDevice, stream), "copy");
            p.in = s->stage_a;
            p.out = s->stage_b;
            if (hold != nullptr) {
                // the mask rides in the third staging vector (n bytes <= 8 n)
                psim::hip_check(hipMemcpyAsync(s->stage_c, hold, static_cast<size_t>(s->graph->n), hipMemcpyHostToDevice,
                                               stream),
                                "copy");
                p.hold = reinterpret_cast<const uint8_t*>(s->stage_c);
            }
        }
        launch(s, p);
        if (where == PSIM_HOST)
            psim::hip_check(hipMemcpyAsync(x, s->stage_b, bytes, hipMemcpyDeviceToHost, stream), "copy");
        if (where == PSIM_HOST || info != nullptr || operator_diagonal != nullptr)
            fill_info(fetch_result(s), info, s->graph->ctx->wall_clock_hz);
    });
}

int psim_system_last_info(psim_system* s, psim_solve_info* info) {
    return guarded([&] {
        if (s == nullptr || info == nullptr) throw std::invalid_argument("null argument");
        fill_info(fetch_result(s), info, s->graph->ctx->wall_clock_hz);
    });
}

int psim_heat_set_matrix(psim_system* s, const double* conductance_per_entry, const double* capacity_area, double dt,
                         int where) {
    return guarded([&] {
        if (s == nullptr || conductance_per_entry == nullptr || capacity_area == nullptr)
            throw std::invalid_argument("null argument");
        psim::check_where(where);
        if (!(dt > 0.0) || !std::isfinite(dt)) throw std::invalid_argument("time step must be finite and positive");
        hipStream_t stream = s->graph->ctx->stream;
        Launch p{};
        p.op = psim::dev::kOpSetMatrix | psim::dev::kOpHeatAssemble;
        p.dt = dt;
        stage_heat_inputs(s, conductance_per_entry, capacity_area, where, p);
        launch(s, p);
        if (where == PSIM_HOST) psim::hip_check(hipStreamSynchronize(stream), "heat set_matrix");
        s->matrix_set = true;
        s->heat_matrix = true;
    });
}

int psim_heat_step(psim_system* s, const double* t_old, const double* forcing, double dt, double tolerance,
                   int32_t max_iterations, double* t_new, psim_solve_info* info, int where) {
    return guarded([&] {
        require_matrix(s);
        if (!s->heat_matrix) throw std::logic_error("psim_heat_set_matrix has not been called");
        psim::check_where(where);
        if (t_old == nullptr || forcing == nullptr || t_new == nullptr) throw std::invalid_argument("null argument");
        if (!(dt > 0.0) || !std::isfinite(dt)) throw std::invalid_argument("time step must be finite and positive");
        if (!(tolerance >= 0.0) || max_iterations < 0) throw std::invalid_argument("invalid solver settings");
        hipStream_t stream = s->graph->ctx->stream;
        const size_t bytes = static_cast<size_t>(s->graph->n) * 8;
        Launch p = solve_params(s, psim::dev::kOpSolve | psim::dev::kOpHeatRhs, tolerance, max_iterations);
        p.dt = dt;
        p.capacity_area = s->capacity_area;
        p.t_old = t_old;
        p.forcing = forcing;
        p.t_new = t_new;
        if (where == PSIM_HOST) {
            psim::hip_check(hipMemcpyAsync(s->stage_a, t_old, bytes, hipMemcpyHostToDevice, stream), "copy");
            psim::hip_check(hipMemcpyAsync(s->stage_b, forcing, bytes, hipMemcpyHostToDevice, stream), "copy");
            p.t_old = s->stage_a;
            p.forcing = s->stage_b;
            p.t_new = s->stage_c;
        }
        launch(s, p);
        if (where == PSIM_HOST)
            psim::hip_check(hipMemcpyAsync(t_new, s->stage_c, bytes, hipMemcpyDeviceToHost, stream), "copy");
        if (where == PSIM_HOST || info != nullptr) fill_info(fetch_result(s), info, s->graph->ctx->wall_clock_hz);
    });
}

int psim_heat_chain(psim_system* s, const double* conductance_per_entry, const double* capacity_area,
                    const double* t_old, const double* forcings, int32_t steps, double dt, double tolerance,
                    int32_t max_iterations, double* t_new, psim_solve_info* step_info, int where) {
    return guarded([&] {
        if (s == nullptr) throw std::invalid_argument("null system");
        const bool assemble = conductance_per_entry != nullptr && capacity_area != nullptr;
        if (!assemble) {
            require_matrix(s);
            if (!s->heat_matrix) throw std::logic_error("psim_heat_set_matrix has not been called");
        }
        psim::check_where(where);
        if (t_old == nullptr || forcings == nullptr || t_new == nullptr) throw std::invalid_argument("null argument");
        if (steps < 1) throw std::invalid_argument("a chain needs at least one step");
        if (!(dt > 0.0) || !std::isfinite(dt)) throw std::invalid_argument("time step must be finite and positive");
        if (!(tolerance >= 0.0) || max_iterations < 0) throw std::invalid_argument("invalid solver settings");
        hipStream_t stream = s->graph->ctx->stream;
        const size_t n = static_cast<size_t>(s->graph->n), bytes = n * 8;
        int op = psim::dev::kOpSolve | psim::dev::kOpHeatRhs | psim::dev::kOpChain;
        if (assemble) op |= psim::dev::kOpSetMatrix | psim::dev::kOpHeatAssemble;
        Launch p = solve_params(s, op, tolerance, max_iterations);
        p.steps = steps;
        p.dt = dt;
        p.capacity_area = s->capacity_area;
        if (assemble) stage_heat_inputs(s, conductance_per_entry, capacity_area, where, p);
        p.t_old = t_old;
        p.forcing = forcings;
        p.t_new = t_new;
        std::vector<void*> staged;
        struct Free {
            std::vector<void*>& v;
            ~Free() {
                for (void* d : v) (void)hipFree(d);
            }
        } release{staged};
        const auto stage = [&](size_t b) {
            void* d = nullptr;
            psim::hip_check(hipMalloc(&d, b), "staging allocation");
            staged.push_back(d);
            return d;
        };
        if (where == PSIM_HOST) {
            double* d_t = static_cast<double*>(stage(bytes));
            double* d_f = static_cast<double*>(stage(bytes * static_cast<size_t>(steps)));
            psim::hip_check(hipMemcpyAsync(d_t, t_old, bytes, hipMemcpyHostToDevice, stream), "copy");
            psim::hip_check(hipMemcpyAsync(d_f, forcings, bytes * static_cast<size_t>(steps), hipMemcpyHostToDevice,
                                           stream), "copy");
            p.t_old = d_t;
            p.forcing = d_f;
            p.t_new = s->stage_c;
        }
        SolveResult* d_steps = nullptr;
        if (step_info != nullptr) {
            d_steps = static_cast<SolveResult*>(stage(sizeof(SolveResult) * static_cast<size_t>(steps)));
            p.step_results = d_steps;
        }
        launch(s, p);
        if (assemble) {
            s->matrix_set = true;
            s->heat_matrix = true;
        }
        if (where == PSIM_HOST)
            psim::hip_check(hipMemcpyAsync(t_new, s->stage_c, bytes, hipMemcpyDeviceToHost, stream), "copy");
        if (step_info != nullptr) {
            std::vector<SolveResult> host(static_cast<size_t>(steps));
            psim::hip_check(hipMemcpyAsync(host.data(), d_steps, host.size() * sizeof(SolveResult),
                                           hipMemcpyDeviceToHost, stream), "copy");
            (void)fetch_result(s);
            for (int32_t k = 0; k < steps; ++k)
                fill_info(host[static_cast<size_t>(k)], step_info + k, s->graph->ctx->wall_clock_hz);
        } else if (!staged.empty()) {
            // staged buffers must outlive the launch
            (void)fetch_result(s);
        }
    });
}

}  // extern "C"
