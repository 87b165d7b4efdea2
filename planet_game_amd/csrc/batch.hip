// Batches: many systems, of any graphs and precisions on one context, in
// one launch of the persistent solver program. Each system owns a group of
// workgroups for the whole launch and runs its own phases and barriers, so
// instances of different sizes and iteration counts never wait for each
// other; a system's results do not depend on its group size, its place in
// the batch or the other members (reductions are per tile, in fixed order).
//
// Workgroup allocation: every instance gets one workgroup, then the rest of
// the device's co-resident workgroups go, one at a time, to the instance
// with the most tiles per workgroup (its phase length). The program runs at
// 3 workgroups per CU (168 VGPRs each) or 2 (256 VGPRs, fewer spills and less
// contention per CU, fewer workgroups). The autotuner times alternative
// occupancies and allocations on the caller's real workload and keeps the
// best.
#include "system.hpp"

#include <array>
#include <cmath>
#include <memory>
#include <queue>
#include <vector>

using psim::guarded;
using psim::dev::Instance;
using psim::dev::Launch;
using psim::dev::SolveResult;

namespace {

constexpr int kSlots = 4;   // launch descriptors in flight per batch

struct Slot {
    Launch* h_launch = nullptr;       // pinned
    SolveResult* h_result = nullptr;  // pinned
    Launch* d_launch = nullptr;
    SolveResult* d_result = nullptr;
    hipEvent_t copied = nullptr;      // after the descriptor upload
    bool used = false;
};

}  // namespace

struct psim_batch {
    psim_ctx* ctx = nullptr;
    std::vector<psim_system*> systems;
    std::vector<int32_t> groups;   // workgroups per instance
    std::vector<int32_t> work;     // level-0 tiles per instance
    int32_t capacity = 0;          // co-resident program workgroups at this occupancy
    int32_t waves = 3;             // program instance: 2 or 3 workgroups per CU
    int32_t map_capacity = 0;      // workgroup-map entries allocated (the largest capacity)
    int32_t blocks = 0;            // sum of groups
    psim::Arena arena;
    Instance* d_inst = nullptr;
    int32_t* d_map = nullptr;
    SolveResult* d_sticky = nullptr;   // per-instance result words (barrier timeouts)
    std::vector<double*> tri;          // packed coarsest scratch for systems that keep it in LDS alone
    std::array<Slot, kSlots> slots{};
    int next = 0;
    ~psim_batch() {
        if (ctx != nullptr) (void)hipStreamSynchronize(ctx->stream);
        for (Slot& s : slots) {
            if (s.h_launch != nullptr) (void)hipHostFree(s.h_launch);
            if (s.h_result != nullptr) (void)hipHostFree(s.h_result);
            if (s.copied != nullptr) (void)hipEventDestroy(s.copied);
        }
        arena.release();
    }
};

namespace {

// Greedy allocation of `budget` workgroups (>= count): one each, then one
// more at a time to the instance with the largest ceil(weight / groups).
std::vector<int32_t> allocate(const std::vector<double>& weight, const std::vector<int32_t>& cap, int32_t budget) {
    const size_t count = weight.size();
    std::vector<int32_t> g(count, 1);
    using Item = std::pair<double, size_t>;
    std::priority_queue<Item> heap;
    for (size_t i = 0; i < count; ++i)
        if (cap[i] > 1) heap.push({weight[i], i});
    int32_t left = budget - static_cast<int32_t>(count);
    while (left > 0 && !heap.empty()) {
        const size_t i = heap.top().second;
        heap.pop();
        ++g[i];
        --left;
        if (g[i] < cap[i]) heap.push({weight[i] / g[i], i});
    }
    return g;
}

void upload_layout(psim_batch* b) {
    const size_t count = b->systems.size();
    std::vector<Instance> inst(count);
    std::vector<int32_t> map;
    map.reserve(static_cast<size_t>(b->blocks));
    int32_t first = 0;
    for (size_t i = 0; i < count; ++i) {
        Instance I = b->systems[i]->host;
        I.group_first = first;
        I.group_size = b->groups[i];
        I.result = b->d_sticky + i;
        I.trace = nullptr;
        I.trace_count = nullptr;
        I.trace_capacity = 0;
        if (I.tri_lds) {
            // a batch launch carries no dynamic LDS: the sweep uses scratch
            I.tri_lds = 0;
            I.tri = b->tri[i];
        }
        inst[i] = I;
        map.insert(map.end(), static_cast<size_t>(b->groups[i]), static_cast<int32_t>(i));
        first += b->groups[i];
    }
    hipStream_t stream = b->ctx->stream;
    psim::hip_check(hipStreamSynchronize(stream), "batch layout");
    psim::hip_check(hipMemcpy(b->d_inst, inst.data(), count * sizeof(Instance), hipMemcpyHostToDevice), "upload");
    psim::hip_check(hipMemcpy(b->d_map, map.data(), map.size() * sizeof(int32_t), hipMemcpyHostToDevice), "upload");
}

void set_groups(psim_batch* b, std::vector<int32_t> groups) {
    int64_t total = 0;
    for (size_t i = 0; i < groups.size(); ++i) {
        if (groups[i] < 1) throw std::invalid_argument("every instance needs at least one workgroup");
        total += groups[i];
    }
    if (total > b->capacity) throw std::invalid_argument("workgroups exceed the device's co-resident capacity");
    b->groups = std::move(groups);
    b->blocks = static_cast<int32_t>(total);
    upload_layout(b);
}

std::vector<int32_t> default_groups(const psim_batch* b, double exponent, double budget_fraction) {
    std::vector<double> weight(b->work.size());
    std::vector<int32_t> cap(b->work.size());
    for (size_t i = 0; i < weight.size(); ++i) {
        weight[i] = std::pow(static_cast<double>(b->work[i]), exponent);
        cap[i] = b->work[i];
    }
    const int32_t budget = std::max<int32_t>(static_cast<int32_t>(b->systems.size()),
                                             static_cast<int32_t>(budget_fraction * b->capacity));
    return allocate(weight, cap, budget);
}

struct StepArgs {
    const double* const* conductance;
    const double* const* capacity;
    const double* const* t_old;
    const double* const* forcing;
    double dt, tolerance;
    int32_t max_iterations;
    double* const* t_new;
};

void check_args(const psim_batch* b, const StepArgs& a) {
    if (a.t_old == nullptr || a.forcing == nullptr || a.t_new == nullptr) throw std::invalid_argument("null argument");
    if (!(a.dt > 0.0) || !std::isfinite(a.dt)) throw std::invalid_argument("time step must be finite and positive");
    if (!(a.tolerance >= 0.0) || a.max_iterations < 0) throw std::invalid_argument("invalid solver settings");
    for (size_t i = 0; i < b->systems.size(); ++i)
        if (a.t_old[i] == nullptr || a.forcing[i] == nullptr || a.t_new[i] == nullptr)
            throw std::invalid_argument("null argument");
    for (size_t i = 0; i < b->systems.size(); ++i) {
        const bool k = a.conductance != nullptr && a.conductance[i] != nullptr;
        const bool c = a.capacity != nullptr && a.capacity[i] != nullptr;
        if (k != c) throw std::invalid_argument("assembly needs both conductance and capacity");
        if (!k && !b->systems[i]->heat_matrix) throw std::logic_error("psim_heat_set_matrix has not been called");
    }
}

// Queue one batched heat step; returns the slot whose records it fills.
Slot& enqueue(psim_batch* b, const StepArgs& a) {
    hipStream_t stream = b->ctx->stream;
    Slot& slot = b->slots[static_cast<size_t>(b->next)];
    b->next = (b->next + 1) % kSlots;
    // the slot's pinned descriptors are free once their last upload ran
    if (slot.used) psim::hip_check(hipEventSynchronize(slot.copied), "batch slot");
    const size_t count = b->systems.size();
    for (size_t i = 0; i < count; ++i) {
        psim_system* s = b->systems[i];
        const bool assemble = a.conductance != nullptr && a.conductance[i] != nullptr;
        int op = psim::dev::kOpSolve | psim::dev::kOpHeatRhs;
        if (assemble) op |= psim::dev::kOpSetMatrix | psim::dev::kOpHeatAssemble;
        Launch p = psim::sys::solve_params(s, op, a.tolerance, a.max_iterations);
        p.dt = a.dt;
        p.capacity_area = s->capacity_area;
        if (assemble) psim::sys::stage_heat_inputs(s, a.conductance[i], a.capacity[i], PSIM_DEVICE, p);
        p.t_old = a.t_old[i];
        p.forcing = a.forcing[i];
        p.t_new = a.t_new[i];
        p.step_results = slot.d_result + i;
        slot.h_launch[i] = p;
    }
    psim::hip_check(hipMemcpyAsync(slot.d_launch, slot.h_launch, count * sizeof(Launch), hipMemcpyHostToDevice, stream),
                    "batch upload");
    psim::hip_check(hipEventRecord(slot.copied, stream), "batch event");
    slot.used = true;
    psim::hip_check(psim::dev::launch_program(b->d_inst, b->d_map, slot.d_launch, Launch{}, b->blocks, 0, stream,
                                              b->blocks > 1, b->waves),
                    "batch launch");
    for (size_t i = 0; i < count; ++i)
        if (a.conductance != nullptr && a.conductance[i] != nullptr) {
            b->systems[i]->matrix_set = true;
            b->systems[i]->heat_matrix = true;
        }
    return slot;
}

// Co-resident program workgroups on the device at `waves` workgroups per CU.
int32_t capacity_at(const psim_ctx* ctx, int waves) {
    return ctx->compute_units * std::min(waves, psim::dev::program_blocks_per_cu(0, waves));
}

// Switch the program instance (2 or 3 workgroups per CU) and go back to the
// default allocation of the new capacity.
void set_waves(psim_batch* b, int waves) {
    if (waves != 2 && waves != 3) throw std::invalid_argument("occupancy must be 2 or 3 workgroups per CU");
    const int32_t capacity = capacity_at(b->ctx, waves);
    if (static_cast<int32_t>(b->systems.size()) > capacity)
        throw std::invalid_argument("batch larger than the device's co-resident workgroups at that occupancy");
    b->waves = waves;
    b->capacity = capacity;
    set_groups(b, default_groups(b, 1.0, 1.0));
}

// Wait for the stream, check the barrier words, copy the slot's records.
void collect(psim_batch* b, Slot& slot, psim_solve_info* infos) {
    hipStream_t stream = b->ctx->stream;
    const size_t count = b->systems.size();
    std::vector<SolveResult> sticky(count);
    psim::hip_check(hipMemcpyAsync(sticky.data(), b->d_sticky, count * sizeof(SolveResult), hipMemcpyDeviceToHost,
                                   stream),
                    "batch results");
    if (infos != nullptr)
        psim::hip_check(hipMemcpyAsync(slot.h_result, slot.d_result, count * sizeof(SolveResult),
                                       hipMemcpyDeviceToHost, stream),
                        "batch results");
    psim::hip_check(hipStreamSynchronize(stream), "batch solve");
    bool timed_out = false;
    for (const SolveResult& r : sticky) timed_out = timed_out || r.barrier_timeouts != 0;
    if (timed_out) {
        for (psim_system* s : b->systems)
            psim::hip_check(hipMemset(s->host.barrier, 0, psim::dev::kBarrierWordCount * sizeof(uint32_t)),
                            "barrier reset");
        psim::hip_check(hipMemset(b->d_sticky, 0, count * sizeof(SolveResult)), "result reset");
        throw psim::DeviceFailure("solver group barrier timed out");
    }
    if (infos != nullptr)
        for (size_t i = 0; i < count; ++i)
            psim::sys::fill_info(slot.h_result[i], infos + i, b->ctx->wall_clock_hz);
}

}  // namespace

extern "C" {

int psim_batch_create(psim_ctx* ctx, psim_system* const* systems, int32_t count, psim_batch** out) {
    return guarded([&] {
        if (ctx == nullptr || systems == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        if (count < 1) throw std::invalid_argument("a batch needs at least one system");
        if (!ctx->cooperative) throw psim::DeviceFailure("cooperative launch unsupported on this device");
        psim::hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        auto b = std::make_unique<psim_batch>();
        b->ctx = ctx;
        b->capacity = capacity_at(ctx, 3);
        b->map_capacity = std::max(b->capacity, capacity_at(ctx, 2));
        if (count > b->capacity) throw std::invalid_argument("batch larger than the device's co-resident workgroups");
        for (int32_t i = 0; i < count; ++i) {
            psim_system* s = systems[i];
            if (s == nullptr) throw std::invalid_argument("null system");
            if (s->graph->ctx != ctx) throw std::invalid_argument("every system of a batch must share its context");
            for (int32_t k = 0; k < i; ++k)
                if (systems[k] == s) throw std::invalid_argument("a system appears twice in the batch");
            b->systems.push_back(s);
            b->work.push_back(s->host.tiles);
        }
        auto& A = b->arena;
        const size_t n = static_cast<size_t>(count);
        A.plan(&b->d_inst, n);
        A.plan(&b->d_map, static_cast<size_t>(b->map_capacity));
        A.plan(&b->d_sticky, n);
        b->tri.assign(n, nullptr);
        for (size_t i = 0; i < n; ++i)
            if (b->systems[i]->host.tri_lds) A.plan(&b->tri[i], b->systems[i]->tri_bytes / sizeof(double));
        for (Slot& s : b->slots) {
            A.plan(&s.d_launch, n);
            A.plan(&s.d_result, n);
        }
        A.allocate();
        psim::hip_check(hipMemset(b->d_sticky, 0, n * sizeof(SolveResult)), "result reset");
        for (Slot& s : b->slots) {
            psim::hip_check(hipHostMalloc(reinterpret_cast<void**>(&s.h_launch), n * sizeof(Launch)), "pinned allocation");
            psim::hip_check(hipHostMalloc(reinterpret_cast<void**>(&s.h_result), n * sizeof(SolveResult)),
                            "pinned allocation");
            psim::hip_check(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming), "event");
        }
        set_groups(b.get(), default_groups(b.get(), 1.0, 1.0));
        *out = b.release();
    });
}

void psim_batch_destroy(psim_batch* batch) { delete batch; }

int psim_batch_size(const psim_batch* batch, int32_t* count, int32_t* capacity) {
    return guarded([&] {
        if (batch == nullptr) throw std::invalid_argument("null batch");
        if (count != nullptr) *count = static_cast<int32_t>(batch->systems.size());
        if (capacity != nullptr) *capacity = batch->capacity;
    });
}

int psim_batch_groups(const psim_batch* batch, int32_t* groups) {
    return guarded([&] {
        if (batch == nullptr || groups == nullptr) throw std::invalid_argument("null argument");
        std::copy(batch->groups.begin(), batch->groups.end(), groups);
    });
}

int psim_batch_set_groups(psim_batch* batch, const int32_t* groups) {
    return guarded([&] {
        if (batch == nullptr) throw std::invalid_argument("null batch");
        if (groups == nullptr)
            set_groups(batch, default_groups(batch, 1.0, 1.0));
        else
            set_groups(batch, std::vector<int32_t>(groups, groups + batch->systems.size()));
    });
}

int psim_batch_set_occupancy(psim_batch* batch, int32_t workgroups_per_cu) {
    return guarded([&] {
        if (batch == nullptr) throw std::invalid_argument("null batch");
        set_waves(batch, workgroups_per_cu);
    });
}

int32_t psim_batch_occupancy(const psim_batch* batch) { return batch ? batch->waves : 0; }

int psim_batch_heat_step(psim_batch* batch, const double* const* conductance_per_entry,
                         const double* const* capacity_area, const double* const* t_old, const double* const* forcing,
                         double dt, double tolerance, int32_t max_iterations, double* const* t_new,
                         psim_solve_info* infos) {
    return guarded([&] {
        if (batch == nullptr) throw std::invalid_argument("null batch");
        const StepArgs a{conductance_per_entry, capacity_area, t_old, forcing, dt, tolerance, max_iterations, t_new};
        check_args(batch, a);
        Slot& slot = enqueue(batch, a);
        if (infos != nullptr) collect(batch, slot, infos);
    });
}

int psim_batch_synchronize(psim_batch* batch) {
    return guarded([&] {
        if (batch == nullptr) throw std::invalid_argument("null batch");
        Slot& last = batch->slots[static_cast<size_t>((batch->next + kSlots - 1) % kSlots)];
        collect(batch, last, nullptr);
    });
}

int psim_batch_autotune(psim_batch* batch, const double* const* conductance_per_entry,
                        const double* const* capacity_area, const double* const* t_old, const double* const* forcing,
                        double dt, double tolerance, int32_t max_iterations, double* const* t_new, int32_t repeats,
                        double* best_ms) {
    return guarded([&] {
        if (batch == nullptr) throw std::invalid_argument("null batch");
        if (repeats < 1) throw std::invalid_argument("repeats must be positive");
        const StepArgs a{conductance_per_entry, capacity_area, t_old, forcing, dt, tolerance, max_iterations, t_new};
        check_args(batch, a);
        // candidates: the program instance (2 or 3 workgroups per CU: a
        // register budget of 256 or 168 VGPRs), and, for each, weight =
        // tiles^e over a share of its co-resident workgroups
        std::vector<std::pair<int, std::vector<int32_t>>> candidates;
        for (int waves : {3, 2}) {
            if (static_cast<int32_t>(batch->systems.size()) > capacity_at(batch->ctx, waves)) continue;
            batch->waves = waves;
            batch->capacity = capacity_at(batch->ctx, waves);
            for (double fraction : {1.0, 0.75, 0.5, 0.25})
                for (double exponent : {1.0, 0.75, 0.5}) {
                    auto g = default_groups(batch, exponent, fraction);
                    bool seen = false;
                    for (const auto& c : candidates) seen = seen || (c.first == waves && c.second == g);
                    if (!seen) candidates.emplace_back(waves, std::move(g));
                }
        }
        hipStream_t stream = batch->ctx->stream;
        hipEvent_t start = nullptr, stop = nullptr;
        psim::hip_check(hipEventCreate(&start), "event");
        psim::hip_check(hipEventCreate(&stop), "event");
        struct Events {
            hipEvent_t a, b;
            ~Events() {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        } release{start, stop};
        double best = 0.0;
        size_t best_index = 0;
        const auto apply = [&](size_t c) {
            batch->waves = candidates[c].first;
            batch->capacity = capacity_at(batch->ctx, batch->waves);
            set_groups(batch, candidates[c].second);
        };
        for (size_t c = 0; c < candidates.size(); ++c) {
            apply(c);
            collect(batch, enqueue(batch, a), nullptr);   // warm
            psim::hip_check(hipEventRecord(start, stream), "event");
            for (int32_t r = 0; r < repeats; ++r) (void)enqueue(batch, a);
            psim::hip_check(hipEventRecord(stop, stream), "event");
            collect(batch, batch->slots[static_cast<size_t>((batch->next + kSlots - 1) % kSlots)], nullptr);
            float ms = 0.0f;
            psim::hip_check(hipEventElapsedTime(&ms, start, stop), "event time");
            const double per = static_cast<double>(ms) / repeats;
            if (c == 0 || per < best) {
                best = per;
                best_index = c;
            }
        }
        apply(best_index);
        if (best_ms != nullptr) *best_ms = best;
    });
}

}  // extern "C"
