// Serving: continuous batching of heat-step requests on one persistent
// launch.
//
// The device's co-resident solver workgroups are split into slots, each
// serving one graph (mesh level) with a fixed number of workgroups; slot
// sizes follow the graph's tiles and slot counts the expected share of
// work per graph. Systems (matrix values, solver vectors) come from a pool
// per graph whose size is set by an HBM budget: a request that starts a
// session is admitted when its graph's pool has a free system, keeps it for
// the session's later steps (T and the matrix stay resident) and frees it
// after the last one. Requests are published to per-graph queues in pinned
// host memory while the server runs; a slot takes the next one as soon as
// it finishes the previous, so requests of every size overlap on the device
// and a late arrival does not wait for a batch to drain.
#include "system.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

using psim::guarded;
using psim::dev::Instance;
using psim::dev::kMaxQueues;
using psim::dev::ServeCompletion;
using psim::dev::ServeParams;
using psim::dev::ServeQueue;
using psim::dev::ServeRequest;
using psim::dev::ServeSlot;
using psim::dev::SolveResult;

namespace {

constexpr int32_t kRing = 1024;           // request entries per queue
constexpr int32_t kTargetTilesPerGroup = 12;   // tiles per workgroup per phase when sizing slots
constexpr uint32_t kMailboxWords = 96;    // generation, position, command (128-byte apart)
constexpr int64_t kMaxPool = 4096;        // systems per graph

template <class T> T* pinned(size_t count) {
    void* p = nullptr;
    psim::hip_check(hipHostMalloc(&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocCoherent | hipHostMallocMapped),
                    "pinned allocation");
    std::memset(p, 0, std::max<size_t>(count, 1) * sizeof(T));
    return static_cast<T*>(p);
}

}  // namespace

struct psim_server {
    psim_ctx* ctx = nullptr;
    int precision = PSIM_MIXED;
    int waves = 3;                     // server program instance: workgroups per CU
    std::vector<psim_graph*> graphs;
    std::vector<int32_t> slot_count, slot_size, pool_size;
    std::vector<size_t> footprint;
    // pool systems, per graph, and their index in the instance table
    std::vector<std::vector<psim_system*>> pool;
    std::vector<std::vector<int32_t>> pool_state;
    // per table index: the matrix key and dt its system last assembled (key 0:
    // none that can be reused); kept across runs
    std::vector<uint64_t> matrix_key;
    std::vector<double> matrix_dt;
    std::vector<psim_system*> owned;
    int32_t blocks = 0;
    psim::Arena arena;
    Instance* d_table = nullptr;
    ServeSlot* d_slots = nullptr;
    int32_t* d_block_slot = nullptr;
    uint64_t* d_heads = nullptr;
    uint32_t* d_words = nullptr;       // per slot: barrier words, then mailbox words
    SolveResult* d_results = nullptr;
    // pinned host memory
    ServeRequest* rings = nullptr;     // graphs x kRing
    uint64_t* tails = nullptr;         // graphs (one per 64-byte line)
    ServeCompletion* completions = nullptr;
    uint32_t* stop = nullptr;
    int32_t completion_capacity = 0;
    ~psim_server() {
        if (ctx != nullptr) (void)hipStreamSynchronize(ctx->stream);
        for (void* p : {static_cast<void*>(rings), static_cast<void*>(tails), static_cast<void*>(completions),
                        static_cast<void*>(stop)})
            if (p != nullptr) (void)hipHostFree(p);
        for (psim_system* s : owned) psim_system_destroy(s);
        arena.release();
    }
};

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// Slot plan: slot size from the graph's tiles, slot counts from each
// graph's share of the expected work (request share x tiles).
void plan_slots(psim_server* s, const std::vector<double>& share, const int32_t* sizes, int32_t capacity) {
    const size_t k = s->graphs.size();
    std::vector<double> work(k);
    double total = 0.0;
    for (size_t q = 0; q < k; ++q) {
        const int32_t tiles = (s->graphs[q]->n + psim::dev::kTile - 1) / psim::dev::kTile;
        s->slot_size[q] = sizes != nullptr && sizes[q] > 0
                              ? std::min(sizes[q], capacity)
                              : std::max<int32_t>(1, (tiles + kTargetTilesPerGroup - 1) / kTargetTilesPerGroup);
        work[q] = share[q] * tiles;
        total += work[q];
    }
    int32_t used = 0;
    for (size_t q = 0; q < k; ++q) {
        const double groups = total > 0.0 ? work[q] / total * capacity : 0.0;
        s->slot_count[q] = std::max<int32_t>(1, static_cast<int32_t>(groups / s->slot_size[q]));
        used += s->slot_count[q] * s->slot_size[q];
    }
    // trim the largest consumers until the plan fits the co-resident workgroups
    while (used > capacity) {
        size_t worst = 0;
        for (size_t q = 1; q < k; ++q)
            if (s->slot_count[q] * s->slot_size[q] > s->slot_count[worst] * s->slot_size[worst]) worst = q;
        if (s->slot_count[worst] <= 1) {
            if (s->slot_size[worst] <= 1) throw std::invalid_argument("too many graphs for the device");
            --s->slot_size[worst];
        } else {
            --s->slot_count[worst];
        }
        used = 0;
        for (size_t q = 0; q < k; ++q) used += s->slot_count[q] * s->slot_size[q];
    }
    s->blocks = used;
}

}  // namespace

extern "C" {

int psim_server_create(psim_ctx* ctx, psim_graph* const* graphs, const double* request_share, const int32_t* slot_sizes,
                       int32_t graph_count, int precision, uint64_t memory_budget, psim_server** out) {
    return psim_server_create_occupancy(ctx, graphs, request_share, slot_sizes, graph_count, precision, memory_budget, 0,
                                        out);
}

int psim_server_create_occupancy(psim_ctx* ctx, psim_graph* const* graphs, const double* request_share,
                                 const int32_t* slot_sizes, int32_t graph_count, int precision, uint64_t memory_budget,
                                 int32_t workgroups_per_cu, psim_server** out) {
    return guarded([&] {
        if (workgroups_per_cu == 0) workgroups_per_cu = 3;
        if (workgroups_per_cu != 2 && workgroups_per_cu != 3)
            throw std::invalid_argument("occupancy must be 2 or 3 workgroups per CU");
        if (ctx == nullptr || graphs == nullptr || out == nullptr) throw std::invalid_argument("null argument");
        if (graph_count < 1 || graph_count > kMaxQueues) throw std::invalid_argument("graph count out of range");
        if (precision != PSIM_PLAIN_FP64 && precision != PSIM_MIXED) throw std::invalid_argument("unknown precision");
        if (!ctx->cooperative) throw psim::DeviceFailure("cooperative launch unsupported on this device");
        psim::hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        auto s = std::make_unique<psim_server>();
        s->ctx = ctx;
        s->precision = precision;
        const size_t k = static_cast<size_t>(graph_count);
        std::vector<double> share(k, 1.0);
        for (size_t q = 0; q < k; ++q) {
            psim_graph* g = graphs[q];
            if (g == nullptr) throw std::invalid_argument("null graph");
            if (g->ctx != ctx) throw std::invalid_argument("every graph must share the server's context");
            if (g->views[g->depth - 1].n > psim::dev::kBlockedCoarse)
                throw std::invalid_argument("served graphs need a coarsest level of at most 168 rows");
            s->graphs.push_back(g);
            if (request_share != nullptr) {
                if (!(request_share[q] >= 0.0) || !std::isfinite(request_share[q]))
                    throw std::invalid_argument("request shares must be finite and non-negative");
                share[q] = request_share[q];
            }
        }
        s->slot_count.assign(k, 0);
        s->slot_size.assign(k, 0);
        s->pool_size.assign(k, 0);
        s->footprint.assign(k, 0);
        s->waves = workgroups_per_cu;
        const int per_cu = std::min(workgroups_per_cu, psim::dev::server_blocks_per_cu(workgroups_per_cu));
        if (per_cu < 1) throw psim::DeviceFailure("the server kernel cannot be resident");
        plan_slots(s.get(), share, slot_sizes, ctx->compute_units * per_cu);
        // pools: one system per slot at least, the rest of the budget split by
        // the expected resident memory (share x footprint)
        for (size_t q = 0; q < k; ++q) {
            psim_system* probe = nullptr;
            if (psim_system_create(s->graphs[q], precision, &probe) != PSIM_OK)
                throw psim::DeviceOutOfMemory(psim_last_error());
            s->footprint[q] = probe->arena.bytes();
            s->owned.push_back(probe);
        }
        // pools: the budget split in proportion to each graph's slots x bytes
        // per system (at least one system per slot; a budget of 0 means
        // exactly that minimum)
        double slot_bytes = 0.0;
        uint64_t minimum = 0;
        for (size_t q = 0; q < k; ++q) {
            slot_bytes += static_cast<double>(s->slot_count[q]) * static_cast<double>(s->footprint[q]);
            minimum += static_cast<uint64_t>(s->slot_count[q]) * s->footprint[q];
        }
        if (memory_budget != 0 && memory_budget < minimum)
            throw std::invalid_argument("memory budget below one system per slot");
        for (size_t q = 0; q < k; ++q) {
            const double fit = static_cast<double>(memory_budget) * s->slot_count[q] / slot_bytes;
            const int64_t count = std::max<int64_t>(s->slot_count[q], static_cast<int64_t>(fit));
            s->pool_size[q] = static_cast<int32_t>(std::min<int64_t>(count, kMaxPool));
        }
        s->pool.resize(k);
        s->pool_state.resize(k);
        std::vector<Instance> table;
        for (size_t q = 0; q < k; ++q) {
            for (int32_t i = 0; i < s->pool_size[q]; ++i) {
                psim_system* sys = nullptr;
                if (i == 0) {
                    sys = s->owned[q];
                } else {
                    if (psim_system_create(s->graphs[q], precision, &sys) != PSIM_OK)
                        throw psim::DeviceOutOfMemory(psim_last_error());
                    s->owned.push_back(sys);
                }
                s->pool[q].push_back(sys);
                s->pool_state[q].push_back(static_cast<int32_t>(table.size()));
                Instance I = sys->host;
                I.tri_lds = 0;   // server slots carry no dynamic LDS (coarsest <= 168 rows)
                table.push_back(I);
            }
        }
        s->matrix_key.assign(table.size(), 0);
        s->matrix_dt.assign(table.size(), 0.0);
        int32_t slots = 0;
        for (size_t q = 0; q < k; ++q) slots += s->slot_count[q];
        auto& A = s->arena;
        A.plan(&s->d_table, table.size());
        A.plan(&s->d_slots, static_cast<size_t>(slots));
        A.plan(&s->d_block_slot, static_cast<size_t>(s->blocks));
        A.plan(&s->d_heads, k * 8);
        A.plan(&s->d_words, static_cast<size_t>(slots) * (psim::dev::kBarrierWordCount + kMailboxWords));
        A.plan(&s->d_results, static_cast<size_t>(slots));
        A.allocate();
        psim::hip_check(hipMemcpy(s->d_table, table.data(), table.size() * sizeof(Instance), hipMemcpyHostToDevice),
                        "upload");
        std::vector<ServeSlot> slot_list;
        std::vector<int32_t> block_slot;
        int32_t first = 0;
        for (size_t q = 0; q < k; ++q)
            for (int32_t c = 0; c < s->slot_count[q]; ++c) {
                const size_t index = slot_list.size();
                ServeSlot sl{};
                sl.first = first;
                sl.size = s->slot_size[q];
                sl.queue = static_cast<int32_t>(q);
                sl.barrier = s->d_words + index * (psim::dev::kBarrierWordCount + kMailboxWords);
                sl.mailbox = sl.barrier + psim::dev::kBarrierWordCount;
                sl.result = s->d_results + index;
                slot_list.push_back(sl);
                block_slot.insert(block_slot.end(), static_cast<size_t>(sl.size), static_cast<int32_t>(index));
                first += sl.size;
            }
        psim::hip_check(hipMemcpy(s->d_slots, slot_list.data(), slot_list.size() * sizeof(ServeSlot),
                                  hipMemcpyHostToDevice),
                        "upload");
        psim::hip_check(hipMemcpy(s->d_block_slot, block_slot.data(), block_slot.size() * sizeof(int32_t),
                                  hipMemcpyHostToDevice),
                        "upload");
        s->rings = pinned<ServeRequest>(k * kRing);
        s->tails = pinned<uint64_t>(k * 8);
        s->completion_capacity = static_cast<int32_t>(k) * kRing;
        s->completions = pinned<ServeCompletion>(static_cast<size_t>(s->completion_capacity));
        s->stop = pinned<uint32_t>(1);
        *out = s.release();
    });
}

void psim_server_destroy(psim_server* server) { delete server; }

int psim_server_plan(const psim_server* s, int32_t* slots, int32_t* slot_size, int32_t* pool, uint64_t* footprint) {
    return guarded([&] {
        if (s == nullptr) throw std::invalid_argument("null server");
        for (size_t q = 0; q < s->graphs.size(); ++q) {
            if (slots != nullptr) slots[q] = s->slot_count[q];
            if (slot_size != nullptr) slot_size[q] = s->slot_size[q];
            if (pool != nullptr) pool[q] = s->pool_size[q];
            if (footprint != nullptr) footprint[q] = s->footprint[q];
        }
    });
}

int psim_server_run(psim_server* s, const psim_serve_request* requests, int32_t count, double dt, double tolerance,
                    int32_t max_iterations, psim_serve_record* records) {
    return guarded([&] {
        if (s == nullptr || (count > 0 && (requests == nullptr || records == nullptr)))
            throw std::invalid_argument("null argument");
        if (count < 0) throw std::invalid_argument("negative request count");
        if (!(dt > 0.0) || !std::isfinite(dt)) throw std::invalid_argument("time step must be finite and positive");
        if (!(tolerance >= 0.0) || max_iterations < 0) throw std::invalid_argument("invalid solver settings");
        const size_t k = s->graphs.size();
        // validate the stream: arrival order, sessions in step order
        std::unordered_map<int32_t, int32_t> open;   // session -> graph while it has a system
        for (int32_t i = 0; i < count; ++i) {
            const psim_serve_request& r = requests[i];
            if (i > 0 && r.arrival_s < requests[i - 1].arrival_s)
                throw std::invalid_argument("requests must be sorted by arrival time");
            if (r.graph < 0 || static_cast<size_t>(r.graph) >= k) throw std::invalid_argument("graph index out of range");
            if (r.temperature_old == nullptr || r.forcing_W_m2 == nullptr || r.temperature_new == nullptr)
                throw std::invalid_argument("null argument");
            const auto it = open.find(r.session);
            if (r.first) {
                if (it != open.end()) throw std::invalid_argument("a session starts twice");
                if (r.conductance_per_entry == nullptr || r.capacity_area == nullptr)
                    throw std::invalid_argument("the first request of a session needs conductance and capacity");
                open[r.session] = r.graph;
            } else {
                if (it == open.end()) throw std::invalid_argument("a session continues before it starts");
                if (it->second != r.graph) throw std::invalid_argument("a session changes graph");
            }
            if (r.last) open.erase(r.session);
        }
        if (!open.empty()) throw std::invalid_argument("a session has no last request");

        // fresh queues and completion records
        hipStream_t stream = s->ctx->stream;
        // queue heads, slot barriers and mailboxes start from zero on every launch
        int32_t slots = 0;
        for (size_t q = 0; q < k; ++q) slots += s->slot_count[q];
        psim::hip_check(hipMemsetAsync(s->d_heads, 0, k * 8 * sizeof(uint64_t), stream), "reset");
        psim::hip_check(hipMemsetAsync(s->d_words, 0,
                                       static_cast<size_t>(slots) * (psim::dev::kBarrierWordCount + kMailboxWords) *
                                           sizeof(uint32_t),
                                       stream),
                        "reset");
        psim::hip_check(hipMemsetAsync(s->d_results, 0, static_cast<size_t>(slots) * sizeof(SolveResult), stream),
                        "reset");
        for (size_t q = 0; q < k; ++q) __atomic_store_n(&s->tails[q * 8], 0ull, __ATOMIC_RELEASE);
        std::memset(s->completions, 0, static_cast<size_t>(s->completion_capacity) * sizeof(ServeCompletion));
        __atomic_store_n(s->stop, 0u, __ATOMIC_RELEASE);

        ServeParams P{};
        P.table = s->d_table;
        P.slots = s->d_slots;
        P.block_slot = s->d_block_slot;
        for (size_t q = 0; q < k; ++q) {
            P.queues[q].ring = s->rings + q * kRing;
            P.queues[q].tail = s->tails + q * 8;
            P.queues[q].head = s->d_heads + q * 8;
            P.queues[q].capacity = kRing;
        }
        P.completions = s->completions;
        P.completion_capacity = s->completion_capacity;
        P.stop = s->stop;
        const double span = count > 0 ? requests[count - 1].arrival_s : 0.0;
        P.lifetime = static_cast<uint64_t>((span + 120.0) * s->ctx->wall_clock_hz);
        psim::hip_check(psim::dev::launch_server(P, s->blocks, stream, s->waves), "server launch");
        // whatever happens below, the server is told to stop and drained
        // before this call returns
        struct Drain {
            psim_server* s;
            ~Drain() {
                __atomic_store_n(s->stop, 1u, __ATOMIC_RELEASE);
                (void)hipStreamSynchronize(s->ctx->stream);
            }
        } drain{s};
        // a run that does not complete forgets every kept matrix: an
        // assembly may have been cut short
        struct ForgetUnlessDone {
            psim_server* s;
            bool done = false;
            ~ForgetUnlessDone() {
                if (!done) std::fill(s->matrix_key.begin(), s->matrix_key.end(), uint64_t{0});
            }
        } forget{s};

        // host loop: arrivals -> admission -> queues; completions -> records
        //
        // Free pooled systems per graph: a queue in release order (a session
        // without a reusable matrix takes the system freed longest ago, which
        // evicts the coldest matrix) and, per matrix key, the free systems
        // that hold it. Entries go stale when their system is taken through
        // the other list; is_free and the system's key tell.
        std::vector<std::deque<int32_t>> free_any(k);
        std::vector<std::unordered_map<uint64_t, std::vector<int32_t>>> free_keyed(k);
        std::vector<int32_t> free_count(k, 0);
        std::vector<char> is_free(s->matrix_key.size(), 0);
        const auto release = [&](size_t q, int32_t st) {
            is_free[static_cast<size_t>(st)] = 1;
            ++free_count[q];
            free_any[q].push_back(st);
            const uint64_t key = s->matrix_key[static_cast<size_t>(st)];
            if (key != 0 && s->matrix_dt[static_cast<size_t>(st)] == dt) free_keyed[q][key].push_back(st);
        };
        const auto take_keyed = [&](size_t q, uint64_t key) -> int32_t {
            if (key == 0) return -1;
            const auto found = free_keyed[q].find(key);
            if (found == free_keyed[q].end()) return -1;
            auto& list = found->second;
            while (!list.empty()) {
                const int32_t st = list.back();
                list.pop_back();
                if (is_free[static_cast<size_t>(st)] && s->matrix_key[static_cast<size_t>(st)] == key) {
                    is_free[static_cast<size_t>(st)] = 0;
                    --free_count[q];
                    return st;
                }
            }
            return -1;
        };
        const auto take_any = [&](size_t q) -> int32_t {
            while (!free_any[q].empty()) {
                const int32_t st = free_any[q].front();
                free_any[q].pop_front();
                if (is_free[static_cast<size_t>(st)]) {
                    is_free[static_cast<size_t>(st)] = 0;
                    --free_count[q];
                    return st;
                }
            }
            return -1;
        };
        for (size_t q = 0; q < k; ++q)
            for (const int32_t st : s->pool_state[q]) release(q, st);
        std::unordered_map<int32_t, int32_t> session_state;   // session -> table index
        std::unordered_map<int32_t, bool> session_busy;
        std::vector<uint64_t> published(k, 0);
        std::vector<std::vector<int32_t>> ring_owner(k, std::vector<int32_t>(kRing, -1));
        std::vector<int32_t> free_completion;
        for (int32_t c = s->completion_capacity - 1; c >= 0; --c) free_completion.push_back(c);
        std::vector<int32_t> completion_of(static_cast<size_t>(count), -1);
        std::vector<int32_t> ring_slot_of(static_cast<size_t>(count), -1);
        std::deque<int32_t> pending;
        std::vector<int32_t> in_flight;
        // state -> system (for C A kept by the assembly)
        std::unordered_map<int32_t, psim_system*> system_of;
        for (size_t q = 0; q < k; ++q)
            for (size_t i = 0; i < s->pool[q].size(); ++i) system_of[s->pool_state[q][i]] = s->pool[q][i];

        const auto t0 = std::chrono::steady_clock::now();
        int32_t next = 0, done = 0;
        bool failed = false;
        std::string failure;
        const double give_up = span + 120.0;
        while (done < count) {
            const double now = seconds_since(t0);
            if (now > give_up) {
                failed = true;
                failure = "served requests did not complete in time";
                break;
            }
            while (next < count && requests[next].arrival_s <= now) {
                records[next] = psim_serve_record{};
                records[next].arrival_s = requests[next].arrival_s;
                pending.push_back(next++);
            }
            // admission and submission in arrival order
            for (auto it = pending.begin(); it != pending.end();) {
                const int32_t i = *it;
                const psim_serve_request& r = requests[i];
                const size_t q = static_cast<size_t>(r.graph);
                int32_t st = -1;
                if (r.first) {
                    if (free_count[q] == 0) {
                        ++it;
                        continue;   // admission: no system free in this graph's pool
                    }
                } else {
                    // the session must hold a system and have no step running
                    const auto held = session_state.find(r.session);
                    if (held == session_state.end() || session_busy[r.session]) {
                        ++it;
                        continue;
                    }
                    st = held->second;
                }
                const uint64_t position = published[q];
                if (ring_owner[q][position % kRing] >= 0 || free_completion.empty()) {
                    ++it;
                    continue;
                }
                bool assemble = false;
                if (r.first) {
                    st = take_keyed(q, r.matrix_key);
                    if (st < 0) {
                        st = take_any(q);
                        assemble = true;
                        s->matrix_key[static_cast<size_t>(st)] = r.matrix_key;
                        s->matrix_dt[static_cast<size_t>(st)] = dt;
                    }
                    session_state[r.session] = st;
                }
                session_busy[r.session] = true;
                psim_system* sys = system_of[st];
                const int op = psim::dev::kOpSolve | psim::dev::kOpHeatRhs |
                               (assemble ? psim::dev::kOpSetMatrix | psim::dev::kOpHeatAssemble : 0);
                const psim::dev::Launch lp = psim::sys::solve_params(sys, op, tolerance, max_iterations);
                ServeRequest e{};
                e.state = st;
                e.op = op;
                e.max_iterations = lp.max_iterations;
                e.max_restarts = lp.max_restarts;
                e.dt = dt;
                e.tolerance = tolerance;
                e.inner_relative = lp.inner_relative;
                e.entry_conductance = assemble ? r.conductance_per_entry : nullptr;
                e.capacity_area = assemble ? r.capacity_area : sys->capacity_area;
                e.capacity_keep = assemble ? sys->capacity_area : nullptr;
                e.t_old = r.temperature_old;
                e.forcing = r.forcing_W_m2;
                e.t_new = r.temperature_new;
                e.ticket = static_cast<uint64_t>(i);
                e.completion = free_completion.back();
                free_completion.pop_back();
                completion_of[static_cast<size_t>(i)] = e.completion;
                __atomic_store_n(&s->completions[e.completion].done, 0ull, __ATOMIC_RELAXED);
                s->rings[q * kRing + position % kRing] = e;
                ring_owner[q][position % kRing] = i;
                ring_slot_of[static_cast<size_t>(i)] = static_cast<int32_t>(position % kRing);
                published[q] = position + 1;
                __atomic_store_n(&s->tails[q * 8], published[q], __ATOMIC_RELEASE);
                records[i].submitted_s = seconds_since(t0);
                in_flight.push_back(i);
                it = pending.erase(it);
            }
            // completions
            for (size_t f = 0; f < in_flight.size();) {
                const int32_t i = in_flight[f];
                const int32_t c = completion_of[static_cast<size_t>(i)];
                ServeCompletion& rec = s->completions[c];
                if (__atomic_load_n(&rec.done, __ATOMIC_ACQUIRE) != static_cast<uint64_t>(i) + 1) {
                    ++f;
                    continue;
                }
                const psim_serve_request& r = requests[i];
                psim_serve_record& out = records[i];
                out.completed_s = seconds_since(t0);
                out.device_ms = static_cast<double>(rec.t_end - rec.t_begin) / s->ctx->wall_clock_hz * 1e3;
                out.status = rec.barrier_timeouts != 0 ? PSIM_DEVICE_ERROR : rec.status;
                out.iterations = rec.iterations;
                out.restarts = rec.restarts;
                out.v_cycles = rec.v_cycles;
                out.relative_residual = rec.relative_residual;
                if (rec.barrier_timeouts != 0) {
                    failed = true;
                    failure = "solver group barrier timed out";
                }
                const size_t q = static_cast<size_t>(r.graph);
                ring_owner[q][static_cast<size_t>(ring_slot_of[static_cast<size_t>(i)])] = -1;
                free_completion.push_back(c);
                session_busy[r.session] = false;
                if (r.last) {
                    release(q, session_state[r.session]);
                    session_state.erase(r.session);
                }
                in_flight[f] = in_flight.back();
                in_flight.pop_back();
                ++done;
            }
            if (failed) break;
        }
        __atomic_store_n(s->stop, 1u, __ATOMIC_RELEASE);
        psim::hip_check(hipStreamSynchronize(stream), "server");
        if (failed) throw psim::DeviceFailure(failure);
        for (int32_t i = 0; i < count; ++i)
            if (records[i].status != PSIM_OK && records[i].status != PSIM_NOT_CONVERGED)
                throw psim::DeviceFailure("a served request failed on the device");
        forget.done = true;
    });
}

}  // extern "C"
