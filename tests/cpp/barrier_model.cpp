// CPU model of the solver program's group barrier (group_sync in
// planet_game_amd/csrc/solver.hip): the same barrier words and the same
// arrival logic, run by std::thread "workgroups" on std::atomic words.
//
// Barrier words (uint32): [0] generation, [32 + 32 p] root arrivals of
// parity p, [96 + 64 l + 32 p] arrivals of leaf l for parity p; workgroup g
// arrives on leaf g % 8, the last arrival of a leaf arrives at the root, the
// last root arrival clears the whole idle set (root and every leaf of the
// other parity) and then moves the generation that every waiter polls.
//
// What the model checks (exit status 0 = all passed):
//   * every workgroup leaves every barrier, for group sizes 1..64 that change
//     from launch to launch, with odd and even numbers of barriers per
//     launch (the idle set must be clean whatever the previous group size);
//   * a value a workgroup publishes before a barrier is what every other
//     workgroup reads after it (each episode overwrites the same slots);
//   * in both forms: fenced (release fence before the arrival, acquire fence
//     after the wait, acq_rel fences between the counters a last arrival
//     observes and updates) and lean (no fences: the published values and
//     the counters are relaxed atomics).
// The lean form's ordering on the GPU rests on the hardware (a thread's
// s_waitcnt vmcnt(0) returns once its write-through stores are acknowledged,
// and loads that bypass the L2 read memory); a CPU model cannot check that
// part. It checks the counting protocol, which both forms share.
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

namespace {

constexpr int kLeaves = 8;
constexpr int kWords = 96 + 64 * kLeaves;
constexpr long kSpinLimit = 2000000;   // polls before a wait counts as timed out

struct Barrier {
    std::atomic<uint32_t> w[kWords];
    Barrier() {
        for (auto& x : w) x.store(0, std::memory_order_relaxed);
    }
};

// One arrival of workgroup g of a group of G; returns false on a timeout.
bool group_sync(Barrier& b, int g, int G, bool lean) {
    if (G == 1) return true;
    const int leaf = g % kLeaves;
    const int leaves = G < kLeaves ? G : kLeaves;
    const uint32_t leaf_size = static_cast<uint32_t>((G - leaf + kLeaves - 1) / kLeaves);
    if (!lean) std::atomic_thread_fence(std::memory_order_release);
    const uint32_t seen = b.w[0].load(std::memory_order_relaxed);
    const uint32_t p = seen & 1u;
    std::atomic<uint32_t>& root = b.w[32 + 32 * p];
    std::atomic<uint32_t>& mine = b.w[96 + 64 * leaf + 32 * p];
    if (mine.fetch_add(1, std::memory_order_relaxed) + 1 == leaf_size) {
        if (!lean) std::atomic_thread_fence(std::memory_order_acq_rel);
        if (root.fetch_add(1, std::memory_order_relaxed) + 1 == static_cast<uint32_t>(leaves)) {
            b.w[32 + 32 * (p ^ 1u)].store(0, std::memory_order_relaxed);
#ifdef CLEAR_OWN_LEAF_ONLY
            // a planted mistake (the tests build it to show the model catches
            // it): only this leaf's idle counter is cleared
            b.w[96 + 64 * leaf + 32 * (p ^ 1u)].store(0, std::memory_order_relaxed);
#else
            for (int l = 0; l < kLeaves; ++l) b.w[96 + 64 * l + 32 * (p ^ 1u)].store(0, std::memory_order_relaxed);
#endif
            // the clears complete before the generation moves (lean: the
            // thread's own wait for its stores; fenced: the acq_rel fence)
            std::atomic_thread_fence(lean ? std::memory_order_seq_cst : std::memory_order_acq_rel);
            b.w[0].store(seen + 1, std::memory_order_relaxed);
            return true;
        }
    }
    long spins = 0;
    while (b.w[0].load(std::memory_order_relaxed) == seen) {
        if (++spins > kSpinLimit) return false;
        std::this_thread::yield();
    }
    if (!lean) std::atomic_thread_fence(std::memory_order_acquire);
    return true;
}

// One launch: G workgroups run `episodes` barriers; in each they publish a
// value, meet, and read everyone's.
bool launch(Barrier& b, std::vector<std::atomic<uint64_t>>& slots, int G, int episodes, uint64_t tag, bool lean) {
    std::atomic<bool> ok{true};
    std::vector<std::thread> threads;
    for (int g = 0; g < G; ++g)
        threads.emplace_back([&, g] {
            for (int e = 0; e < episodes && ok.load(std::memory_order_relaxed); ++e) {
                const uint64_t mark = (tag << 32) | (static_cast<uint64_t>(e) << 8);
                slots[static_cast<size_t>(g)].store(mark | static_cast<uint64_t>(g), std::memory_order_relaxed);
                if (!group_sync(b, g, G, lean)) {
                    std::fprintf(stderr, "timeout: G=%d episode %d workgroup %d\n", G, e, g);
                    ok = false;
                    return;
                }
                for (int h = 0; h < G; ++h) {
                    const uint64_t v = slots[static_cast<size_t>(h)].load(std::memory_order_relaxed);
                    if (v != (mark | static_cast<uint64_t>(h))) {
                        std::fprintf(stderr, "stale read: G=%d episode %d reader %d slot %d\n", G, e, g, h);
                        ok = false;
                        return;
                    }
                }
                // the second barrier of the episode keeps the next episode's
                // writes after every read of this one
                if (!group_sync(b, g, G, lean)) {
                    std::fprintf(stderr, "timeout: G=%d episode %d workgroup %d (second)\n", G, e, g);
                    ok = false;
                    return;
                }
            }
        });
    for (auto& t : threads) t.join();
    return ok.load();
}

}  // namespace

int main(int argc, char** argv) {
    const int launches = argc > 1 ? std::atoi(argv[1]) : 60;
    std::vector<std::atomic<uint64_t>> slots(64);
    int failures = 0;
    for (const bool lean : {false, true}) {
        Barrier b;   // one barrier-word block, reused by every launch (as a system's)
        uint64_t tag = 1;
        unsigned state = lean ? 12345u : 777u;
        for (int k = 0; k < launches; ++k) {
            state = state * 1103515245u + 12345u;
            const int G = 1 + static_cast<int>((state >> 16) % 64);
            // odd episodes: an odd number of barriers, so the next launch
            // starts on the other parity
            const int episodes = 1 + static_cast<int>((state >> 8) % 5);
            if (!launch(b, slots, G, episodes, tag++, lean)) {
                ++failures;
                std::fprintf(stderr, "%s form failed at launch %d (G=%d)\n", lean ? "lean" : "fenced", k, G);
                break;
            }
            // an extra single barrier flips the parity for the next launch
            if ((state >> 4) & 1) {
                std::vector<std::thread> threads;
                std::atomic<bool> ok{true};
                for (int g = 0; g < G; ++g)
                    threads.emplace_back([&, g] {
                        if (!group_sync(b, g, G, lean)) ok = false;
                    });
                for (auto& t : threads) t.join();
                if (!ok) {
                    ++failures;
                    std::fprintf(stderr, "%s form: parity flip timed out (G=%d)\n", lean ? "lean" : "fenced", G);
                    break;
                }
            }
        }
        std::printf("%s form: %s\n", lean ? "lean" : "fenced", failures == 0 ? "ok" : "FAILED");
    }
    return failures == 0 ? 0 : 1;
}
