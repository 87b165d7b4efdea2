// CPU model of the solver program's group barrier (group_sync and
// group_exit in planet_game_amd/csrc/solver.hip): the same barrier words and
// the same arrival, wait and exit logic, run by std::thread "workgroups" on
// std::atomic words.
//
// Barrier words (uint32): [32] workgroups that have left the launch,
// [96 + 64 l] arrivals on leaf l in this launch. Workgroup g adds 1 to leaf
// g % 8 at every barrier (episode k = 1, 2, ... of the launch) and waits
// until every leaf l holds at least k x (the number of workgroups on leaf l);
// the counters only grow within a launch. The last workgroup to leave the
// launch clears them.
//
// What the model checks (exit status 0 = all passed):
//   * every workgroup leaves every barrier, for group sizes 1..64 that change
//     from launch to launch, with 1 to 10 barriers per launch (the counters
//     must be clean at every launch's start whatever the previous group size);
//   * a value a workgroup publishes before a barrier is what every other
//     workgroup reads after it (each episode overwrites the same slots);
//   * in both forms: fenced (release fence before the arrival, acquire fence
//     after the wait) and lean (no fences: the published values and the
//     counters are relaxed atomics).
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

// One arrival of workgroup g of a group of G at its episode-th barrier
// (`episode` counts this workgroup's barriers in the launch, as grp.episode
// does); returns false on a timeout.
bool group_sync(Barrier& b, int g, int G, uint32_t& episode, bool lean) {
    if (G == 1) return true;
    const uint32_t k = ++episode;
    if (!lean) std::atomic_thread_fence(std::memory_order_release);
    b.w[96 + 64 * (g % kLeaves)].fetch_add(1, std::memory_order_relaxed);
    for (int l = 0; l < kLeaves && l < G; ++l) {
        const uint32_t target = k * static_cast<uint32_t>((G - l + kLeaves - 1) / kLeaves);
        long spins = 0;
        while (static_cast<int32_t>(b.w[96 + 64 * l].load(std::memory_order_relaxed) - target) < 0) {
            if (++spins > kSpinLimit) return false;
            std::this_thread::yield();
        }
    }
    if (!lean) std::atomic_thread_fence(std::memory_order_acquire);
    return true;
}

// Workgroup g leaves the launch; the last to leave clears the counters.
void group_exit(Barrier& b, int G) {
    if (G == 1) return;
    const uint32_t before = b.w[32].fetch_add(1, std::memory_order_relaxed);
#ifdef CLEAR_AT_FIRST_EXIT
    // a planted mistake (the tests build it to show the model catches it):
    // the first workgroup to leave clears the leaves, while others may still
    // wait on them
    if (before == 0)
        for (int l = 0; l < kLeaves; ++l) b.w[96 + 64 * l].store(0, std::memory_order_relaxed);
    if (before + 1 == static_cast<uint32_t>(G)) b.w[32].store(0, std::memory_order_relaxed);
#else
    if (before + 1 == static_cast<uint32_t>(G)) {
        for (int l = 0; l < kLeaves; ++l) b.w[96 + 64 * l].store(0, std::memory_order_relaxed);
        b.w[32].store(0, std::memory_order_relaxed);
    }
#endif
}

// One launch: G workgroups run `episodes` barriers; in each they publish a
// value, meet, and read everyone's.
bool launch(Barrier& b, std::vector<std::atomic<uint64_t>>& slots, int G, int episodes, uint64_t tag, bool lean) {
    std::atomic<bool> ok{true};
    std::vector<std::thread> threads;
    for (int g = 0; g < G; ++g)
        threads.emplace_back([&, g] {
            uint32_t episode = 0;
            for (int e = 0; e < episodes && ok.load(std::memory_order_relaxed); ++e) {
                const uint64_t mark = (tag << 32) | (static_cast<uint64_t>(e) << 8);
                slots[static_cast<size_t>(g)].store(mark | static_cast<uint64_t>(g), std::memory_order_relaxed);
                if (!group_sync(b, g, G, episode, lean)) {
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
                if (!group_sync(b, g, G, episode, lean)) {
                    std::fprintf(stderr, "timeout: G=%d episode %d workgroup %d (second)\n", G, e, g);
                    ok = false;
                    return;
                }
            }
            group_exit(b, G);
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
            // 2..10 barriers per launch (two per episode)
            const int episodes = 1 + static_cast<int>((state >> 8) % 5);
            if (!launch(b, slots, G, episodes, tag++, lean)) {
                ++failures;
                std::fprintf(stderr, "%s form failed at launch %d (G=%d)\n", lean ? "lean" : "fenced", k, G);
                break;
            }
            // now and then a launch of a single barrier: an odd count, whose
            // counters must be clean for the next launch as well
            if ((state >> 4) & 1) {
                std::vector<std::thread> threads;
                std::atomic<bool> ok{true};
                for (int g = 0; g < G; ++g)
                    threads.emplace_back([&, g] {
                        uint32_t episode = 0;
                        if (!group_sync(b, g, G, episode, lean)) ok = false;
                        group_exit(b, G);
                    });
                for (auto& t : threads) t.join();
                if (!ok) {
                    ++failures;
                    std::fprintf(stderr, "%s form: single-barrier launch timed out (G=%d)\n", lean ? "lean" : "fenced",
                                 G);
                    break;
                }
            }
        }
        std::printf("%s form: %s\n", lean ? "lean" : "fenced", failures == 0 ? "ok" : "FAILED");
    }
    return failures == 0 ? 0 : 1;
}
