// Host-thread model of a work-sharing protocol for the ensemble kernel
// (idle workgroups running 256-sink chunks of the pair passes of systems
// still running; the device side is not in the library: its first GPU run
// did not finish and was taken out, tests/test_gpu_ensemble_sharing.py):
// the slot layout, claim word and steps, with std::atomic in place of the
// device's agent-scope atomics and fences, threads in place of workgroups.  Test
// infrastructure: built and run by tests/test_share_protocol.py, also under
// ThreadSanitizer, whose happens-before checking covers the chunk outputs
// (plain memory written by helpers, read by the owner).
//
// Protocol (one slot per owner workgroup at a time):
//   claim word = seq << 32 | C << 16 | next   (C chunks, next = claimed)
//   owner  : write the pass descriptor, done = 0, then store the claim word
//            (seq + 1, C, 0) with release; claim chunks itself by CAS like a
//            helper; after next == C wait for done == C (bounded: on time-out
//            the slot is marked dead, never reused, and the pass fails);
//   helper : load a claim word (acquire); if next < C and the slot is not
//            dead, read the descriptor, CAS (seq, C, next) -> (seq, C,
//            next + k) with acq_rel; on success run chunks [next, next + k)
//            with that descriptor, then done += k (release).
// Why a successful CAS implies the descriptor read belongs to seq: the
// owner rewrites the descriptor only after done == C, i.e. after every
// claim of seq (including this one) has been made; the claim (a release
// RMW ordered after the descriptor reads) and the done increments
// (release) happen-before the owner's acquire of done == C, hence before
// its next descriptor writes.  A helper that read a newer descriptor, or a
// mix, read it after next reached C, and its CAS from next < C fails.
//
// Nobody waits on unclaimed work: the owner claims chunks too, so it only
// ever waits for chunks a helper is running.  Helpers never wait on
// anything: they exit when every owner has finished, or after an idle
// bound.
//
// Planted mistakes (compile-time, for the test that the model catches
// them): -DPLANT_NO_WAIT (the owner reads the outputs without waiting for
// done), -DPLANT_RELAXED_DONE (done += k without release), -DPLANT_NO_SEQ
// (the claim word keeps its sequence number across passes).
//
// Usage: share_model <owners> <helpers> <jobs_per_owner> <seed> [stall [n]]
//   stall = 1: one helper stalls after its first claim for longer than the
//   owner's bound; the run must still end, with that one pass reported
//   failed and every other pass correct.  n > 0: every pass has n sinks
//   (the same chunk count, which is what a claim word without its sequence
//   number cannot tell apart).
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kSlots = 8;
constexpr int kChunk = 256;   // sinks per chunk
constexpr int kMaxClaim = 4;  // chunks per claim
constexpr uint64_t kNextMask = 0xFFFF;

struct Slot {
    std::atomic<uint64_t> claim{0};
    std::atomic<uint32_t> done{0};
    std::atomic<int> owner{-1};
    std::atomic<int> dead{0};
    // pass descriptor
    std::atomic<int> job{0};       // owner-side job id (stands in for the ranges)
    std::atomic<int> n_sinks{0};
    std::atomic<int *> out{nullptr};
};

struct World {
    Slot slots[kSlots];
    std::atomic<int> idle{0};
    std::atomic<int> owners_left{0};
    std::atomic<int> stall_armed{0};
    std::chrono::milliseconds owner_bound{2000};
    std::chrono::milliseconds helper_idle_bound{3000};
    std::chrono::milliseconds stall_for{2600};
};

uint64_t word(uint32_t seq, int C, int next) {
    return (static_cast<uint64_t>(seq) << 32) | (static_cast<uint64_t>(C) << 16) |
           static_cast<uint64_t>(next);
}
int w_C(uint64_t w) { return static_cast<int>((w >> 16) & 0xFFFF); }
int w_next(uint64_t w) { return static_cast<int>(w & kNextMask); }
uint32_t w_seq(uint64_t w) { return static_cast<uint32_t>(w >> 32); }

// The value a chunk writes for sink i of job j (the "pass result").
int value(int job, int i) { return job * 7919 + i * 31 + 1; }

void jitter(std::mt19937 &rng) {
    const unsigned r = rng() % 64;
    if (r < 4) std::this_thread::sleep_for(std::chrono::microseconds(rng() % 50));
    else if (r < 16) std::this_thread::yield();
}

void run_chunks(int job, int n_sinks, int *out, int c0, int k) {
    const int b = c0 * kChunk, e = std::min(n_sinks, (c0 + k) * kChunk);
    for (int i = b; i < e; ++i) out[i] = value(job, i);
}

// Claim size: more chunks per claim while many are left (the device's
// sinks-per-lane), never more than kMaxClaim.
int claim_size(int left) { return std::max(1, std::min(kMaxClaim, left / 4)); }

// One claim attempt on slot s whose claim word was read as `cur`.  Returns
// the number of chunks claimed (0: nothing, or the CAS lost) and the first.
int try_claim(Slot &s, uint64_t cur, int &first) {
    const int C = w_C(cur), next = w_next(cur);
    if (next >= C) return 0;
    const int k = std::min(claim_size(C - next), C - next);
    uint64_t exp = cur;
    if (!s.claim.compare_exchange_strong(exp, cur + static_cast<uint64_t>(k),
                                         std::memory_order_acq_rel, std::memory_order_relaxed))
        return 0;
    first = next;
    return k;
}

struct OwnerResult {
    int passes = 0, shared = 0, failed = 0, bad = 0;
};

void owner_main(World &W, int id, int jobs, unsigned seed, int fixed_n, OwnerResult *res) {
    std::mt19937 rng(seed);
    int my = -1;
    for (int j = 0; j < jobs; ++j) {
        const int job = id * 100000 + j;
        const int n_sinks = fixed_n > 0 ? fixed_n : 1 + static_cast<int>(rng() % 2000);
        const int C = (n_sinks + kChunk - 1) / kChunk;
        int *out = new int[n_sinks];
        for (int i = 0; i < n_sinks; ++i) out[i] = 0;
        res->passes++;
        // acquire a slot once (first pass that wants to share)
        if (my < 0 && C >= 2) {
            for (int s = 0; s < kSlots && my < 0; ++s) {
                int expv = -1;
                if (W.slots[s].dead.load(std::memory_order_relaxed) == 0 &&
                    W.slots[s].owner.compare_exchange_strong(expv, id, std::memory_order_acq_rel))
                    my = s;
            }
        }
        const bool share = my >= 0 && C >= 2 && W.idle.load(std::memory_order_relaxed) > 0;
        if (!share) {
            run_chunks(job, n_sinks, out, 0, C);
        } else {
            Slot &S = W.slots[my];
            res->shared++;
#ifdef PLANT_NO_SEQ
            const uint32_t seq = w_seq(S.claim.load(std::memory_order_relaxed));
#else
            const uint32_t seq = w_seq(S.claim.load(std::memory_order_relaxed)) + 1;
#endif
            S.job.store(job, std::memory_order_relaxed);
            S.n_sinks.store(n_sinks, std::memory_order_relaxed);
            S.out.store(out, std::memory_order_relaxed);
            S.done.store(0, std::memory_order_relaxed);
            jitter(rng);
            S.claim.store(word(seq, C, 0), std::memory_order_release);
            for (;;) {
                const uint64_t cur = S.claim.load(std::memory_order_relaxed);
                if (w_next(cur) >= C) break;
                int first = 0;
                const int k = try_claim(S, cur, first);
                if (k > 0) {
                    run_chunks(job, n_sinks, out, first, k);
                    S.done.fetch_add(static_cast<uint32_t>(k), std::memory_order_release);
                }
                jitter(rng);
            }
            const auto t0 = Clock::now();
            bool ok = true;
#ifdef PLANT_NO_WAIT
            while (false) {
#else
            while (S.done.load(std::memory_order_acquire) != static_cast<uint32_t>(C)) {
#endif
                if (Clock::now() - t0 > W.owner_bound) {
                    ok = false;
                    break;
                }
                std::this_thread::yield();
            }
            if (!ok) {
                // a helper holds a chunk past the bound: the pass fails, the
                // slot is never used again, the output buffer is abandoned
                S.dead.store(1, std::memory_order_relaxed);
                my = -1;
                res->failed++;
                continue;  // `out` is leaked on purpose: a late writer may still touch it
            }
        }
        for (int i = 0; i < n_sinks; ++i)
            if (out[i] != value(job, i)) res->bad++;
        delete[] out;
        jitter(rng);
    }
    W.owners_left.fetch_sub(1, std::memory_order_acq_rel);
}

struct HelperResult {
    long claims = 0, chunks = 0;
    bool idle_exit = false;
};

void helper_main(World &W, int id, unsigned seed, bool may_stall, HelperResult *res) {
    std::mt19937 rng(seed);
    W.idle.fetch_add(1, std::memory_order_relaxed);
    auto last = Clock::now();
    int start = id % kSlots;
    for (;;) {
        if (W.owners_left.load(std::memory_order_acquire) == 0) return;
        bool got = false;
        for (int q = 0; q < kSlots && !got; ++q) {
            Slot &S = W.slots[(start + q) % kSlots];
            const uint64_t cur = S.claim.load(std::memory_order_acquire);
            if (w_next(cur) >= w_C(cur) || S.dead.load(std::memory_order_relaxed)) continue;
            const int job = S.job.load(std::memory_order_relaxed);
            const int n_sinks = S.n_sinks.load(std::memory_order_relaxed);
            int *out = S.out.load(std::memory_order_relaxed);
            jitter(rng);
            int first = 0;
            const int k = try_claim(S, cur, first);
            if (k == 0) continue;
            got = true;
            res->claims++;
            res->chunks += k;
            if (may_stall && W.stall_armed.exchange(0) == 1)
                std::this_thread::sleep_for(W.stall_for);
            run_chunks(job, n_sinks, out, first, k);
#ifdef PLANT_RELAXED_DONE
            S.done.fetch_add(static_cast<uint32_t>(k), std::memory_order_relaxed);
#else
            S.done.fetch_add(static_cast<uint32_t>(k), std::memory_order_release);
#endif
        }
        start = (start + 1) % kSlots;
        if (got) {
            last = Clock::now();
        } else {
            if (Clock::now() - last > W.helper_idle_bound) {
                res->idle_exit = true;
                return;
            }
            std::this_thread::yield();
        }
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s owners helpers jobs seed [stall]\n", argv[0]);
        return 2;
    }
    const int n_own = std::atoi(argv[1]), n_help = std::atoi(argv[2]), jobs = std::atoi(argv[3]);
    const unsigned seed = static_cast<unsigned>(std::atoi(argv[4]));
    const bool stall = argc > 5 && std::atoi(argv[5]) != 0;
    const int fixed_n = argc > 6 ? std::atoi(argv[6]) : 0;
    World W;
    W.owners_left.store(n_own);
    W.stall_armed.store(stall ? 1 : 0);
    std::vector<OwnerResult> ores(n_own);
    std::vector<HelperResult> hres(n_help);
    std::vector<std::thread> th;
    const auto t0 = Clock::now();
    for (int h = 0; h < n_help; ++h)
        th.emplace_back(helper_main, std::ref(W), h, seed * 977 + h, h == 0, &hres[h]);
    for (int o = 0; o < n_own; ++o)
        th.emplace_back(owner_main, std::ref(W), o, jobs, seed * 131 + o, fixed_n, &ores[o]);
    for (auto &t : th) t.join();
    const double secs = std::chrono::duration<double>(Clock::now() - t0).count();
    long passes = 0, shared = 0, failed = 0, bad = 0, claims = 0, chunks = 0, idle_exits = 0;
    for (auto &r : ores) {
        passes += r.passes;
        shared += r.shared;
        failed += r.failed;
        bad += r.bad;
    }
    for (auto &r : hres) {
        claims += r.claims;
        chunks += r.chunks;
        idle_exits += r.idle_exit ? 1 : 0;
    }
    std::printf("passes %ld shared %ld failed %ld bad %ld helper_claims %ld helper_chunks %ld "
                "idle_exits %ld seconds %.3f\n",
                passes, shared, failed, bad, claims, chunks, idle_exits, secs);
    // normal runs: no failure, every output correct; stall runs: exactly one
    // failed pass, every completed pass correct
    if (bad != 0) return 1;
    if (!stall && failed != 0) return 1;
    if (stall && failed != 1) return 1;
    return 0;
}
