// HOLD recursion driver (host side) over the gfx950 kernels.
//
// Reference: hold_step (src/hold.cuh:51-181), nbodysystem_globals /
// nbodysystem_buffers (src/particle_system.cuh:65-132); hold() is declared
// but never defined there (src/hold.cuh:20-35) and is defined here.
//
// Differences from the reference, by design:
//  * per-level slow/fast buffers are allocated once per level (on first
//    use) and reused; the reference allocates two full particle systems on
//    every call and never frees them (src/hold.cuh:56);
//  * the S<-S and S<-F kicks of the slow particles are one launch over two
//    source segments (slow, fast), so each slow particle's force is summed
//    once (src/hold.cuh:137-145), and the F<-S kick of the fast particles
//    rides in the same launch as a second group of sinks; the exact FP64
//    kick mode keeps the reference's three kicks so its sums follow the
//    CPU order;
//  * the only host wait per call is the count read after the split (the
//    reference blocks at the same point, src/split.cuh:430): the split
//    publishes the count in host-mapped memory and the host spins on it,
//    for a bounded time, instead of synchronising the stream;
//  * the two slow half-drifts ride on other launches: the first on the
//    split's scatter of the slow set, the second on the scatter-back (the
//    fast recursions in between never read the slow set), and both sets
//    are scattered back by one launch;
//  * a max_level guard returns HHS_ERR_DEPTH instead of recursing without
//    bound when timesteps collapse (two coincident particles with eps = 0).
#include <immintrin.h>

#include <chrono>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "hhs_internal.h"
#include "pair_math.h"

namespace hhs {

// One allocation per level holding its slow and fast sets, carved for FP64
// records and, over the same memory, for FP32 records (half the size); a
// hold_step call uses one precision throughout.
struct LevelBuffers {
    void *block = nullptr;
    Particles<double> slow{}, fast{};
    Particles<float> slow32{}, fast32{};
    template <typename T> const Particles<T> &slow_set() const;
    template <typename T> const Particles<T> &fast_set() const;
};
template <> inline const Particles<double> &LevelBuffers::slow_set<double>() const { return slow; }
template <> inline const Particles<double> &LevelBuffers::fast_set<double>() const { return fast; }
template <> inline const Particles<float> &LevelBuffers::slow_set<float>() const { return slow32; }
template <> inline const Particles<float> &LevelBuffers::fast_set<float>() const { return fast32; }

}  // namespace hhs

struct hhs_globals {
    int n_total = 0;
    double eta = 0.0, eps = 0.0, eps2 = 0.0;
    int max_level = 0;
    int *predicate = nullptr;
    int *scanned = nullptr;
    void *scan_ws = nullptr;
    size_t scan_ws_bytes = 0;
    // per level: the split's ticket (seq << 32 | raw fast count) and its
    // partition, in host-mapped pinned memory (host / device address)
    void *part_block = nullptr;
    unsigned long long *ticket_host = nullptr, *ticket_map = nullptr;
    hhs_partition *part_host = nullptr, *part_map = nullptr;
    uint32_t split_seq = 0;  // the last ticket number handed to a split
    // device words, per level: the reported fast count of the level's last
    // split, i.e. the set size of the next level's call (read by the
    // speculative launches), and the last set size seen per level (the
    // host's prediction of which launch to speculate)
    int *next_n_dev = nullptr;
    int last_n[64];
    bool speculation = true;
    bool fuse_reduce = true;  // kick partials applied by their readers
    bool chain_scatter = true;  // chained scatter-backs in one launch
    std::vector<hhs::LevelBuffers> levels;
    double simtime = 0.0;
    hhs_hold_stats stats{};
    int timestep_mode = HHS_TIMESTEP_EXACT;
    int kick_precision = -1;  // HHS_KICK_*; -1 = the process-wide setting
    int active_kick = HHS_KICK_MIXED32;  // the mode resolved at hold_step entry
    long long root_steps = 0; // root steps completed by hhs_evolve_f64
    // Calls with at most subtree_max particles run their whole subtree in one
    // device launch (0 = never).
    int subtree_max = 0;
    void *ts_ws = nullptr;  // partial minima of the symmetric timestep kernel
    size_t ts_ws_bytes = 0;
    void *sub_ws = nullptr;
    size_t sub_ws_bytes = 0;
    // subtree statistics: added up on the device by the subtree kernels,
    // fetched once at the end of the hold_step call that launched them
    hhs::SubtreeAcc *sub_acc = nullptr, *sub_acc_host = nullptr;
    bool sub_pending = false;
};

namespace hhs {
namespace {

// Carves one particle set of capacity n out of `p` (96 B per particle in
// FP64, 56 B in FP32).
template <typename T> char *carve(char *p, int n, Particles<T> &out) {
    auto take = [&](size_t bytes) {
        char *r = p;
        p += (bytes + 255) & ~static_cast<size_t>(255);
        return r;
    };
    const size_t nn = static_cast<size_t>(n);
    out.pos = reinterpret_cast<T *>(take(3 * nn * sizeof(T)));
    out.vel = reinterpret_cast<T *>(take(3 * nn * sizeof(T)));
    out.acc = reinterpret_cast<T *>(take(3 * nn * sizeof(T)));
    out.timestep = reinterpret_cast<T *>(take(nn * sizeof(T)));
    out.m = reinterpret_cast<T *>(take(nn * sizeof(T)));
    out.id = reinterpret_cast<int *>(take(nn * sizeof(int)));
    out.parent_id = reinterpret_cast<int *>(take(nn * sizeof(int)));
    out.n = n;
    return p;
}

size_t set_bytes(int n) {
    const size_t nn = static_cast<size_t>(n);
    auto r = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    return 3 * r(3 * nn * 8) + 2 * r(nn * 8) + 2 * r(nn * 4);
}

// The level table is sized once (max_level + 1 entries) at creation and
// never resized, so level descriptors stay valid across the recursion.
int ensure_level(hhs_globals *g, int level) {
    if (level >= static_cast<int>(g->levels.size())) {
        set_error("hold_step: level %d outside the level table", level);
        return HHS_ERR_DEPTH;
    }
    if (g->levels[level].block) return HHS_OK;
    LevelBuffers &lb = g->levels[level];
    const size_t bytes = 2 * set_bytes(g->n_total);
    HHS_HIP(hipMalloc(&lb.block, bytes));
    char *p = static_cast<char *>(lb.block);
    p = carve(p, g->n_total, lb.slow);
    carve(p, g->n_total, lb.fast);
    p = static_cast<char *>(lb.block);
    p = carve(p, g->n_total, lb.slow32);
    carve(p, g->n_total, lb.fast32);
    return HHS_OK;
}

// Default size limit of device-run subtrees, from the MI355X sweep over
// {0, 64, 256, 1024, 2048}: 256 was fastest at every size measured (N = 1024
// integration 0.64 -> 0.20 s, N = 2048 1.08 -> 0.35 s, N = 8192 0.59 ->
// 0.30 s, one N = 131072 root step 3.45 -> 3.02 s).  Larger subtrees lose
// because one workgroup then does work the whole GPU would do faster.
constexpr int kDefaultSubtreeMax = 256;

// The kick mode of one hold_step call: the globals' own, or the
// process-wide setting when that is -1, read once at entry so that every
// level of the recursion uses the same arithmetic.
int resolve_kick(const hhs_globals *g) {
    return g->kick_precision >= 0 ? g->kick_precision : hhs_kick_get_precision();
}

// The device subtree runs the MIXED32 kick; other kick modes keep the host
// recursion so that their arithmetic is the one requested.
bool subtree_level(const hhs_globals *g, int clevel) {
    return g->subtree_max > 0 && g->sub_ws && g->active_kick == HHS_KICK_MIXED32 &&
           g->max_level - clevel <= kSubtreeLevels - 3;
}
bool use_subtree(const hhs_globals *g, int n, int clevel) {
    return n <= g->subtree_max && subtree_level(g, clevel);
}

// Band of the split's decision guard after a timestep evaluation in the
// globals' mode (pair_math.h): approximate modes only.
double guard_band(const hhs_globals *g) {
    return g->timestep_mode == HHS_TIMESTEP_FP32   ? kGuardBandFp32
           : g->timestep_mode == HHS_TIMESTEP_FAST ? kGuardBandFast
                                                   : 0.0;
}

// The subtree runs in stream order with no host synchronisation: nothing
// the host recursion decides depends on it, and its statistics (and a
// failure) are fetched by fetch_subtree_stats at the end of the hold_step
// call.
int queue_subtree(hhs_globals *g, const Particles<double> &total, int n, int clevel, double dt,
                  bool calc, hipStream_t st, const int *n_dev = nullptr,
                  const KickPending &kick = KickPending{}) {
    if (!g->sub_pending) {
        HHS_HIP(hipMemsetAsync(g->sub_acc, 0, sizeof(SubtreeAcc), st));
        g->sub_pending = true;
    }
    return launch_subtree(total, n, clevel, dt, calc ? 1 : 0, g->eta, g->eps2, g->timestep_mode,
                          g->max_level - clevel, g->sub_ws, g->sub_ws_bytes, g->sub_acc, st,
                          n_dev, kick);
}
int run_subtree(hhs_globals *g, const Particles<double> &total, int n, int clevel, double dt,
                bool calc, hipStream_t st, const KickPending &kick) {
    if (int rc = queue_subtree(g, total, n, clevel, dt, calc, st, nullptr, kick)) return rc;
    g->simtime += dt;  // the leaves of a completed subtree tile its dt
    return HHS_OK;
}

// Adds the subtrees' statistics to the globals' and reports a subtree that
// failed (one sync, only when subtrees ran).
int fetch_subtree_stats(hhs_globals *g, hipStream_t st) {
    if (!g->sub_pending) return HHS_OK;
    g->sub_pending = false;
    HHS_HIP(hipMemcpyAsync(g->sub_acc_host, g->sub_acc, sizeof(SubtreeAcc),
                           hipMemcpyDeviceToHost, st));
    HHS_HIP(hipStreamSynchronize(st));
    const SubtreeAcc &o = *g->sub_acc_host;
    hhs_hold_stats &S = g->stats;
    S.calls += o.calls;
    S.interactions += o.interactions;
    S.ts_pairs += o.ts_pairs;
    for (int l = 0; l < 64; ++l) S.calls_per_level[l] += o.level_calls[l];
    if (o.max_level > S.max_level) S.max_level = o.max_level;
    if (o.status != HHS_OK) {
        set_error("hold_step: recursion deeper than max_level %d (timesteps collapsed?)",
                  g->max_level);
        return o.status;
    }
    return HHS_OK;
}

int alloc_subtree(hhs_globals *g, int n_max) {
    if (g->sub_ws) HHS_HIP(hipFree(g->sub_ws));
    g->sub_ws = nullptr;
    g->sub_ws_bytes = 0;
    g->subtree_max = 0;
    if (n_max <= 0) return HHS_OK;
    const int cap = n_max < g->n_total ? n_max : g->n_total;
    const size_t bytes = subtree_workspace_bytes(cap);
    if (!g->sub_acc) {
        HHS_HIP(hipMalloc(reinterpret_cast<void **>(&g->sub_acc), sizeof(SubtreeAcc)));
        HHS_HIP(hipHostMalloc(reinterpret_cast<void **>(&g->sub_acc_host), sizeof(SubtreeAcc),
                              hipHostMallocDefault));
    }
    HHS_HIP(hipMalloc(&g->sub_ws, bytes));
    g->sub_ws_bytes = bytes;
    g->subtree_max = cap;
    return HHS_OK;
}

// Timesteps of the call's set: FP64 sets in the globals' mode (symmetric
// kernel from kSymMinN particles when its workspace exists); FP32 sets in
// FP32 with the formula's IEEE operations (the reference with T = float).
// After the FP32 symmetric kernel the decision guard for pivots dt * 2^-k
// runs as a stage of the call (launch_timesteps_sym); *guard_dev then names
// the device word that tells the split whether it still has to guard.
int call_timesteps(hhs_globals *g, const Particles<double> &total, int n, double dt,
                   hipStream_t st, const int **guard_dev) {
    *guard_dev = nullptr;
    const int sym_min = g->timestep_mode == HHS_TIMESTEP_EXACT ? kSymMinNExact : kSymMinN;
    if (g->ts_ws && n >= sym_min && timesteps_sym_ws_bytes(n) <= g->ts_ws_bytes) {
        const double band = guard_band(g);
        const TsGuard gd = make_ts_guard(dt, band > 0.0 ? band : 0.0);
        const bool stage = g->timestep_mode == HHS_TIMESTEP_FP32 && band > 0.0;
        return launch_timesteps_sym(g->timestep_mode, total.pos, total.vel, total.m,
                                    total.timestep, n, g->eta, g->eps2, g->ts_ws, g->ts_ws_bytes,
                                    st, stage ? &gd : nullptr, stage ? guard_dev : nullptr);
    }
    return launch_timesteps_mode(g->timestep_mode, total.pos, total.vel, total.m, total.timestep,
                                 n, g->eta, g->eps2, st);
}
int call_timesteps(hhs_globals *g, const Particles<float> &total, int n, float, hipStream_t st,
                   const int **guard_dev) {
    *guard_dev = nullptr;
    return launch_timesteps<float>(total.pos, total.vel, total.m, total.timestep, n,
                                   static_cast<float>(g->eta), static_cast<float>(g->eps2), st);
}

// Ticket numbers of the HOLD splits: nonzero, so that a zeroed ticket word
// never matches.
uint32_t next_seq(hhs_globals *g) {
    if (++g->split_seq == 0) g->split_seq = 1;
    return g->split_seq;
}

// The count read-back of src/split.cuh:430.  The split's last workgroup
// stores the ticket word (seq << 32 | raw fast count) into host-mapped
// memory with system-scope release; the host spins on it for at most
// kTicketSpin and then blocks on the stream, which always ends.  Kernels
// still in flight on the stream need no wait: everything launched next is
// stream-ordered after them.  On MI355X the spin returns the count ~15 us
// sooner than a stream synchronisation whose wake-up follows the
// end-of-kernel signal.
constexpr auto kTicketSpin = std::chrono::microseconds(2000);

int wait_split(hhs_globals *g, int clevel, uint32_t seq, hipStream_t st, int *raw) {
    unsigned long long *w = g->ticket_host + clevel;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned i = 1;; ++i) {
        const unsigned long long v = __atomic_load_n(w, __ATOMIC_ACQUIRE);
        if (static_cast<uint32_t>(v >> 32) == seq) {
            *raw = static_cast<int>(static_cast<uint32_t>(v));
            return HHS_OK;
        }
        if ((i & 255) == 0 && std::chrono::steady_clock::now() - t0 > kTicketSpin) break;
        _mm_pause();
    }
    HHS_HIP(hipStreamSynchronize(st));
    const unsigned long long v = __atomic_load_n(w, __ATOMIC_ACQUIRE);
    if (static_cast<uint32_t>(v >> 32) != seq) {
        set_error("hold_step: the split of level %d published no count", clevel);
        return HHS_ERR_HIP;
    }
    *raw = static_cast<int>(static_cast<uint32_t>(v));
    return HHS_OK;
}

// Launches already queued for a call before the host knew its set size.
struct Queued {
    bool split = false;    // its split, whose ticket carries `seq`
    uint32_t seq = 0;
    bool subtree = false;  // its whole subtree
};

// Speculation.  The first fast recursion of a call (src/hold.cuh:116-121)
// starts with no timestep evaluation, so its first launch -- the split of
// the fast set, or the set's device subtree -- needs nothing but the set
// size, which the call's own split has just written to next_n_dev.  Queued
// before the host waits for that split, with the size read on the device,
// it runs with no host round trip in between.  Each launch checks the size
// itself and does nothing unless the host recursion would make that same
// launch for it (the subtree: size in [1, subtree_max]; the split: above
// that, or above 0 where no subtree runs), so the host only has to skip it
// later.  The launch the level's last set size predicts is queued, both when
// there is none yet.  FP64 subtrees need subtree_max <= 256 (one variant for
// every size); splits need the call's size, their bound, <= 2^18.
template <typename T>
int speculate(hhs_globals *g, const Particles<T> &set, int n_bound, int clevel, T dt,
              hipStream_t st, Queued *q) {
    *q = Queued{};
    if (!g->speculation || clevel > g->max_level || clevel >= 64) return HHS_OK;
    constexpr bool f64 = std::is_same<T, double>::value;
    const bool sub = f64 && subtree_level(g, clevel) && g->subtree_max <= 256;
    const int n_min = (f64 && subtree_level(g, clevel)) ? g->subtree_max : 0;
    const int last = g->last_n[clevel];
    const bool want_sub = sub && (last < 0 || last <= g->subtree_max);
    const bool want_split = n_bound > n_min && n_bound <= (1 << 18) && (last < 0 || last > n_min);
    const int *n_dev = g->next_n_dev + clevel;
    int rc;
    if constexpr (f64) {
        if (want_sub) {
            if ((rc = queue_subtree(g, set, g->subtree_max, clevel, dt, false, st, n_dev)))
                return rc;
            q->subtree = true;
        }
    }
    if (want_split) {
        if ((rc = ensure_level(g, clevel))) return rc;
        const LevelBuffers &lb = g->levels[clevel];
        q->seq = next_seq(g);
        rc = launch_split_hold<T>(set, lb.slow_set<T>(), lb.fast_set<T>(), dt, n_bound,
                                  g->part_map + clevel, g->ticket_map + clevel, q->seq,
                                  g->next_n_dev + clevel + 1, dt * T(0.5), g->scan_ws,
                                  g->scan_ws_bytes, st, 0.0, 0.0, 0.0, n_dev, n_min);
        if (rc) return rc;
        q->split = true;
    }
    return HHS_OK;
}

// kick (FP64 sets): total's pending kick from the parent call's split-K
// launch, when this call runs as a device subtree that applies it.
// up: this call is its parent's second fast recursion, whose scatter-back
// comes right after it; a host call then hands its own scatter-back (and
// those deferred to it) up in *up instead of launching it, and the parent
// launches the whole chain at once (one launch for the lot).
template <typename T>
int hold_rec(hhs_globals *g, const Particles<T> &total, int n, int clevel, T dt, bool calc,
             hipStream_t st, Queued queued = Queued{}, const KickPending &kick = KickPending{},
             ScatterChain<T> *up = nullptr) {
    constexpr bool f64 = std::is_same<T, double>::value;
    if (clevel > g->max_level) {
        set_error("hold_step: recursion level %d exceeds max_level %d (timesteps collapsed?)",
                  clevel, g->max_level);
        return HHS_ERR_DEPTH;
    }
    if (clevel < 64) g->last_n[clevel] = n;
    hhs_hold_stats &S = g->stats;
    if constexpr (f64) {
        if (use_subtree(g, n, clevel)) {
            if (!queued.subtree) return run_subtree(g, total, n, clevel, dt, calc, st, kick);
            g->simtime += dt;  // queued by the parent, and it ran with this n
            return HHS_OK;
        }
    }
    S.calls++;
    if (clevel < 64) S.calls_per_level[clevel]++;
    if (clevel > S.max_level) S.max_level = clevel;

    int rc;
    // src/hold.cuh:74-76
    const int *guard_dev = nullptr;
    if (calc) {
        if ((rc = call_timesteps(g, total, n, dt, st, &guard_dev))) return rc;
        S.ts_pairs += static_cast<long long>(n) * (n - 1);
    }
    if ((rc = ensure_level(g, clevel))) return rc;
    const Particles<T> slow = g->levels[clevel].slow_set<T>();  // by value: plain pointers
    const Particles<T> fast = g->levels[clevel].fast_set<T>();
    const T h = dt * T(0.5);  // the half-drift of src/hold.cuh:123-124,158-159
    // src/hold.cuh:79-83 with the first slow half-drift, then the count
    // read-back (src/split.cuh:430); FP64 sets run the decision guard after
    // a timestep evaluation in an approximate mode, so that the slow/fast
    // decisions are the exact mode's.  A split the parent queued has run
    // (or is running) already.
    uint32_t seq = queued.seq;
    if (!queued.split) {
        const double band = (f64 && calc) ? guard_band(g) : 0.0;
        seq = next_seq(g);
        rc = launch_split_hold<T>(total, slow, fast, dt, n, g->part_map + clevel,
                                  g->ticket_map + clevel, seq, g->next_n_dev + clevel + 1, h,
                                  g->scan_ws, g->scan_ws_bytes, st, band, g->eta, g->eps2,
                                  nullptr, 0, guard_dev);
        if (rc) return rc;
    }
    Queued child;
    if ((rc = speculate<T>(g, fast, n, clevel + 1, dt / T(2), st, &child))) return rc;
    int raw = 0;
    if ((rc = wait_split(g, clevel, seq, st, &raw))) return rc;
    // a lone fast particle stays slow (src/split.cuh:321-324): the slow set
    // becomes the whole set, drifted
    const int nf = raw == 1 ? 0 : raw;
    const int ns = n - nf;
    if (raw == 1 && (rc = launch_lone_fast_fixup<T>(total, slow, n, h, st))) return rc;
    KickDefer kdefer;  // the kick's partials left to their readers (below)

    // src/hold.cuh:106-114
    if (nf == 0) g->simtime += static_cast<double>(dt);
    // src/hold.cuh:116-121
    if (nf > 0 && (rc = hold_rec<T>(g, fast, nf, clevel + 1, dt / T(2), false, st, child)))
        return rc;
    if (ns > 0) {
        // src/hold.cuh:123-124: done by the split
        // src/hold.cuh:137-145: slow sinks <- {slow, fast}; fast sinks <- slow.
        // In the exact FP64 mode the S<-S and S<-F kicks are two velocity
        // updates, as in the reference, so the sums match the CPU order.
        // Otherwise one launch kicks both: the fast sinks ride along as the
        // kick's second sink group, fed by the slow segment only.  FP32 sets
        // always use the MIXED32 arithmetic.
        const int prec = f64 ? g->active_kick : HHS_KICK_MIXED32;
        const bool exact = prec == HHS_KICK_F64_EXACT;
        KickArgs<T> ks{slow.pos, slow.vel, slow.acc, ns,  slow.pos, slow.m, ns,      fast.pos,
                       fast.m,   exact ? 0 : nf,      dt, g->eps2,  prec};
        if (!exact && nf > 0) {
            ks.sb_pos = fast.pos;
            ks.sb_vel = fast.vel;
            ks.sb_acc = fast.acc;
            ks.n_sink_b = nf;
        }
        // When the second fast recursion is a device subtree (or there is
        // none), no host kick runs before this call's scatter-back, so the
        // kick's split-K partials may wait for their readers: the fast
        // sinks' sum is applied by that subtree as it loads the set, the
        // slow sinks' by the scatter-back (one reduce launch fewer).
        bool defer = false;
        if constexpr (f64)
            defer = !exact && g->fuse_reduce && (nf == 0 || use_subtree(g, nf, clevel + 1));
        if ((rc = launch_kick<T>(ks, st, defer ? &kdefer : nullptr))) return rc;
        if (exact && nf > 0) {
            KickArgs<T> ksf{slow.pos, slow.vel, slow.acc, ns, fast.pos, fast.m, nf, nullptr,
                            nullptr,  0,        dt,       g->eps2,  prec};
            if ((rc = launch_kick<T>(ksf, st))) return rc;
            KickArgs<T> kf{fast.pos, fast.vel, fast.acc, nf, slow.pos, slow.m, ns, nullptr,
                           nullptr,  0,        dt,       g->eps2,  prec};
            if ((rc = launch_kick<T>(kf, st))) return rc;
        }
        S.interactions += static_cast<long long>(ns) * (ns + nf) + static_cast<long long>(nf) * ns;
        // src/hold.cuh:158-159: done by the scatter-back below
    }
    // src/hold.cuh:168-173
    ScatterChain<T> below;  // scatter-backs the second recursion left to this call
    if (nf > 0 && (rc = hold_rec<T>(g, fast, nf, clevel + 1, dt / T(2), true, st, Queued{},
                                    kdefer.b, g->chain_scatter ? &below : nullptr)))
        return rc;
    // src/hold.cuh:177-178 (the deeper levels reuse the level's sets, so
    // scatter back now), with the second slow half-drift
    if (below.depth == 0 && !up)
        return launch_update_pair<T>(total, fast, nf, slow, ns, h, st, kdefer.a);
    ScatterChain<T> ch;
    ch.lv[0] = ScatterLevel<T>{slow, ns, h, kdefer.a, fast.parent_id, nf};
    ch.bottom_fast = fast;
    ch.depth = 1;
    if (below.depth > 0) {
        for (int k = 0; k < below.depth; ++k) ch.lv[1 + k] = below.lv[k];
        ch.bottom_fast = below.bottom_fast;
        ch.depth = 1 + below.depth;
    }
    // the parent prepends its own level, so there must be room for it
    if (up && ch.depth < kScatterChainMax) {
        *up = ch;
        return HHS_OK;
    }
    if (ch.depth == 1) return launch_update_pair<T>(total, fast, nf, slow, ns, h, st, kdefer.a);
    return launch_update_chain<T>(total, ch, st);
}

// hold_step on one particle set of either precision (the C entry points).
template <typename T, typename P>
int hold_step_c(hhs_globals *g, P *total, int n, int clevel, T dt, int calc_timestep,
                void *stream) {
    HHS_REQUIRE(g && total, "hold_step: null argument");
    HHS_REQUIRE(n > 0 && n <= total->n && n <= g->n_total,
                "hold_step: n=%d outside [1, min(%d, %d)]", n, total->n, g->n_total);
    HHS_REQUIRE(clevel >= 0, "hold_step: negative level");
    HHS_REQUIRE(dt > T(0), "hold_step: dt must be positive");
    HHS_REQUIRE(total->pos && total->vel && total->m && total->timestep,
                "hold_step: null field in total");
    hipStream_t st = as_stream(stream);
    g->active_kick = resolve_kick(g);
    const int rc = hold_rec<T>(g, from_c(*total), n, clevel, dt, calc_timestep != 0, st);
    const int rs = fetch_subtree_stats(g, st);
    return rc != HHS_OK ? rc : rs;
}

}  // namespace
}  // namespace hhs

using namespace hhs;

extern "C" {

int hhs_globals_create(hhs_globals **out, int n_total, double eta, double eps, int max_level) {
    HHS_REQUIRE(out, "globals_create: null output");
    *out = nullptr;
    HHS_REQUIRE(n_total > 0, "globals_create: n_total must be positive");
    HHS_REQUIRE(eta > 0.0, "globals_create: eta must be positive");
    HHS_REQUIRE(eps >= 0.0, "globals_create: eps must be >= 0");
    if (max_level <= 0) max_level = 48;
    HHS_REQUIRE(max_level < 64, "globals_create: max_level must be < 64");
    hhs_globals *g = new (std::nothrow) hhs_globals();
    if (!g) {
        set_error("globals_create: out of host memory");
        return HHS_ERR_ALLOC;
    }
    g->n_total = n_total;
    g->eta = eta;
    g->eps = eps;
    g->eps2 = eps * eps;
    g->max_level = max_level;
    g->levels.resize(static_cast<size_t>(max_level) + 1);
    g->scan_ws_bytes = scan_ws_bytes(n_total);
    hipError_t e = hipMalloc(&g->predicate, sizeof(int) * static_cast<size_t>(n_total));
    if (e == hipSuccess) e = hipMalloc(&g->scanned, sizeof(int) * static_cast<size_t>(n_total));
    if (e == hipSuccess) e = hipMalloc(&g->scan_ws, g->scan_ws_bytes);
    const size_t levels = static_cast<size_t>(max_level) + 1;
    if (e == hipSuccess)
        e = hipHostMalloc(&g->part_block,
                          (sizeof(unsigned long long) + sizeof(hhs_partition)) * levels,
                          hipHostMallocMapped | hipHostMallocCoherent);
    void *map = nullptr;
    if (e == hipSuccess) e = hipHostGetDevicePointer(&map, g->part_block, 0);
    if (e == hipSuccess) {
        g->ticket_host = static_cast<unsigned long long *>(g->part_block);
        g->ticket_map = static_cast<unsigned long long *>(map);
        g->part_host = reinterpret_cast<hhs_partition *>(g->ticket_host + levels);
        g->part_map = reinterpret_cast<hhs_partition *>(g->ticket_map + levels);
        for (size_t l = 0; l < levels; ++l) g->ticket_host[l] = 0;
    }
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&g->next_n_dev), sizeof(int) * (levels + 1));
    if (e == hipSuccess) e = hipMemset(g->next_n_dev, 0, sizeof(int) * (levels + 1));
    for (int &v : g->last_n) v = -1;
    if (e == hipSuccess) e = hipMemset(g->scan_ws, 0, g->scan_ws_bytes);
    if (e != hipSuccess) {
        int rc = hip_status(e, "globals_create", __FILE__, __LINE__);
        hhs_globals_destroy(g);
        return rc;
    }
    if (int rc = alloc_subtree(g, kDefaultSubtreeMax)) {
        hhs_globals_destroy(g);
        return rc;
    }
    // Symmetric timesteps from kSymMinN particles, when the records and
    // partial minima (358 MB at n = 131072, n^2 / 48 bytes) stay within
    // 8 GiB (n_total up to ~640k).
    const size_t tsb = timesteps_sym_ws_bytes(n_total);
    if (n_total >= kSymMinN && tsb <= (size_t{8} << 30)) {
        if (hipMalloc(&g->ts_ws, tsb) != hipSuccess) {
            int rc = hip_status(hipErrorOutOfMemory, "globals_create", __FILE__, __LINE__);
            hhs_globals_destroy(g);
            return rc;
        }
        g->ts_ws_bytes = tsb;
    }
    *out = g;
    return HHS_OK;
}

int hhs_globals_destroy(hhs_globals *g) {
    if (!g) return HHS_OK;
    for (auto &lb : g->levels)
        if (lb.block) (void)hipFree(lb.block);
    if (g->predicate) (void)hipFree(g->predicate);
    if (g->scanned) (void)hipFree(g->scanned);
    if (g->scan_ws) (void)hipFree(g->scan_ws);
    if (g->next_n_dev) (void)hipFree(g->next_n_dev);
    if (g->part_block) (void)hipHostFree(g->part_block);
    if (g->sub_ws) (void)hipFree(g->sub_ws);
    if (g->ts_ws) (void)hipFree(g->ts_ws);
    if (g->sub_acc) (void)hipFree(g->sub_acc);
    if (g->sub_acc_host) (void)hipHostFree(g->sub_acc_host);
    delete g;
    return HHS_OK;
}

int *hhs_globals_predicate(hhs_globals *g) { return g ? g->predicate : nullptr; }
int *hhs_globals_scanned_predicate(hhs_globals *g) { return g ? g->scanned : nullptr; }
double hhs_globals_simtime(const hhs_globals *g) { return g ? g->simtime : 0.0; }
void hhs_globals_set_simtime(hhs_globals *g, double t) {
    if (g) g->simtime = t;
}
double hhs_globals_eta(const hhs_globals *g) { return g ? g->eta : 0.0; }
double hhs_globals_eps(const hhs_globals *g) { return g ? g->eps : 0.0; }

int hhs_globals_set_timestep_mode(hhs_globals *g, int mode) {
    HHS_REQUIRE(g, "set_timestep_mode: null globals");
    HHS_REQUIRE(mode == HHS_TIMESTEP_EXACT || mode == HHS_TIMESTEP_FAST ||
                    mode == HHS_TIMESTEP_FP32,
                "set_timestep_mode: mode must be HHS_TIMESTEP_EXACT, _FAST or _FP32");
    g->timestep_mode = mode;
    return HHS_OK;
}

int hhs_globals_set_kick_precision(hhs_globals *g, int mode) {
    HHS_REQUIRE(g, "set_kick_precision: null globals");
    HHS_REQUIRE(mode >= -1 && mode <= HHS_KICK_F64_EXACT,
                "set_kick_precision: mode must be -1 or an HHS_KICK_* value");
    g->kick_precision = mode;
    return HHS_OK;
}

int hhs_globals_set_subtree_max(hhs_globals *g, int n_max) {
    HHS_REQUIRE(g, "set_subtree_max: null globals");
    HHS_REQUIRE(n_max >= 0, "set_subtree_max: n_max must be >= 0");
    HHS_HIP(hipDeviceSynchronize());  // the old workspace may still be in use
    return alloc_subtree(g, n_max);
}
int hhs_globals_subtree_max(const hhs_globals *g) { return g ? g->subtree_max : 0; }
int hhs_globals_set_speculation(hhs_globals *g, int on) {
    HHS_REQUIRE(g, "set_speculation: null globals");
    g->speculation = on != 0;
    return HHS_OK;
}

int hhs_globals_set_chain_scatter(hhs_globals *g, int on) {
    HHS_REQUIRE(g, "set_chain_scatter: null globals");
    g->chain_scatter = on != 0;
    return HHS_OK;
}

int hhs_globals_set_fused_reduce(hhs_globals *g, int on) {
    HHS_REQUIRE(g, "set_fused_reduce: null globals");
    g->fuse_reduce = on != 0;
    return HHS_OK;
}

int hhs_globals_timestep_mode(const hhs_globals *g) { return g ? g->timestep_mode : -1; }
int hhs_globals_kick_precision(const hhs_globals *g) { return g ? g->kick_precision : -1; }
long long hhs_globals_root_steps(const hhs_globals *g) { return g ? g->root_steps : 0; }
void hhs_globals_set_root_steps(hhs_globals *g, long long k) {
    if (g) g->root_steps = k;
}

int hhs_globals_stats(const hhs_globals *g, hhs_hold_stats *out) {
    HHS_REQUIRE(g && out, "globals_stats: null argument");
    *out = g->stats;
    return HHS_OK;
}

int hhs_globals_reset_stats(hhs_globals *g) {
    HHS_REQUIRE(g, "globals_reset_stats: null globals");
    std::memset(&g->stats, 0, sizeof(g->stats));
    return HHS_OK;
}

int hhs_hold_step_f64(hhs_globals *g, hhs_particles_f64 *total, int n, int clevel,
                      double stime, double etime, double dt, int calc_timestep, void *stream) {
    (void)stime;
    (void)etime;  // the system is autonomous; the reference never reads them
    return hold_step_c<double>(g, total, n, clevel, dt, calc_timestep, stream);
}

int hhs_hold_step_f32(hhs_globals *g, hhs_particles_f32 *total, int n, int clevel, float stime,
                      float etime, float dt, int calc_timestep, void *stream) {
    (void)stime;
    (void)etime;
    return hold_step_c<float>(g, total, n, clevel, dt, calc_timestep, stream);
}

int hhs_hold_f32(hhs_globals *g, hhs_particles_f32 *sys, float t_end, float dt_root,
                 void *stream) {
    HHS_REQUIRE(g && sys, "hold: null argument");
    HHS_REQUIRE(dt_root > 0.0f, "hold: dt_root must be positive");
    float t = 0.0f;
    while (t < t_end) {
        int rc = hhs_hold_step_f32(g, sys, sys->n, 0, t, t + dt_root, dt_root, 1, stream);
        if (rc) return rc;
        t += dt_root;
    }
    return HHS_OK;
}

int hhs_hold_f64(hhs_globals *g, hhs_particles_f64 *sys, double t_end, double dt_root,
                 void *stream) {
    HHS_REQUIRE(g && sys, "hold: null argument");
    HHS_REQUIRE(dt_root > 0.0, "hold: dt_root must be positive");
    double t = 0.0;
    while (t < t_end) {
        int rc = hhs_hold_step_f64(g, sys, sys->n, 0, t, t + dt_root, dt_root, 1, stream);
        if (rc) return rc;
        t += dt_root;
    }
    return HHS_OK;
}

}  // extern "C"
