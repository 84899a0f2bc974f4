// Internal declarations shared by the libhhs translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "hhs.h"

namespace hhs {

// Thread-local last-error message (hhs_last_error()).
void set_error(const char *fmt, ...);
void clear_error();

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Per-thread device scratch reused across launches, grown on demand.  A
// launch on another stream than the previous user, or one that regrows it,
// first waits for the device to drain, so one host thread's launches on
// several streams never use it concurrently; launches on one stream pay
// nothing.  Usage: acquire, queue the kernels on `st`, release(st).
class ThreadScratch {
  public:
    ~ThreadScratch();
    int acquire(size_t bytes, hipStream_t st, void **out);
    int release(hipStream_t st);

  private:
    void *ptr_ = nullptr;
    size_t bytes_ = 0;
    hipStream_t stream_ = nullptr;
    bool used_ = false;
};

// Returns HHS_ERR_HIP (and records the message) when `e` is not hipSuccess.
int hip_status(hipError_t e, const char *what, const char *file, int line);

#define HHS_HIP(call)                                                         \
    do {                                                                      \
        hipError_t e__ = (call);                                              \
        if (e__ != hipSuccess)                                                \
            return ::hhs::hip_status(e__, #call, __FILE__, __LINE__);         \
    } while (0)

// Checks the launch that was just issued.
#define HHS_LAUNCH_CHECK(name)                                                \
    do {                                                                      \
        hipError_t e__ = hipGetLastError();                                   \
        if (e__ != hipSuccess)                                                \
            return ::hhs::hip_status(e__, name, __FILE__, __LINE__);          \
    } while (0)

#define HHS_REQUIRE(cond, ...)                                                \
    do {                                                                      \
        if (!(cond)) {                                                        \
            ::hhs::set_error(__VA_ARGS__);                                    \
            return HHS_ERR_INVALID;                                           \
        }                                                                     \
    } while (0)

// ------------------------------------------------------------ particle sets
template <typename T> struct Particles {
    T *pos, *vel, *acc, *timestep, *m;
    int *id, *parent_id;
    int n;
};

inline Particles<double> from_c(const hhs_particles_f64 &p) {
    return {p.pos, p.vel, p.acc, p.timestep, p.m, p.id, p.parent_id, p.n};
}
inline Particles<float> from_c(const hhs_particles_f32 &p) {
    return {p.pos, p.vel, p.acc, p.timestep, p.m, p.id, p.parent_id, p.n};
}

// ------------------------------------------------------------------- kick
// Sinks [0, n_sink) of `si` kicked by up to two source segments.  Source
// segment k is (pos, m, n); a segment with n == 0 is skipped.
template <typename T> struct KickArgs {
    const T *si_pos;
    T *si_vel;
    T *si_acc;    // may be null
    int n_sink;
    const T *s0_pos, *s0_m;
    int n0;
    const T *s1_pos, *s1_m;
    int n1;
    T h;          // velocity update factor: vel += acc * h   (h = fac * dt)
    double eps2;
    int precision = -1;  // HHS_KICK_*; -1 = the process-wide setting
    // Optional second group of sinks in the same launch, kicked by segment 0
    // only (HOLD: the fast sinks from the slow sources beside the slow sinks
    // from both).  Not in the exact FP64 mode.
    const T *sb_pos = nullptr;
    T *sb_vel = nullptr;
    T *sb_acc = nullptr;
    int n_sink_b = 0;
};

struct KickConfig {
    int sinks_per_lane;  // 1, 2, 4
    int waves;           // 1, 2, 4, 8
};

KickConfig kick_pick_config(int n_sink, int n_src);
// A kick's chunk partials whose sum the next reader applies instead of a
// reduce launch: rows [k][i][3] FP64 partial accelerations of K chunks,
// added in chunk order from 0.0 and applied as vel = vel + T(a) * h with
// no contraction (kick_reduce_kernel's arithmetic), so the bits are those
// of the reduce.  part null: nothing pending.
struct KickPending {
    const double *part = nullptr;
    int K = 0;
    int rows = 0;
    double h = 0.0;
};

// Row i's sum over the K chunk partials part[k][rows][3], added in chunk
// order from 0.0 (the split-K reduce's arithmetic, so the same bits as a
// plain loop).  Eight chunks' loads are issued before their adds: the plain
// loop waited for each chunk's load in turn, one memory round trip per
// chunk.  K is uniform over the wave.
__device__ __forceinline__ void sum_kick_partials(const double *__restrict__ part, int K,
                                                  int rows, int i, double &ax, double &ay,
                                                  double &az) {
    constexpr int B = 8;
    ax = 0.0;
    ay = 0.0;
    az = 0.0;
    for (int k = 0; k < K; k += B) {
        double v[B][3];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (k + j < K) {
                const double *q = part + 3 * (static_cast<size_t>(k + j) * rows + i);
                v[j][0] = q[0];
                v[j][1] = q[1];
                v[j][2] = q[2];
            }
        }
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (k + j < K) {
                ax += v[j][0];
                ay += v[j][1];
                az += v[j][2];
            }
        }
    }
}
// defer non-null: a split-K launch leaves its reduce to the caller, which
// gets each sink group's pending partials (a: the first group, b: the
// second) and must have them applied before anything reads those
// velocities, and before the next kick launch on this host thread reuses
// the partial buffer.
struct KickDefer {
    KickPending a, b;
};
template <typename T>
int launch_kick(const KickArgs<T> &a, hipStream_t s, KickDefer *defer = nullptr);

// -------------------------------------------------------------- timesteps
template <typename T>
int launch_timesteps(const T *pos, const T *vel, const T *m, T *ts, int n, T eta, T eps2,
                     hipStream_t s);
// Autotuned one-sided timestep configurations (timestep.hip).
void ts_tuned_install(int mode, int cls, int w, int k);
void ts_tuned_clear();
bool ts_tuned_valid(int mode, int cls, int w, int k);
int launch_timesteps_mode(int mode, const double *pos, const double *vel, const double *m,
                          double *ts, int n, double eta, double eps2, hipStream_t s);
// Symmetric evaluation (each unordered pair once) for the EXACT and FP32
// modes (FAST falls back to the one-sided kernel); ws holds
// timesteps_sym_ws_bytes(n) bytes of workspace.  HOLD uses it from kSymMinN
// particles up, where it beats the one-sided kernel (MI355X, FP32 on an
// equal-mass Plummer sphere, us, symmetric with 64-source units vs
// one-sided: n = 6144 62 vs 62, 8192 76 vs 101, 12288 141 vs 215, 16384 215
// vs 358, 32768 688 vs 1263, 131072 10240 vs 19443; EXACT: n = 32768 6.02
// vs 6.21 ms, 65536 17.1 vs 22.4, 131072 56.2 vs 92.4).
constexpr int kSymMinN = 8192;        // FP32 mode
constexpr int kSymMinNExact = 32768;  // exact mode
size_t timesteps_sym_ws_bytes(int n);
// guard (FP32 mode, with guard_overflow): the decision guard's exact
// recomputation of the timesteps within its band of a pivot runs as a stage
// of this call on the whole GPU; *guard_overflow then names a device word
// that is nonzero when the stage could not take them all, in which case the
// split's own guard must run (launch_split_hold's guard_dev).
struct TsGuard;
int launch_timesteps_sym(int mode, const double *pos, const double *vel, const double *m,
                         double *ts, int n, double eta, double eps2, void *ws, size_t ws_bytes,
                         hipStream_t s, const TsGuard *guard = nullptr,
                         const int **guard_overflow = nullptr);

// ----------------------------------------------------------- scan / split
// Workspace layout: u32 header[16] ([0] look-back epoch, [1] its completion
// counter, [2] raw fast count of the last split), then one status word per
// 1024-element tile of the look-back scan (and, for the reduce-then-scan of
// in-place scans, one int count per chunk after them).
constexpr size_t kScanHeaderBytes = 64;
size_t scan_ws_bytes(int n_max);
int scan_ws_init(void *ws, size_t bytes, hipStream_t s);

// lookback_ws false: a workspace that only the reduce-then-scan uses (the
// allocation-free split's fixed scratch).
template <typename T>
int launch_split(const Particles<T> &total, const Particles<T> &slow, const Particles<T> &fast,
                 int *predicate, int *scanned, T pivot, int n, hhs_partition *part_dev,
                 void *ws, size_t ws_bytes, hipStream_t s, bool lookback_ws = true);

// ------------------------------------------------------------ drift/update
template <typename T> int launch_drift(T *pos, const T *vel, int n, T h, hipStream_t s);
// HOLD-internal fused forms (hold.hip): the split that also half-drifts the
// slow set and reports the raw fast count, its lone-particle follow-up, and
// the scatter-back of both sets with the second slow half-drift.
// guard_band > 0 runs the decision guard of the approximate timestep modes
// (pair_math.h) on the set's timesteps before the predicates are formed;
// with guard_dev, only when that device word is nonzero (the timestep call
// ran the guard itself, launch_timesteps_sym).
// fast_count (optional device word) receives the reported fast count.
// n_dev non-null: a speculative split whose count is read on the device
// (n is then an upper bound, at most 2^18) and which runs only when that
// count exceeds n_min.
template <typename T>
int launch_split_hold(const Particles<T> &total, const Particles<T> &slow,
                      const Particles<T> &fast, T pivot, int n, hhs_partition *part,
                      unsigned long long *ticket, uint32_t seq, int *fast_count, T h_slow,
                      void *ws, size_t ws_bytes, hipStream_t st, double guard_band = 0.0,
                      double eta = 0.0, double eps2 = 0.0, const int *n_dev = nullptr,
                      int n_min = 0, const int *guard_dev = nullptr);
template <typename T>
int launch_lone_fast_fixup(const Particles<T> &total, const Particles<T> &slow, int n, T h_slow,
                           hipStream_t st);
// slow_kick: the slow set's pending kick (KickDefer), applied to its
// velocities before the half-drift.
template <typename T>
int launch_update_pair(const Particles<T> &prev, const Particles<T> &fast, int nf,
                       const Particles<T> &slow, int ns, T h_slow, hipStream_t s,
                       const KickPending &slow_kick = KickPending{});
template <typename T>
int launch_update(const Particles<T> &prev, const Particles<T> &cur, int n, hipStream_t s);
// Scatter-backs of a chain of HOLD calls, each the second fast recursion of
// the one above it, in one launch.  Level 0 is the top call, whose own set
// receives every element; level k's set is level k-1's fast set.  Each
// level's slow records are drifted by its h after its pending kick, as its
// own scatter-back would; the deepest level's fast records are copied.  An
// element of level k's slow set goes to lv[k].slow.parent_id, then up
// through lv[k-1 .. 0].fast_pid.  The same bits as the calls' scatter-backs
// one after the other, deepest first.
constexpr int kScatterChainMax = 8;
template <typename T> struct ScatterLevel {
    Particles<T> slow;
    int ns;
    T h;
    KickPending kick;
    const int *fast_pid;  // parent_id of this level's fast set
    int nf;
};
template <typename T> struct ScatterChain {
    ScatterLevel<T> lv[kScatterChainMax];
    Particles<T> bottom_fast;  // the deepest level's fast set
    int depth = 0;
};
template <typename T>
int launch_update_chain(const Particles<T> &prev, const ScatterChain<T> &c, hipStream_t s);

// ------------------------------------------------- device-side subtrees
// Statistics of the device subtrees run since the accumulator was cleared,
// added up by the subtree kernels themselves (absolute levels; launches on
// one stream run one after the other, so plain read-modify-writes do).
constexpr int kSubtreeLevels = 50;
struct SubtreeAcc {
    long long calls, interactions, ts_pairs;
    int max_level;  // deepest absolute level reached (0 if none)
    int status;     // HHS_OK, or the status of the first failed subtree
    long long level_calls[64];
};
size_t subtree_workspace_bytes(int n_max);
// hold_step(total[0..n), dt, calc) at level clevel with its whole recursion
// in one workgroup-wide launch (MIXED32 kicks, timestep mode ts_mode); its
// statistics are added into acc_dev.  max_level is relative to this call
// (<= kSubtreeLevels - 3).  n_dev non-null: a speculative subtree whose
// count is read on the device (n is then its upper bound, at most 256),
// which runs only when that count is in [1, n].
// kick: total's pending kick (KickDefer), applied to the velocities as the
// subtree loads them.
int launch_subtree(const Particles<double> &total, int n, int clevel, double dt, int calc,
                   double eta, double eps2, int ts_mode, int max_level, void *ws,
                   size_t ws_bytes, SubtreeAcc *acc_dev, hipStream_t st,
                   const int *n_dev = nullptr, const KickPending &kick = KickPending{});

}  // namespace hhs
