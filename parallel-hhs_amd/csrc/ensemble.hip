// Batched HOLD: many independent systems integrated in one persistent launch.
//
// No reference counterpart (the reference integrates one system, host
// driven, with a blocking sync per recursion call, src/split.cuh:430).  Here
// a "team" -- a whole 256-thread workgroup, or one 64-lane wave for small
// systems -- owns one system at a time and runs its entire HOLD integration
// (root steps, the recursion of src/hold.cuh:51-181 as an explicit stack in
// LDS, timesteps, splits, drifts, kicks) without returning to the host.
//
// * Work distribution: workgroups pull work items from one atomic queue,
//   ordered on the device by a cost estimate from the initial state (largest
//   first): a "big" system takes the whole workgroup, a bundle of four
//   "small" systems one wave each.
// * A system's numbers depend only on its own data, not on the team size
//   (per-sink source order; a call with <= 256 sources sums and minimises
//   over four source quarters combined in order, a larger one over
//   64-source kick blocks from the range start; a stable partition;
//   energies from the first wave), so a system integrates to the same bits
//   alone or in any mix, on a wave or a workgroup team, and a device
//   subtree to the same bits as the ensemble.
// * In-place hierarchy: a recursion level owns a contiguous index range
//   [b, e) of the system's working copy; the split partitions it stably,
//   fast first, through a scratch copy.  No scatter-back is needed; the
//   original particle order is restored by id at the end.
// * Same numerics as the per-system path: FP64 positions/velocities/drift,
//   FP64-derived FP32 separations with two-sum compensated FP32 force sums,
//   bit-exact FP64 pair timesteps (pair_math.h).
#include <algorithm>
#include <mutex>
#include <numeric>
#include <set>
#include <tuple>
#include <vector>

#include "hhs_internal.h"
#include "pair_math.h"

namespace hhs {
namespace {

constexpr int kWG = 256;
constexpr int kWaves = kWG / 64;
// Team of an LDS-resident device subtree (<= kSubLds particles): four
// waves per SIMD, so the one workgroup on the GPU hides its own latency.
constexpr int kSubTeam = 1024;
constexpr int kMaxFrames = kSubtreeLevels;  // max_level <= 47 plus root and one child
constexpr int kFold = 64;
constexpr int kCounters = 16;  // queue counters in the workspace

struct Frame {
    int b, e, nf, stage;
    double dt;
    int calc, pad;
};

template <int TEAM> struct TeamShared {
    Frame stack[kMaxFrames];
    int level_calls[kMaxFrames];  // calls per recursion level (relative)
    int sp;
    int status;
    int wave_tot[TEAM / 64];
    double red[64];
};

struct EnsArgs {
    const int *offsets;
    double *pos, *vel;
    const double *m;
    const double *t_end, *dt_root;
    double eta, c_ff, eps2;
    float eps2f;
    int max_level;
    long long max_calls;
    double *wpos, *wvel, *wm, *wts;
    int *wid;
    double *tpos, *tvel, *tm, *tts;
    int *tid;
    const int *big_list;
    int n_big;
    const int *small_list;
    int n_small;
    double *est;        // cost estimate per system
    int *small_sorted;  // small systems, largest estimate first
    int *queue;         // work items, largest estimate first: >= 0 a big
                        // system, < 0 a bundle of four small ones
    int n_items;
    unsigned int *counters;
    hhs_system_result *results;
    int *done;
    int ts_mode;
    // Decision guard of the approximate timestep modes (pair_math.h): the
    // band of every call's TsGuard; guard == 0 in the exact mode.
    double guard_qmax;
    unsigned long long guard_lo, guard_hi;
    int guard;
};

// The guard's band for a timestep mode (none for the exact mode).
inline void set_guard(EnsArgs &a, int ts_mode) {
    const double band = ts_mode == HHS_TIMESTEP_FP32   ? kGuardBandFp32
                        : ts_mode == HHS_TIMESTEP_FAST ? kGuardBandFast
                                                       : 0.0;
    const TsGuard g = make_ts_guard(1.0, band);
    a.guard = band > 0.0 ? 1 : 0;
    a.guard_qmax = g.qmax;
    a.guard_lo = g.lo;
    a.guard_hi = g.hi;
}

// A workgroup team (TEAM = the block size) or a one-wave team (TEAM = 64).
template <int TEAM> struct Team {
    int t;
    __device__ Team() : t(threadIdx.x) {}
    __device__ void sync() const { __syncthreads(); }
};
template <> struct Team<64> {
    int t;
    __device__ Team() : t(threadIdx.x & 63) {}
    __device__ void sync() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};

__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Sum of one int per thread over the team; every thread gets the total.
template <int TEAM>
__device__ int team_sum(int x, TeamShared<TEAM> &sh, const Team<TEAM> &tm) {
    x = wave_sum(x);
    if constexpr (TEAM == 64) {
        return x;
    } else {
        if ((tm.t & 63) == 0) sh.wave_tot[tm.t >> 6] = x;
        tm.sync();
        int s = 0;
#pragma unroll
        for (int k = 0; k < TEAM / 64; ++k) s += sh.wave_tot[k];
        tm.sync();
        return s;
    }
}

// Exclusive rank of p among the team's threads in thread order, and the
// team total.  TAIL: a barrier after the wave totals are read, so that the
// next call may overwrite them; a caller that passes another barrier before
// its next call leaves it out.
template <int TEAM, bool TAIL = true>
__device__ void team_scan_bits(bool p, int &excl, int &tot, TeamShared<TEAM> &sh,
                               const Team<TEAM> &tm) {
    const int lane = tm.t & 63;
    const uint64_t mask = __ballot(p);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int in_wave = __popcll(mask & lt);
    if constexpr (TEAM == 64) {
        excl = in_wave;
        tot = __popcll(mask);
    } else {
        const int w = tm.t >> 6;
        if (lane == 0) sh.wave_tot[w] = __popcll(mask);
        tm.sync();
        int before = 0, all = 0;
#pragma unroll
        for (int k = 0; k < TEAM / 64; ++k) {
            const int c = sh.wave_tot[k];
            before += (k < w) ? c : 0;
            all += c;
        }
        if constexpr (TAIL) tm.sync();
        excl = before + in_wave;
        tot = all;
    }
}

__device__ __forceinline__ void copy_rec(const EnsArgs &a, bool to_tmp, int src, int dst) {
    const double *sp = to_tmp ? a.wpos : a.tpos;
    const double *sv = to_tmp ? a.wvel : a.tvel;
    const double *sm = to_tmp ? a.wm : a.tm;
    const double *st = to_tmp ? a.wts : a.tts;
    const int *si = to_tmp ? a.wid : a.tid;
    double *dp = to_tmp ? a.tpos : a.wpos;
    double *dv = to_tmp ? a.tvel : a.wvel;
    double *dm = to_tmp ? a.tm : a.wm;
    double *dt = to_tmp ? a.tts : a.wts;
    int *di = to_tmp ? a.tid : a.wid;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dp[3 * dst + k] = sp[3 * src + k];
        dv[3 * dst + k] = sv[3 * src + k];
    }
    dm[dst] = sm[src];
    dt[dst] = st[src];
    di[dst] = si[src];
}

// Source tiles: the team copies TEAM sources at a time into its LDS tile
// (one coalesced load per thread), then every lane reads them as LDS
// broadcasts -- one L2 round trip per tile instead of one per source.
// Timestep tiles hold x, y, z, m, vx, vy, vz (stride kTsStride doubles);
// kick tiles hold x, y, z, m (stride kKickStride).
constexpr int kTsStride = 8;
constexpr int kKickStride = 4;

template <int TEAM>
__device__ __forceinline__ void load_ts_tile(const EnsArgs &a, int s0, int je, double *tile,
                                             const Team<TEAM> &tm) {
    tm.sync();  // the previous tile has been read by every lane
    const int j = s0 + tm.t;
    if (j < je) {
        double *q = tile + kTsStride * tm.t;
        q[0] = a.wpos[3 * j];
        q[1] = a.wpos[3 * j + 1];
        q[2] = a.wpos[3 * j + 2];
        q[3] = a.wm[j];
        q[4] = a.wvel[3 * j];
        q[5] = a.wvel[3 * j + 1];
        q[6] = a.wvel[3 * j + 2];
    }
    tm.sync();
}

// One pass over the sources [jb, je) (index order, through LDS tiles) for S
// sinks per lane: sinks i0 + t + s * TEAM.  Each sink's arithmetic and
// source order are those of one sink per lane, so the result does not
// depend on S; several sinks per lane give each wave independent chains to
// interleave, which a team running alone on its CU needs.
template <int TEAM, int S, int MODE>
__device__ void team_timesteps_pass(const EnsArgs &a, int jb, int je, int i0, double *tile,
                                    const Team<TEAM> &tm) {
    int ii[S];
    double xi[S], yi[S], zi[S], vxi[S], vyi[S], vzi[S], mi[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        ii[s] = i0 + tm.t + s * TEAM;
        const int ic = ii[s] < je ? ii[s] : je - 1;
        xi[s] = a.wpos[3 * ic];
        yi[s] = a.wpos[3 * ic + 1];
        zi[s] = a.wpos[3 * ic + 2];
        vxi[s] = a.wvel[3 * ic];
        vyi[s] = a.wvel[3 * ic + 1];
        vzi[s] = a.wvel[3 * ic + 2];
        mi[s] = a.wm[ic];
    }
    double minval[S], bden[S];
    float bnum32[S], bden32[S], mif[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        minval[s] = TsConst<double>::huge;
        bden[s] = 1.0;
        bnum32[s] = static_cast<float>(TsConst<double>::huge);
        bden32[s] = 1.0f;
        mif[s] = static_cast<float>(mi[s]);
    }
    const float etaf = static_cast<float>(a.eta), cff = static_cast<float>(a.c_ff);
    for (int s0 = jb; s0 < je; s0 += TEAM) {
        load_ts_tile<TEAM>(a, s0, je, tile, tm);
        const int cnt = uni(min(TEAM, je - s0));
#pragma unroll 2
        for (int k = 0; k < cnt; ++k) {
            const double *q = tile + kTsStride * k;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const bool other = s0 + k != ii[s];
                if constexpr (MODE == HHS_TIMESTEP_FP32)
                    pair_timestep_f32(xi[s], yi[s], zi[s], vxi[s], vyi[s], vzi[s], mif[s], q[0],
                                      q[1], q[2], q[4], q[5], q[6], static_cast<float>(q[3]), etaf,
                                      cff, a.eps2f, other, bnum32[s], bden32[s]);
                else if constexpr (MODE == HHS_TIMESTEP_FAST)
                    pair_timestep_fast(xi[s], yi[s], zi[s], vxi[s], vyi[s], vzi[s], mi[s], q[0],
                                       q[1], q[2], q[4], q[5], q[6], q[3], a.eta, a.c_ff, a.eps2,
                                       other, minval[s], bden[s]);
                else
                    pair_timestep<double>(xi[s], yi[s], zi[s], vxi[s], vyi[s], vzi[s], mi[s], q[0],
                                          q[1], q[2], q[4], q[5], q[6], q[3], a.eta, a.c_ff,
                                          a.eps2, other, minval[s]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        double v = minval[s];
        if constexpr (MODE == HHS_TIMESTEP_FP32)
            v = static_cast<double>(bnum32[s]) / static_cast<double>(bden32[s]);
        else if constexpr (MODE == HHS_TIMESTEP_FAST)
            v = minval[s] / bden[s];
        if (ii[s] < je) a.wts[ii[s]] = v;
    }
}

// Sinks per lane for a pass with `rem` sinks left: enough to cover them,
// at most kMaxSpl (FP32 timesteps and kicks; the FP64 timestep modes keep
// one, their pair state would cost the kernel a wave of occupancy).
constexpr int kMaxSpl = 4;
template <int TEAM> __device__ __forceinline__ int sinks_per_lane(int rem, int cap) {
    return min(cap, (rem + TEAM - 1) / TEAM);
}

// ------------------------------------------------ calls with few sources
// A call whose source range holds at most kSmallSrc particles (the deep
// levels of every system, and every call of a device subtree) has few
// sinks too, so one lane per sink would leave most of a workgroup idle:
// at 50 sinks one wave does all the pair work while three wait.  Such calls
// use one fixed sum structure instead: the sources [jb, je) are cut into
// kQuarters quarters of ceil(L / kQuarters); each quarter's minimum (a
// fresh fold from 1e30) or force (a plain FP32 sum from zero) is formed on
// its own, and the quarters are combined in quarter order (a fold of the
// fractions; FP64 adds of the sums).  A workgroup team with at least two
// waves per group of 64 sinks (<= 128 sinks on kWG threads, <= 256 on the
// kSubTeam threads of an LDS subtree) forms the quarters in parallel, one
// or two per wave, and combines them through LDS; other teams form them
// one after the other.  The bits are the same either way.
constexpr int kQuarters = 4;
constexpr int kSmallSrc = kWG;  // the whole source range fits one workgroup tile
// A workgroup team's quarter partials sit in its LDS tile array after the
// first kTileDoubles.  Groups of 64 sinks whose quarters are spread over
// waves (2 waves or more per group) number at most kMaxSplitGroups; the
// area holds two kicks' FP32 quarter sums or one pass's quarter minima.
constexpr int kTileDoubles = kWG * kTsStride;
template <int TEAM> constexpr int kMaxSplitGroups = TEAM / 128 < 4 ? TEAM / 128 : 4;
template <int TEAM> constexpr int kPartDoubles = 2 * kMaxSplitGroups<TEAM> * kQuarters * 3 * 64 / 2;

struct Span {
    int b, e;
};
__device__ __forceinline__ Span quarter(int jb, int je, int q) {
    const int qs = (je - jb + kQuarters - 1) / kQuarters;
    const int b = min(jb + q * qs, je);
    return Span{b, min(b + qs, je)};
}

// Running minimum of one sink in the representation of each mode (FP32 and
// FAST: a fraction num/den; EXACT: the value), as in team_timesteps_pass.
template <int MODE> struct TsFold {
    using V = typename std::conditional<MODE == HHS_TIMESTEP_FP32, float, double>::type;
    V num, den;
    __device__ void init() {
        num = static_cast<V>(TsConst<double>::huge);
        den = V(1);
    }
    __device__ void merge(V onum, V oden) {
        if constexpr (MODE == HHS_TIMESTEP_FP32)
            fold_fraction32(onum, oden, true, num, den);
        else if constexpr (MODE == HHS_TIMESTEP_FAST)
            fold_fraction(onum, oden, true, num, den);
        else if (onum < num)
            num = onum;
    }
    __device__ double value() const {
        if constexpr (MODE == HHS_TIMESTEP_FP32)
            return static_cast<double>(num) / static_cast<double>(den);
        else if constexpr (MODE == HHS_TIMESTEP_FAST)
            return num / den;
        else
            return num;
    }
};

struct TsSink {
    double x, y, z, vx, vy, vz, m;
    float mf;
    int i;
};

__device__ __forceinline__ TsSink ts_sink(const EnsArgs &a, int i, int je) {
    const int ic = i < je ? i : je - 1;
    TsSink s;
    s.x = a.wpos[3 * ic];
    s.y = a.wpos[3 * ic + 1];
    s.z = a.wpos[3 * ic + 2];
    s.vx = a.wvel[3 * ic];
    s.vy = a.wvel[3 * ic + 1];
    s.vz = a.wvel[3 * ic + 2];
    s.m = a.wm[ic];
    s.mf = static_cast<float>(s.m);
    s.i = i;
    return s;
}

// acc[s] <- the fold, from 1e30, of sources [qb, qe) (tile entry j - t0) for
// S sinks.
template <int MODE, int S>
__device__ __forceinline__ void ts_fold_span(const EnsArgs &a, const TsSink (&s)[S],
                                             const double *tile, int t0, int qb, int qe,
                                             float etaf, float cff, TsFold<MODE> (&acc)[S]) {
#pragma unroll
    for (int k = 0; k < S; ++k) acc[k].init();
#pragma unroll 2
    for (int j = qb; j < qe; ++j) {
        const double *q = tile + kTsStride * (j - t0);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const TsSink &p = s[k];
            const bool other = j != p.i;
            if constexpr (MODE == HHS_TIMESTEP_FP32)
                pair_timestep_f32(p.x, p.y, p.z, p.vx, p.vy, p.vz, p.mf, q[0], q[1], q[2], q[4],
                                  q[5], q[6], static_cast<float>(q[3]), etaf, cff, a.eps2f, other,
                                  acc[k].num, acc[k].den);
            else if constexpr (MODE == HHS_TIMESTEP_FAST)
                pair_timestep_fast(p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m, q[0], q[1], q[2], q[4],
                                   q[5], q[6], q[3], a.eta, a.c_ff, a.eps2, other, acc[k].num,
                                   acc[k].den);
            else
                pair_timestep<double>(p.x, p.y, p.z, p.vx, p.vy, p.vz, p.m, q[0], q[1], q[2],
                                      q[4], q[5], q[6], q[3], a.eta, a.c_ff, a.eps2, other,
                                      acc[k].num);
        }
    }
}

// Wave team: S sinks per lane (i0 + lane + 64 s), the quarters one after
// the other, from one tile when L <= 64, else each through its own tile (a
// quarter holds <= 64 sources).
template <int MODE, int S>
__device__ __forceinline__ void ts_small_wave_pass(const EnsArgs &a, int jb, int je, int i0, double *tile,
                                   const Team<64> &tm, float etaf, float cff) {
    TsSink s[S];
    TsFold<MODE> tot[S], acc[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
        s[k] = ts_sink(a, i0 + tm.t + 64 * k, je);
        tot[k].init();
    }
    const bool whole = je - jb <= 64;  // one tile holds every source
    if (whole) load_ts_tile<64>(a, jb, je, tile, tm);
#pragma unroll 1
    for (int q = 0; q < kQuarters; ++q) {
        const Span Q = quarter(jb, je, q);
        if (!whole) load_ts_tile<64>(a, Q.b, Q.e, tile, tm);
        ts_fold_span<MODE, S>(a, s, tile, whole ? jb : Q.b, Q.b, Q.e, etaf, cff, acc);
#pragma unroll
        for (int k = 0; k < S; ++k) tot[k].merge(acc[k].num, acc[k].den);
    }
#pragma unroll
    for (int k = 0; k < S; ++k)
        if (s[k].i < je) a.wts[s[k].i] = tot[k].value();
}

// Waves per sink group for a workgroup team with G groups of 64 sinks: 4,
// 2 or 1 (each wave then forms 1, 2 or 4 quarters), as many as the team
// holds; 1 beyond kMaxSplitGroups groups (the partial area's size).
template <int TEAM> __device__ __forceinline__ int waves_per_group(int G) {
    constexpr int nw = TEAM / 64;
    if (G > kMaxSplitGroups<TEAM>) return 1;
    return 4 * G <= nw ? 4 : (2 * G <= nw ? 2 : 1);
}

// Workgroup team (TEAM = kWG, or kSubTeam for LDS subtrees): the L <=
// kSmallSrc sources in one tile (entry j - jb), one sink per lane in G =
// ceil(L / 64) groups of 64.  With R = waves_per_group > 1, the R waves of
// a group form 4 / R quarters each, and the group's first wave combines
// them from the partial area.
template <int TEAM, int MODE>
__device__ __forceinline__ void ts_small_wg(const EnsArgs &a, int jb, int je, double *tile,
                                            const Team<TEAM> &tm, float etaf, float cff) {
    using V = typename TsFold<MODE>::V;
    constexpr int GM = kMaxSplitGroups<TEAM>;
    static_assert(2 * GM * kQuarters * 64 * sizeof(V) <= kPartDoubles<TEAM> * sizeof(double),
                  "the quarter minima fit the partial area");
    load_ts_tile<TEAM>(a, jb, je, tile, tm);
    const int G = (je - jb + 63) >> 6;
    const int w = uni(tm.t >> 6), lane = tm.t & 63;
    const int R = waves_per_group<TEAM>(G);
    if (R == 1) {
        if (w < G) {
            TsSink s[1] = {ts_sink(a, jb + 64 * w + lane, je)};
            TsFold<MODE> tot, acc[1];
            tot.init();
#pragma unroll 1
            for (int q = 0; q < kQuarters; ++q) {
                const Span Q = quarter(jb, je, q);
                ts_fold_span<MODE, 1>(a, s, tile, jb, Q.b, Q.e, etaf, cff, acc);
                tot.merge(acc[0].num, acc[0].den);
            }
            if (s[0].i < je) a.wts[s[0].i] = tot.value();
        }
        return;
    }
    const int per = kQuarters / R, g = w / R, r = w % R;  // per: 1 or 2
    TsSink s[1] = {ts_sink(a, jb + 64 * g + lane, je)};
    V *pn = reinterpret_cast<V *>(tile + kTileDoubles);
    V *pd = pn + GM * kQuarters * 64;
    if (g < G) {
        TsFold<MODE> acc[1];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (k < per) {
                const Span Q = quarter(jb, je, r * per + k);
                ts_fold_span<MODE, 1>(a, s, tile, jb, Q.b, Q.e, etaf, cff, acc);
                const int slot = (g * kQuarters + r * per + k) * 64 + lane;
                pn[slot] = acc[0].num;
                pd[slot] = acc[0].den;
            }
        }
    }
    tm.sync();
    if (r == 0 && g < G) {
        TsFold<MODE> tot;
        tot.init();
#pragma unroll
        for (int q = 0; q < kQuarters; ++q) {
            const int slot = (g * kQuarters + q) * 64 + lane;
            tot.merge(pn[slot], pd[slot]);
        }
        if (s[0].i < je) a.wts[s[0].i] = tot.value();
    }
}

template <int TEAM, int MODE>
__device__ __forceinline__ void team_timesteps_small(const EnsArgs &a, int jb, int je, double *tile,
                                     const Team<TEAM> &tm) {
    const float etaf = static_cast<float>(a.eta), cff = static_cast<float>(a.c_ff);
    if constexpr (TEAM == 64) {
        constexpr int cap = MODE == HHS_TIMESTEP_FP32 ? kMaxSpl : 1;
        for (int i0 = jb; i0 < je;) {
            const int S = sinks_per_lane<64>(je - i0, cap);
            if (S >= 4)
                ts_small_wave_pass<MODE, (cap >= 4 ? 4 : 1)>(a, jb, je, i0, tile, tm, etaf, cff);
            else if (S == 3)
                ts_small_wave_pass<MODE, (cap >= 3 ? 3 : 1)>(a, jb, je, i0, tile, tm, etaf, cff);
            else if (S == 2)
                ts_small_wave_pass<MODE, (cap >= 2 ? 2 : 1)>(a, jb, je, i0, tile, tm, etaf, cff);
            else
                ts_small_wave_pass<MODE, 1>(a, jb, je, i0, tile, tm, etaf, cff);
            i0 += S * 64;
        }
    } else {
        ts_small_wg<TEAM, MODE>(a, jb, je, tile, tm, etaf, cff);
    }
}

template <int TEAM, int MODE>
__device__ __forceinline__ void team_timesteps_mode(const EnsArgs &a, int jb, int je, double *tile,
                                    const Team<TEAM> &tm) {
    constexpr int cap = MODE == HHS_TIMESTEP_FP32 ? kMaxSpl : 1;
    if constexpr (TEAM > kWG) {
        // teams wider than kWG run LDS subtrees only: <= kSmallSrc particles
        if (je > jb) team_timesteps_small<TEAM, MODE>(a, jb, je, tile, tm);
    } else {
        if (je - jb <= kSmallSrc) {
            if (je > jb) team_timesteps_small<TEAM, MODE>(a, jb, je, tile, tm);
            return;
        }
        for (int i0 = jb; i0 < je;) {
            const int S = sinks_per_lane<TEAM>(je - i0, cap);
            if (S >= 4) {
                team_timesteps_pass<TEAM, (cap >= 4 ? 4 : 1), MODE>(a, jb, je, i0, tile, tm);
            } else if (S == 3) {
                team_timesteps_pass<TEAM, (cap >= 3 ? 3 : 1), MODE>(a, jb, je, i0, tile, tm);
            } else if (S == 2) {
                team_timesteps_pass<TEAM, (cap >= 2 ? 2 : 1), MODE>(a, jb, je, i0, tile, tm);
            } else {
                team_timesteps_pass<TEAM, 1, MODE>(a, jb, je, i0, tile, tm);
            }
            i0 += S * TEAM;
        }
    }
}

// timesteps of particles [o+b, o+e) against each other (src/timestep.cuh),
// sources in index order through LDS tiles.
template <int TEAM>
__device__ __forceinline__ void team_timesteps(const EnsArgs &a, int o, int b, int e, double *tile,
                               const Team<TEAM> &tm) {
    const int jb = uni(o + b), je = uni(o + e);
    if (a.ts_mode == HHS_TIMESTEP_FP32)
        team_timesteps_mode<TEAM, HHS_TIMESTEP_FP32>(a, jb, je, tile, tm);
    else if (a.ts_mode == HHS_TIMESTEP_FAST)
        team_timesteps_mode<TEAM, HHS_TIMESTEP_FAST>(a, jb, je, tile, tm);
    else
        team_timesteps_mode<TEAM, HHS_TIMESTEP_EXACT>(a, jb, je, tile, tm);
}

// Decision guard (pair_math.h) on timesteps the partition has just loaded:
// lane values `ts` of sinks `i` (valid when `in`) of the call whose sources
// are [jb, je).  guard_mask gives the wave's flagged lanes; guard_fix
// recomputes each flagged timestep with the exact formula on its wave (64
// sources per step, a wave minimum) and the owner lane takes the value in
// its register and stores it, so the partition -- which reads each sink on
// its owner thread -- needs no barrier.  A minimum does not depend on
// order, so the value is the exact mode's bit for bit.  Both are called by
// every lane of the wave.
__device__ __forceinline__ uint64_t guard_mask(const TsGuard &g, bool in, double ts) {
    return __ballot(in && guard_near_pivot(g, ts));
}

__device__ __forceinline__ double guard_fix(const EnsArgs &a, uint64_t mask, int jb, int je,
                                            int i, double ts) {
    const int lane = threadIdx.x & 63;
    while (mask) {
        const int l = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        mask &= mask - 1;
        const int s = __shfl(i, l, 64);
        const double xi = a.wpos[3 * s], yi = a.wpos[3 * s + 1], zi = a.wpos[3 * s + 2];
        const double vxi = a.wvel[3 * s], vyi = a.wvel[3 * s + 1], vzi = a.wvel[3 * s + 2];
        const double mi = a.wm[s];
        double best = TsConst<double>::huge;
        for (int j = jb + lane; j < je; j += 64)
            pair_timestep<double>(xi, yi, zi, vxi, vyi, vzi, mi, a.wpos[3 * j], a.wpos[3 * j + 1],
                                  a.wpos[3 * j + 2], a.wvel[3 * j], a.wvel[3 * j + 1],
                                  a.wvel[3 * j + 2], a.wm[j], a.eta, a.c_ff, a.eps2, j != s,
                                  best);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const double o = __shfl_xor(best, d, 64);
            best = o < best ? o : best;
        }
        if (lane == l) {
            ts = best;
            a.wts[s] = best;
        }
    }
    return ts;
}

// Stable partition of [o+b, o+e) by timestep <= pivot, fast first.  Returns
// the fast count after the lone-particle rule (src/split.cuh:321-324).
//
// A range that fits the team (e - b <= TEAM) is partitioned in registers:
// each thread holds its particle's record, one scan gives the fast count
// and the ranks, and the records are stored at their places after it (the
// scan's barrier orders every load before every store).  No barrier
// follows the scan's reads of the wave totals: the caller passes another
// (the end of the frame's stage) before the next scan.  Larger ranges go
// through the scratch copy.  guard (calls that have just computed their
// timesteps in an approximate mode): the decision guard gd runs on the
// loaded timesteps before any predicate is formed.
template <int TEAM, bool GUARD>
__device__ int team_partition(const EnsArgs &a, int o, int b, int e, double pivot,
                              TeamShared<TEAM> &sh, const Team<TEAM> &tm, bool guard,
                              TsGuard gd) {
    if (e - b <= TEAM) {
        const int i = o + b + tm.t;
        const bool in = i < o + e;
        const int ic = in ? i : o + b;
        double x = a.wpos[3 * ic], y = a.wpos[3 * ic + 1], z = a.wpos[3 * ic + 2];
        double vx = a.wvel[3 * ic], vy = a.wvel[3 * ic + 1], vz = a.wvel[3 * ic + 2];
        double m = a.wm[ic], ts = a.wts[ic];
        int id = a.wid[ic];
        const uint64_t fl = (GUARD && guard) ? guard_mask(gd, in, ts) : 0ull;
        if (GUARD && fl) {
            // rare: the record is loaded again after the exact recompute, so
            // that the loop does not keep the loaded copy in registers
            ts = guard_fix(a, fl, o + b, o + e, ic, ts);
            x = a.wpos[3 * ic];
            y = a.wpos[3 * ic + 1];
            z = a.wpos[3 * ic + 2];
            vx = a.wvel[3 * ic];
            vy = a.wvel[3 * ic + 1];
            vz = a.wvel[3 * ic + 2];
            m = a.wm[ic];
            id = a.wid[ic];
        }
        const bool p = in && ts <= pivot;
        int excl, nf;
        team_scan_bits<TEAM, false>(p, excl, nf, sh, tm);
        if (nf == 1) return 0;                  // lone fast particle stays slow, order kept
        if (nf == 0 || nf == e - b) return nf;  // stable partition is the identity
        if (in) {
            const int d = o + b + (p ? excl : nf + tm.t - excl);
            a.wpos[3 * d] = x;
            a.wpos[3 * d + 1] = y;
            a.wpos[3 * d + 2] = z;
            a.wvel[3 * d] = vx;
            a.wvel[3 * d + 1] = vy;
            a.wvel[3 * d + 2] = vz;
            a.wm[d] = m;
            a.wts[d] = ts;
            a.wid[d] = id;
        }
        tm.sync();
        return nf;
    }
    int c = 0;
    if (GUARD && guard) {
        for (int c0 = o + b; c0 < o + e; c0 += TEAM) {  // uniform trip count
            const int i = c0 + tm.t;
            const bool in = i < o + e;
            double ts = in ? a.wts[i] : 0.0;
            const uint64_t fl = guard_mask(gd, in, ts);
            if (fl) ts = guard_fix(a, fl, o + b, o + e, i, ts);
            c += (in && ts <= pivot) ? 1 : 0;
        }
    } else {
        for (int i = o + b + tm.t; i < o + e; i += TEAM) c += (a.wts[i] <= pivot) ? 1 : 0;
    }
    const int nf = team_sum<TEAM>(c, sh, tm);
    if (nf == 1) return 0;               // lone fast particle stays slow, order kept
    if (nf == 0 || nf == e - b) return nf;  // stable partition is the identity
    int fbase = 0, sbase = 0;
    for (int c0 = b; c0 < e; c0 += TEAM) {
        const int i = c0 + tm.t;
        const bool p = i < e && a.wts[o + i] <= pivot;
        int excl, tot;
        team_scan_bits<TEAM>(p, excl, tot, sh, tm);
        if (i < e) {
            const int d = p ? (b + fbase + excl) : (b + nf + sbase + (i - c0) - excl);
            copy_rec(a, true, o + i, o + d);
        }
        fbase += tot;
        sbase += min(TEAM, e - c0) - tot;
    }
    tm.sync();
    for (int i = o + b + tm.t; i < o + e; i += TEAM) copy_rec(a, false, i, i);
    tm.sync();
    return nf;
}

template <int TEAM>
__device__ void team_drift(const EnsArgs &a, int o, int b, int e, double h,
                           const Team<TEAM> &tm) {
#pragma clang fp contract(off)
    for (int k = 3 * (o + b) + tm.t; k < 3 * (o + e); k += TEAM)
        a.wpos[k] = a.wpos[k] + a.wvel[k] * h;  // src/drift.cuh:25
}

// vel[i] += a * h, the same expression on every path, with the FP32 force
// sum a debiased for v_rsq_f32's mean error (kRsqDebias).
__device__ __forceinline__ void kick_apply(const EnsArgs &a, int i, double ax, double ay,
                                           double az, double h) {
    const double hc = h * kRsqDebias;
    a.wvel[3 * i + 0] = a.wvel[3 * i + 0] + ax * hc;
    a.wvel[3 * i + 1] = a.wvel[3 * i + 1] + ay * hc;
    a.wvel[3 * i + 2] = a.wvel[3 * i + 2] + az * hc;
}

// One pass of team_kick for S sinks per lane (sinks i0 + t + s * TEAM); the
// per-sink sum is the same for every S.
template <int TEAM, int S, bool SOFT>
__device__ void team_kick_pass(const EnsArgs &a, int i0, int se, int jb, int je, double h,
                               double *tile, const Team<TEAM> &tm) {
    int ii[S];
    double xi[S], yi[S], zi[S];
    float hx[S], lx[S], hy[S], ly[S], hz[S], lz[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        ii[s] = i0 + tm.t + s * TEAM;
        const int ic = ii[s] < se ? ii[s] : se - 1;
        xi[s] = a.wpos[3 * ic];
        yi[s] = a.wpos[3 * ic + 1];
        zi[s] = a.wpos[3 * ic + 2];
        hx[s] = lx[s] = hy[s] = ly[s] = hz[s] = lz[s] = 0.f;
    }
    for (int s0 = jb; s0 < je; s0 += TEAM) {
        tm.sync();  // the previous tile has been read by every lane
        const int j = s0 + tm.t;
        if (j < je) {
            double *q = tile + kKickStride * tm.t;
            q[0] = a.wpos[3 * j];
            q[1] = a.wpos[3 * j + 1];
            q[2] = a.wpos[3 * j + 2];
            q[3] = a.wm[j];
        }
        tm.sync();
        const int cnt = uni(min(TEAM, je - s0));
        for (int k0 = 0; k0 < cnt; k0 += kFold) {
            const int k1 = min(k0 + kFold, cnt);
            float px[S], py[S], pz[S];
#pragma unroll
            for (int s = 0; s < S; ++s) px[s] = py[s] = pz[s] = 0.f;
#pragma unroll 4
            for (int k = k0; k < k1; ++k) {
                const double *q = tile + kKickStride * k;
                const float mj = static_cast<float>(q[3]);
#pragma unroll
                for (int s = 0; s < S; ++s)
                    pair_accel<SOFT>(q[0], q[1], q[2], mj, xi[s], yi[s], zi[s], a.eps2f, px[s],
                                     py[s], pz[s]);
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                two_sum_add(hx[s], lx[s], px[s]);
                two_sum_add(hy[s], ly[s], py[s]);
                two_sum_add(hz[s], lz[s], pz[s]);
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = ii[s];
        if (i < se) {
            const double ax = static_cast<double>(hx[s]) + static_cast<double>(lx[s]);
            const double ay = static_cast<double>(hy[s]) + static_cast<double>(ly[s]);
            const double az = static_cast<double>(hz[s]) + static_cast<double>(lz[s]);
            kick_apply(a, i, ax, ay, az, h);
        }
    }
}

// Kick tile: sources [s0, min(s0 + TEAM, je)) as x, y, z, m (entry j - s0).
template <int TEAM>
__device__ __forceinline__ void load_kick_tile(const EnsArgs &a, int s0, int je, double *tile,
                                               const Team<TEAM> &tm) {
    tm.sync();  // the previous tile has been read by every lane
    const int j = s0 + tm.t;
    if (j < je) {
        double *q = tile + kKickStride * tm.t;
        q[0] = a.wpos[3 * j];
        q[1] = a.wpos[3 * j + 1];
        q[2] = a.wpos[3 * j + 2];
        q[3] = a.wm[j];
    }
    tm.sync();
}


// Plain FP32 force sums from zero over sources [qb, qe) (tile entry j - t0)
// for S sinks.
template <int S, bool SOFT>
__device__ __forceinline__ void kick_span(const EnsArgs &a, const double (&xi)[S],
                                          const double (&yi)[S], const double (&zi)[S],
                                          const double *tile, int t0, int qb, int qe,
                                          float (&px)[S], float (&py)[S], float (&pz)[S]) {
#pragma unroll
    for (int s = 0; s < S; ++s) px[s] = py[s] = pz[s] = 0.f;
#pragma unroll 4
    for (int j = qb; j < qe; ++j) {
        const double *q = tile + kKickStride * (j - t0);
        const float mj = static_cast<float>(q[3]);
#pragma unroll
        for (int s = 0; s < S; ++s)
            pair_accel<SOFT>(q[0], q[1], q[2], mj, xi[s], yi[s], zi[s], a.eps2f, px[s], py[s],
                             pz[s]);
    }
}

// S sinks per lane (i0 + t + TEAM s), the quarters one after the other.
// WHOLE: the tile already holds every source (entry j - tb); otherwise the
// sources are loaded here, all at once when they fit one tile, else each
// quarter into its own tile (wave teams: a quarter holds <= 64 sources).
template <int TEAM, int S, bool SOFT, bool WHOLE>
__device__ __forceinline__ void kick_small_pass(const EnsArgs &a, int i0, int se, int jb, int je,
                                                int tb, double h, double *tile,
                                                const Team<TEAM> &tm) {
    int ii[S];
    double xi[S], yi[S], zi[S], ax[S], ay[S], az[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        ii[s] = i0 + tm.t + s * TEAM;
        const int ic = ii[s] < se ? ii[s] : se - 1;
        xi[s] = a.wpos[3 * ic];
        yi[s] = a.wpos[3 * ic + 1];
        zi[s] = a.wpos[3 * ic + 2];
        ax[s] = ay[s] = az[s] = 0.0;
    }
    // a wave team's tile holds TEAM sources: one load when they all fit
    const bool whole = WHOLE || je - jb <= TEAM;
    if (!WHOLE && whole) load_kick_tile<TEAM>(a, jb, je, tile, tm);
    const int t0 = WHOLE ? tb : jb;
#pragma unroll 1
    for (int q = 0; q < kQuarters; ++q) {
        const Span Q = quarter(jb, je, q);
        if (!whole) load_kick_tile<TEAM>(a, Q.b, Q.e, tile, tm);
        float px[S], py[S], pz[S];
        kick_span<S, SOFT>(a, xi, yi, zi, tile, whole ? t0 : Q.b, Q.b, Q.e, px, py, pz);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            ax[s] += static_cast<double>(px[s]);
            ay[s] += static_cast<double>(py[s]);
            az[s] += static_cast<double>(pz[s]);
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (ii[s] < se) kick_apply(a, ii[s], ax[s], ay[s], az[s], h);
}

template <int TEAM, bool SOFT, bool WHOLE>
__device__ __forceinline__ void kick_small_passes(const EnsArgs &a, int sb, int se, int jb, int je,
                                                  int tb, double h, double *tile,
                                                  const Team<TEAM> &tm) {
    for (int i0 = sb; i0 < se;) {
        const int S = sinks_per_lane<TEAM>(se - i0, kMaxSpl);
        if (S >= 4)
            kick_small_pass<TEAM, 4, SOFT, WHOLE>(a, i0, se, jb, je, tb, h, tile, tm);
        else if (S == 3)
            kick_small_pass<TEAM, 3, SOFT, WHOLE>(a, i0, se, jb, je, tb, h, tile, tm);
        else if (S == 2)
            kick_small_pass<TEAM, 2, SOFT, WHOLE>(a, i0, se, jb, je, tb, h, tile, tm);
        else
            kick_small_pass<TEAM, 1, SOFT, WHOLE>(a, i0, se, jb, je, tb, h, tile, tm);
        i0 += S * TEAM;
    }
}

// Workgroup team (TEAM >= kWG), one kick with the quarter sums: sinks
// [s0, s1) from L <= kSmallSrc sources [j0, j1) already in the tile (entry
// j - tb).  With R = waves_per_group > 1 the R waves of each group of 64
// sinks form 4 / R quarters each and leave them in the partial area (kick
// slot k); returns whether it did (kick_small_wg_add then adds them after a
// barrier).  Otherwise every lane sums its sinks' quarters itself.
template <int TEAM, bool SOFT>
__device__ __forceinline__ bool kick_small_wg_quarters(const EnsArgs &a, int k, int s0, int s1,
                                                       int j0, int j1, int tb, double h,
                                                       double *tile, float *part,
                                                       const Team<TEAM> &tm) {
    constexpr int GM = kMaxSplitGroups<TEAM>;
    const int G = (s1 - s0 + 63) >> 6, R = waves_per_group<TEAM>(G);
    if (R == 1) {
        kick_small_passes<TEAM, SOFT, true>(a, s0, s1, j0, j1, tb, h, tile, tm);
        return false;
    }
    const int w = uni(tm.t >> 6), lane = tm.t & 63;
    const int per = kQuarters / R, g = w / R, r = w % R;  // per: 1 or 2
    if (g < G) {
        const int i = s0 + 64 * g + lane;
        const int ic = i < s1 ? i : s1 - 1;
        const double xi[1] = {a.wpos[3 * ic]}, yi[1] = {a.wpos[3 * ic + 1]},
                     zi[1] = {a.wpos[3 * ic + 2]};
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            if (kk < per) {
                const Span Q = quarter(j0, j1, r * per + kk);
                float px[1], py[1], pz[1];
                kick_span<1, SOFT>(a, xi, yi, zi, tile, tb, Q.b, Q.e, px, py, pz);
                const int slot = (k * GM + g) * kQuarters + r * per + kk;
                part[(3 * slot + 0) * 64 + lane] = px[0];
                part[(3 * slot + 1) * 64 + lane] = py[0];
                part[(3 * slot + 2) * 64 + lane] = pz[0];
            }
        }
    }
    return true;
}

// The quarters of kick slot k added in quarter order (FP64) by each group's
// first wave, and the velocity update.
template <int TEAM>
__device__ __forceinline__ void kick_small_wg_add(const EnsArgs &a, int k, int s0, int s1,
                                                  double h, const float *part,
                                                  const Team<TEAM> &tm) {
    constexpr int GM = kMaxSplitGroups<TEAM>;
    const int G = (s1 - s0 + 63) >> 6, R = waves_per_group<TEAM>(G);
    const int w = uni(tm.t >> 6), lane = tm.t & 63;
    const int g = w / R;
    if (w % R != 0 || g >= G) return;
    const int i = s0 + 64 * g + lane;
    double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll
    for (int q = 0; q < kQuarters; ++q) {
        const int slot = (k * GM + g) * kQuarters + q;
        ax += static_cast<double>(part[(3 * slot + 0) * 64 + lane]);
        ay += static_cast<double>(part[(3 * slot + 1) * 64 + lane]);
        az += static_cast<double>(part[(3 * slot + 2) * 64 + lane]);
    }
    if (i < s1) kick_apply(a, i, ax, ay, az, h);
}

// Stage-1 work of a call on [b, e) with <= kSmallSrc particles, fast
// [b, sb), slow [sb, e), for a workgroup team: the slow particles' first
// half-drift (src/hold.cuh:123-124, the expression of team_drift) rides on
// the tile load, each source being loaded by one thread; then slow sinks
// from every particle and fast sinks from the slow ones (src/hold.cuh:
// 137-145), both from that tile; both kicks' quarters are added behind one
// barrier.
template <int TEAM, bool SOFT>
__device__ __forceinline__ void team_kicks_small_wg(const EnsArgs &a, int b, int sb, int e,
                                                    double hd, double h, double *tile,
                                                    const Team<TEAM> &tm) {
    float *part = reinterpret_cast<float *>(tile + kTileDoubles);
    tm.sync();  // every lane is done with the previous tile and partials
    {
#pragma clang fp contract(off)
        const int j = b + tm.t;
        if (j < e) {
            double x = a.wpos[3 * j], y = a.wpos[3 * j + 1], z = a.wpos[3 * j + 2];
            if (j >= sb) {
                x = x + a.wvel[3 * j] * hd;
                y = y + a.wvel[3 * j + 1] * hd;
                z = z + a.wvel[3 * j + 2] * hd;
                a.wpos[3 * j] = x;
                a.wpos[3 * j + 1] = y;
                a.wpos[3 * j + 2] = z;
            }
            double *q = tile + kKickStride * tm.t;
            q[0] = x;
            q[1] = y;
            q[2] = z;
            q[3] = a.wm[j];
        }
    }
    tm.sync();
    const bool qs = kick_small_wg_quarters<TEAM, SOFT>(a, 0, sb, e, b, e, b, h, tile, part, tm);
    const bool qf = sb > b &&
                    kick_small_wg_quarters<TEAM, SOFT>(a, 1, b, sb, sb, e, b, h, tile, part, tm);
    if (!qs && !qf) return;
    tm.sync();
    if (qs) kick_small_wg_add<TEAM>(a, 0, sb, e, h, part, tm);
    if (qf) kick_small_wg_add<TEAM>(a, 1, b, sb, h, part, tm);
}

// Sinks [sb, se) from L <= kSmallSrc sources [jb, je) with the quarter
// sums: a wave team one quarter after the other, a workgroup team from one
// tile.
template <int TEAM, bool SOFT>
__device__ __forceinline__ void team_kick_small(const EnsArgs &a, int sb, int se, int jb, int je,
                                                double h, double *tile, const Team<TEAM> &tm) {
    if constexpr (TEAM == 64) {
        kick_small_passes<64, SOFT, false>(a, sb, se, jb, je, jb, h, tile, tm);
    } else {
        float *part = reinterpret_cast<float *>(tile + kTileDoubles);
        load_kick_tile<TEAM>(a, jb, je, tile, tm);
        if (kick_small_wg_quarters<TEAM, SOFT>(a, 0, sb, se, jb, je, jb, h, tile, part, tm)) {
            tm.sync();
            kick_small_wg_add<TEAM>(a, 0, sb, se, h, part, tm);
        }
    }
}

// vel[i] += h * sum_j m_j (x_j - x_i) / (r^2 + eps2)^{3/2}, sinks [sb, se),
// sources [jb, je) (absolute indices) in index order through LDS tiles.
// Up to kSmallSrc sources: the quarter sums above.  More: summed in FP32
// over blocks of kFold sources from jb (TEAM is a multiple of kFold, so
// tile edges fall on block edges) folded into a two-sum accumulator.
template <int TEAM, bool SOFT>
__device__ __forceinline__ void team_kick(const EnsArgs &a, int sb, int se, int jb, int je, double h,
                          double *tile, const Team<TEAM> &tm) {
    static_assert(TEAM % kFold == 0, "tiles must hold whole fold blocks");
    static_assert(TEAM <= kWG, "wider teams take team_kicks_small_wg");
    jb = uni(jb);
    je = uni(je);
    if (je - jb <= kSmallSrc) {
        if (se > sb && je > jb) team_kick_small<TEAM, SOFT>(a, uni(sb), uni(se), jb, je, h, tile, tm);
        return;
    }
    for (int i0 = uni(sb); i0 < se;) {
        const int S = sinks_per_lane<TEAM>(se - i0, kMaxSpl);
        if (S >= 4)
            team_kick_pass<TEAM, 4, SOFT>(a, i0, se, jb, je, h, tile, tm);
        else if (S == 3)
            team_kick_pass<TEAM, 3, SOFT>(a, i0, se, jb, je, h, tile, tm);
        else if (S == 2)
            team_kick_pass<TEAM, 2, SOFT>(a, i0, se, jb, je, h, tile, tm);
        else
            team_kick_pass<TEAM, 1, SOFT>(a, i0, se, jb, je, h, tile, tm);
        i0 += S * TEAM;
    }
}

// E = sum 1/2 m v^2 - sum_{i<j} m_i m_j / sqrt(r^2 + eps2), FP64, fixed
// reduction order.  The team's first wave evaluates it alone (lane l takes
// particles l, l + 64, ...; the 64 lane sums are added in lane order), so
// the value does not depend on the team size: twice per system, it is a
// negligible share of the integration.
template <int TEAM>
__device__ double team_energy(const EnsArgs &a, int o, int n, TeamShared<TEAM> &sh,
                              const Team<TEAM> &tm) {
#pragma clang fp contract(off)
    if (tm.t < 64) {
        double acc = 0.0;
        const int jb = uni(o), je = uni(o + n);
        for (int i0 = jb; i0 < je; i0 += 64) {
            const int i = i0 + tm.t;
            const int ic = i < je ? i : je - 1;
            const double xi = a.wpos[3 * ic], yi = a.wpos[3 * ic + 1], zi = a.wpos[3 * ic + 2];
            double s = 0.0;
            for (int j = i0 + 1; j < je; ++j) {
                const double dx = a.wpos[3 * j] - xi, dy = a.wpos[3 * j + 1] - yi,
                             dz = a.wpos[3 * j + 2] - zi;
                const double r2 = dx * dx + dy * dy + dz * dz + a.eps2;
                const double t = a.wm[j] / sqrt(r2);
                if (j > i && r2 != 0.0) s += t;
            }
            if (i < je) {
                const double vx = a.wvel[3 * i], vy = a.wvel[3 * i + 1], vz = a.wvel[3 * i + 2];
                acc += 0.5 * a.wm[i] * (vx * vx + vy * vy + vz * vz) - a.wm[i] * s;
            }
        }
        sh.red[tm.t] = acc;
    }
    tm.sync();
    double tot = 0.0;
#pragma unroll 1
    for (int k = 0; k < 64; ++k) tot += sh.red[k];
    tm.sync();
    return tot;
}

struct SysStats {
    long long calls = 0, interactions = 0, ts_pairs = 0;
    int max_level = 0;
};

// hold_step (src/hold.cuh:51-181) on particles [o, o+n) with time step dt
// and the given calc_timestep flag, run to completion with an explicit
// stack: a root step is (dt_root, calc = 1).  Levels are counted from 0 at
// this call.  Returns HHS_OK or a status code (uniform).
//
// Thread 0 alone writes the stack.  Every thread reads the top frame at the
// head of an iteration; each stage passes a barrier before thread 0 writes
// (so every read of this iteration is done) and ends with one after (so
// the next iteration reads the new state).  A call without fast particles
// (a leaf) finishes in its stage 1.
template <int TEAM, bool SOFT, bool GUARD>
__device__ int team_hold_frame(const EnsArgs &a, int o, int n, double dt, int calc,
                               TeamShared<TEAM> &sh, double *tile, const Team<TEAM> &tm,
                               SysStats &st) {
    if (tm.t == 0) {
        sh.sp = 1;
        sh.status = HHS_OK;
        sh.stack[0] = Frame{0, n, 0, 0, dt, calc, 0};
    }
    tm.sync();
    for (;;) {
        const int sp = sh.sp;
        if (sp == 0 || sh.status != HHS_OK) break;
        const Frame f = sh.stack[sp - 1];
        const int level = sp - 1;
        if (f.stage == 0) {
            if (level > a.max_level || (a.max_calls > 0 && st.calls >= a.max_calls)) {
                tm.sync();
                if (tm.t == 0) sh.status = HHS_ERR_DEPTH;
                tm.sync();
                break;
            }
            st.calls++;
            st.max_level = max(st.max_level, level);
            if (tm.t == 0) sh.level_calls[level]++;
            if (f.calc) {  // src/hold.cuh:74-76
                team_timesteps<TEAM>(a, o, f.b, f.e, tile, tm);
                st.ts_pairs += static_cast<long long>(f.e - f.b) * (f.e - f.b - 1);
                tm.sync();
            }
            // :79-83, with the decision guard after an approximate timestep
            // evaluation
            const TsGuard gd{1.0 / f.dt, a.guard_qmax, a.guard_lo, a.guard_hi};
            const int nf = team_partition<TEAM, GUARD>(a, o, f.b, f.e, f.dt, sh, tm,
                                                       f.calc && a.guard, gd);
            if (tm.t == 0) {
                sh.stack[sp - 1].nf = nf;
                sh.stack[sp - 1].stage = 1;
                if (nf > 0) {  // :116-121
                    sh.stack[sp] = Frame{f.b, f.b + nf, 0, 0, f.dt / 2, 0, 0};
                    sh.sp = sp + 1;
                }
            }
            tm.sync();
        } else if (f.stage == 1) {
            const int sb = f.b + f.nf;
            if (f.e > sb) {
                // :123-124 and :137-145 -- slow sinks from every particle,
                // fast sinks from slow ones
                if constexpr (TEAM > kWG) {
                    team_kicks_small_wg<TEAM, SOFT>(a, o + f.b, o + sb, o + f.e, f.dt * 0.5, f.dt,
                                                    tile, tm);
                } else {
                    if (TEAM == kWG && f.e - f.b <= kSmallSrc) {
                        team_kicks_small_wg<TEAM, SOFT>(a, o + f.b, o + sb, o + f.e, f.dt * 0.5,
                                                        f.dt, tile, tm);
                    } else {
                        team_drift<TEAM>(a, o, sb, f.e, f.dt * 0.5, tm);
                        tm.sync();
                        team_kick<TEAM, SOFT>(a, o + sb, o + f.e, o + f.b, o + f.e, f.dt, tile, tm);
                        if (f.nf > 0)
                            team_kick<TEAM, SOFT>(a, o + f.b, o + sb, o + sb, o + f.e, f.dt, tile,
                                                  tm);
                    }
                }
                st.interactions += static_cast<long long>(f.e - sb) * (f.e - f.b) +
                                   static_cast<long long>(f.nf) * (f.e - sb);
            } else {
                tm.sync();
            }
            if (f.nf == 0) {
                // a leaf: :158-159 (the whole range is slow) once every
                // kick has read the positions, and the call is done
                tm.sync();
                team_drift<TEAM>(a, o, sb, f.e, f.dt * 0.5, tm);
                if (tm.t == 0) sh.sp = sp - 1;
            } else if (tm.t == 0) {
                sh.stack[sp - 1].stage = 2;
                sh.stack[sp] = Frame{f.b, sb, 0, 0, f.dt / 2, 1, 0};  // :168-173
                sh.sp = sp + 1;
            }
            tm.sync();
        } else {
            // :158-159 -- the slow particles' second half-drift, due before
            // anything outside this call reads them (the fast recursion in
            // between touches only [b, sb))
            const int sb = f.b + f.nf;
            if (f.e > sb) team_drift<TEAM>(a, o, sb, f.e, f.dt * 0.5, tm);
            tm.sync();
            if (tm.t == 0) sh.sp = sp - 1;
            tm.sync();
        }
    }
    const int status = sh.status;
    tm.sync();
    return status;
}

template <int TEAM, bool SOFT, bool GUARD>
__device__ void team_run_system(const EnsArgs &a, int s, TeamShared<TEAM> &sh, double *tile,
                                const Team<TEAM> &tm) {
    const int o = uni(a.offsets[s]);
    const int n = uni(a.offsets[s + 1]) - o;
    for (int i = tm.t; i < n; i += TEAM) {
        const int g = o + i;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a.wpos[3 * g + k] = a.pos[3 * g + k];
            a.wvel[3 * g + k] = a.vel[3 * g + k];
        }
        a.wm[g] = a.m[g];
        a.wts[g] = 0.0;
        a.wid[g] = i;
    }
    if (tm.t == 0)
        for (int l = 0; l < kMaxFrames; ++l) sh.level_calls[l] = 0;
    tm.sync();
    SysStats st;
    int status = HHS_OK;
    double e0 = 0.0, e1 = 0.0, t = 0.0;
    if (n > 0) {
        e0 = team_energy<TEAM>(a, o, n, sh, tm);
        const double t_end = a.t_end[s], dtr = a.dt_root[s];
        if (!(dtr > 0.0)) status = HHS_ERR_INVALID;
        while (status == HHS_OK && t < t_end) {
            status = team_hold_frame<TEAM, SOFT, GUARD>(a, o, n, dtr, 1, sh, tile, tm, st);
            if (status == HHS_OK) t += dtr;
        }
        e1 = team_energy<TEAM>(a, o, n, sh, tm);
        // restore the original particle order
        for (int i = tm.t; i < n; i += TEAM) {
            const int g = o + i;
            const int u = o + a.wid[g];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                a.pos[3 * u + k] = a.wpos[3 * g + k];
                a.vel[3 * u + k] = a.wvel[3 * g + k];
            }
        }
    }
    if (a.done) {
        __threadfence_system();  // this thread's pos/vel stores, before the flag
        tm.sync();
    }
    if (tm.t == 0) {
        hhs_system_result r;
        r.e0 = e0;
        r.e1 = e1;
        r.t = t;
        r.calls = st.calls;
        r.interactions = st.interactions;
        r.ts_pairs = st.ts_pairs;
        r.max_level = st.max_level;
        r.status = status;
        a.results[s] = r;
        if (a.done)
            __hip_atomic_store(&a.done[s], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    tm.sync();
}

// ------------------------------------------------------------ scheduling
// A system's cost is not a function of its size alone: one tight binary
// makes a 55-particle system recurse deep enough to cost as much as a
// 1500-particle one.  So the batch is ordered by an estimate from the
// initial state, largest first (list scheduling), and teams pull work items
// in that order: a big system on the whole workgroup, or a bundle of four
// small systems with neighbouring estimates, one per wave.
//
// Estimate: the initial timesteps give each particle the level it descends
// to (tau <= dt_root / 2^(L-1) reaches level L); with n_L particles at
// level L or deeper, a root step makes about 2^L calls on n_L particles,
// each costing ~n_L^2 pair evaluations plus a fixed overhead.
constexpr double kCallCost = 16384.0;  // per-call overhead, in pair evaluations

// (device, stream, kernel) triples whose first ensemble launch ran the
// scratch warm-up (see hhs_ensemble_run_f64)
std::mutex g_warm_mu;
std::set<std::tuple<int, hipStream_t, const void *>> g_warmed;
constexpr int kEstSinks = 1024;        // sinks sampled per system

__global__ __launch_bounds__(64) void ensemble_estimate_kernel(EnsArgs a) {
    __shared__ int hist[kMaxFrames];
    const int s = blockIdx.x, lane = threadIdx.x;
    const int o = a.offsets[s], n = a.offsets[s + 1] - o;
    for (int l = lane; l < kMaxFrames; l += 64) hist[l] = 0;
    __syncthreads();
    const double dtr = a.dt_root[s], t_end = a.t_end[s];
    const int lmax = min(a.max_level, kMaxFrames - 1);
    if (n >= 2 && dtr > 0.0) {
        const float etaf = static_cast<float>(a.eta), cff = static_cast<float>(a.c_ff);
        // At most kEstSinks sinks (every stride-th particle), each counted
        // stride times: O(n) work per sink keeps large systems cheap.
        const int stride = (n + kEstSinks - 1) / kEstSinks;
        const int ns = (n + stride - 1) / stride;
        for (int q0 = 0; q0 < ns; q0 += 64) {
            const int q = q0 + lane;
            const int i = (q < ns ? q : ns - 1) * stride;
            const int ic = o + i;
            const double xi = a.pos[3 * ic], yi = a.pos[3 * ic + 1], zi = a.pos[3 * ic + 2];
            const double vxi = a.vel[3 * ic], vyi = a.vel[3 * ic + 1], vzi = a.vel[3 * ic + 2];
            const float mi = static_cast<float>(a.m[ic]);
            float bnum = static_cast<float>(TsConst<double>::huge), bden = 1.0f;
            for (int j = o; j < o + n; ++j)
                pair_timestep_f32(xi, yi, zi, vxi, vyi, vzi, mi, a.pos[3 * j], a.pos[3 * j + 1],
                                  a.pos[3 * j + 2], a.vel[3 * j], a.vel[3 * j + 1],
                                  a.vel[3 * j + 2], static_cast<float>(a.m[j]), etaf, cff,
                                  a.eps2f, j != ic, bnum, bden);
            if (q < ns) {
                const double tau = static_cast<double>(bnum) / static_cast<double>(bden);
                int l = 0;
                for (double dt = dtr; l < lmax && tau <= dt; dt *= 0.5) ++l;
                atomicAdd(&hist[l], stride);
            }
        }
    }
    __syncthreads();
    if (lane == 0) {
        double work = 0.0;
        int nl = 0;
        for (int l = kMaxFrames - 1; l >= 0; --l) {
            nl += hist[l];
            if (nl >= 2) work += ldexp(1.0, l) * (static_cast<double>(nl) * nl + kCallCost);
        }
        const double steps = (dtr > 0.0 && t_end > 0.0) ? ceil(t_end / dtr) : 0.0;
        a.est[s] = work * steps + n;
    }
}

// Rank of item k among all items by (estimate descending, index ascending).
__device__ __forceinline__ bool before(double ej, int j, double ek, int k) {
    return ej > ek || (ej == ek && j < k);
}

__global__ __launch_bounds__(256) void rank_small_kernel(EnsArgs a) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= a.n_small) return;
    const double ek = a.est[a.small_list[k]];
    int r = 0;
    for (int j = 0; j < a.n_small; ++j) r += before(a.est[a.small_list[j]], j, ek, k) ? 1 : 0;
    a.small_sorted[r] = a.small_list[k];
}

__device__ __forceinline__ double item_est(const EnsArgs &a, int k) {
    return k < a.n_big ? a.est[a.big_list[k]] : a.est[a.small_sorted[4 * (k - a.n_big)]];
}

__global__ __launch_bounds__(256) void rank_items_kernel(EnsArgs a) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= a.n_items) return;
    const double ek = item_est(a, k);
    int r = 0;
    for (int j = 0; j < a.n_items; ++j) r += before(item_est(a, j), j, ek, k) ? 1 : 0;
    a.queue[r] = k < a.n_big ? a.big_list[k] : -1 - (k - a.n_big);
}

// Three workgroups per CU (<= 168 VGPRs): the batch is throughput-bound,
// and two per CU measured 5-10 % slower.  flatten inlines the whole engine
// (both team sizes), which the register allocator needs to stay within
// that budget without spilling.
template <bool SOFT, bool GUARD>
// Four waves per SIMD (<= 128 VGPRs, four workgroups per CU; the LDS of
// four, 4 x 40 KB, still fits): the compiler then spills ~200 bytes per lane
// to scratch, but the batch runs faster than at three waves per SIMD with
// no spills (MI355X, 4096-system workload: 5.24-5.55 s against 6.39-6.47 s,
// same bits).  Fewer sinks per lane to avoid the spills lost more than the
// occupancy gained (two: 6.8 s, one: 7.3 s).
__global__ __launch_bounds__(kWG, 4) __attribute__((flatten)) void ensemble_kernel(EnsArgs a) {
    __shared__ TeamShared<kWG> sh_wg;
    __shared__ TeamShared<64> sh[kWaves];
    // Source tiles: the whole array for a workgroup team (kWG sources), a
    // quarter per wave team (64 sources each).
    __shared__ double tiles[kTileDoubles + kPartDoubles<kWG>];
    __shared__ int s_item;
    const int w = threadIdx.x >> 6;
    for (;;) {
        if (threadIdx.x == 0)
            s_item = static_cast<int>(__hip_atomic_fetch_add(&a.counters[0], 1u, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT));
        __syncthreads();
        const int k = s_item;
        __syncthreads();
        if (k >= a.n_items) break;
        const int code = a.queue[k];
        if (code >= 0) {
            Team<kWG> tm;
            team_run_system<kWG, SOFT, GUARD>(a, code, sh_wg, tiles, tm);
        } else {
            const int idx = 4 * (-1 - code) + w;
            if (idx < a.n_small) {
                Team<64> tw;
                team_run_system<64, SOFT, GUARD>(a, a.small_sorted[idx], sh[w],
                                          tiles + w * 64 * kTsStride, tw);
            }
            __syncthreads();  // the bundle is done before the tiles change hands
        }
    }
}

// Subtrees of at most kSubLds particles keep their working and scratch
// copies in LDS (2 x 68 B per particle, 34 KB at 256): the subtree is the
// only work on the GPU while it runs, so each of a call's dependent memory
// round trips (tile loads, partition, drifts, velocity updates) costs LDS
// latency instead of L2 latency.  The engine reaches them through generic
// pointers; the arithmetic and its order are those of the global-memory
// variant, so both give the same bits.  They run on a kSubTeam-thread
// team, whose every call has <= kSmallSrc sources: each of a call's source
// quarters gets a wave of its own (four waves per SIMD instead of one).
constexpr int kSubLds = 256;
static_assert(kSubLds <= kSmallSrc, "an LDS subtree's calls take the quarter-split path");

template <bool SOFT, bool LDS, bool GUARD>
__global__ __launch_bounds__(LDS ? kSubTeam : kWG) __attribute__((flatten)) void subtree_kernel(
    EnsArgs a, Particles<double> total, int n, int clevel, double dt, int calc, SubtreeAcc *acc,
    const int *n_dev, KickPending kp) {
    constexpr int TEAM = LDS ? kSubTeam : kWG;
    if (n_dev) {  // speculative: the count an earlier kernel wrote, n its bound
        const int nd = __builtin_amdgcn_readfirstlane(*n_dev);
        if (nd < 1 || nd > n) return;
        n = nd;
    }
    __shared__ TeamShared<TEAM> sh;
    __shared__ double tile[kTileDoubles + kPartDoubles<TEAM>];
    constexpr int C = LDS ? kSubLds : 1;
    __shared__ double l_pos[2][3 * C], l_vel[2][3 * C], l_m[2][C], l_ts[2][C];
    __shared__ int l_id[2][C];
    if constexpr (LDS) {
        a.wpos = l_pos[0];
        a.wvel = l_vel[0];
        a.wm = l_m[0];
        a.wts = l_ts[0];
        a.wid = l_id[0];
        a.tpos = l_pos[1];
        a.tvel = l_vel[1];
        a.tm = l_m[1];
        a.tts = l_ts[1];
        a.tid = l_id[1];
    }
    Team<TEAM> tm;
    for (int i = tm.t; i < n; i += TEAM) {
        // the set's pending kick (the parent call's split-K partials), as
        // kick_reduce_kernel would have applied it
        double kv[3] = {0.0, 0.0, 0.0};
        if (kp.part) {
            sum_kick_partials(kp.part, kp.K, kp.rows, i, kv[0], kv[1], kv[2]);
#pragma unroll
            for (int c = 0; c < 3; ++c) kv[c] = __dmul_rn(kv[c], kp.h);
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            a.wpos[3 * i + k] = total.pos[3 * i + k];
            a.wvel[3 * i + k] = kp.part ? __dadd_rn(total.vel[3 * i + k], kv[k])
                                        : total.vel[3 * i + k];
        }
        a.wm[i] = total.m[i];
        a.wts[i] = total.timestep[i];
        a.wid[i] = i;
    }
    for (int l = tm.t; l < kMaxFrames; l += TEAM) sh.level_calls[l] = 0;
    tm.sync();
    SysStats st;
    const int status = team_hold_frame<TEAM, SOFT, GUARD>(a, 0, n, dt, calc, sh, tile, tm, st);
    for (int g = tm.t; g < n; g += TEAM) {
        const int u = a.wid[g];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            total.pos[3 * u + k] = a.wpos[3 * g + k];
            total.vel[3 * u + k] = a.wvel[3 * g + k];
        }
        total.m[u] = a.wm[g];
        total.timestep[u] = a.wts[g];
    }
    // the statistics, added to the accumulator (absolute levels)
    if (tm.t == 0) {
        acc->calls += st.calls;
        acc->interactions += st.interactions;
        acc->ts_pairs += st.ts_pairs;
        acc->max_level = max(acc->max_level, clevel + st.max_level);
        if (status != HHS_OK && acc->status == HHS_OK) acc->status = status;
    }
    for (int l = tm.t; l < kMaxFrames && clevel + l < 64; l += TEAM)
        acc->level_calls[clevel + l] += sh.level_calls[l];
}

__global__ void zero_counters_kernel(unsigned int *c) {
    for (int i = threadIdx.x; i < kCounters; i += blockDim.x) c[i] = 0u;
}

size_t align256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

struct EnsLayout {
    size_t counters, big, small, est, small_sorted, queue, wpos, wvel, wm, wts, wid, tpos, tvel, tm, tts, tid, total;
};

EnsLayout layout(int nsys, long long np) {
    EnsLayout L;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t r = off;
        off += align256(bytes);
        return r;
    };
    const size_t n = static_cast<size_t>(np > 0 ? np : 1);
    const size_t s = static_cast<size_t>(nsys > 0 ? nsys : 1);
    L.counters = take(4 * kCounters);
    L.big = take(s * 4);
    L.small = take(s * 4);
    L.est = take(s * 8);
    L.small_sorted = take(s * 4);
    L.queue = take(s * 4);
    L.wpos = take(3 * n * 8);
    L.wvel = take(3 * n * 8);
    L.wm = take(n * 8);
    L.wts = take(n * 8);
    L.wid = take(n * 4);
    L.tpos = take(3 * n * 8);
    L.tvel = take(3 * n * 8);
    L.tm = take(n * 8);
    L.tts = take(n * 8);
    L.tid = take(n * 4);
    L.total = off;
    return L;
}

}  // namespace

size_t subtree_workspace_bytes(int n_max) {
    const EnsLayout L = layout(1, n_max);
    return L.total;
}

int launch_subtree(const Particles<double> &total, int n, int clevel, double dt, int calc,
                   double eta, double eps2, int ts_mode, int max_level, void *ws,
                   size_t ws_bytes, SubtreeAcc *acc_dev, hipStream_t st, const int *n_dev,
                   const KickPending &kick) {
    HHS_REQUIRE(n > 0 && clevel >= 0 && max_level >= 0 && max_level <= kMaxFrames - 3 && acc_dev,
                "subtree: n=%d, clevel=%d, max_level=%d", n, clevel, max_level);
    HHS_REQUIRE(!kick.part || (!n_dev && kick.rows == n),
                "subtree: a pending kick needs the host-known set size");
    // a speculative launch must pick the variant every count up to n picks
    HHS_REQUIRE(!n_dev || n <= kSubLds, "subtree: a device-side count needs n <= %d", kSubLds);
    const EnsLayout L = layout(1, n);
    HHS_REQUIRE(ws && ws_bytes >= L.total, "subtree: workspace of %zu bytes, need %zu",
                ws_bytes, L.total);
    char *w = static_cast<char *>(ws);
    EnsArgs a{};
    a.eta = eta;
    a.c_ff = eta / TsConst<double>::sqrt2;
    a.eps2 = eps2;
    a.eps2f = static_cast<float>(eps2);
    a.max_level = max_level;
    a.max_calls = 0;
    a.wpos = reinterpret_cast<double *>(w + L.wpos);
    a.wvel = reinterpret_cast<double *>(w + L.wvel);
    a.wm = reinterpret_cast<double *>(w + L.wm);
    a.wts = reinterpret_cast<double *>(w + L.wts);
    a.wid = reinterpret_cast<int *>(w + L.wid);
    a.tpos = reinterpret_cast<double *>(w + L.tpos);
    a.tvel = reinterpret_cast<double *>(w + L.tvel);
    a.tm = reinterpret_cast<double *>(w + L.tm);
    a.tts = reinterpret_cast<double *>(w + L.tts);
    a.tid = reinterpret_cast<int *>(w + L.tid);
    a.ts_mode = ts_mode;
    set_guard(a, ts_mode);
    const bool lds = n <= kSubLds;
    // SOFT: eps^2 > 0 in FP32, so r^2 > 0 in every FP32 pair formula.
    // GUARD: the decision guard of the approximate timestep modes, always on
    // here, so that a subtree decides as the host recursion does.
    const bool soft = static_cast<float>(eps2) > 0.0f;
    auto kern = a.guard ? (soft ? (lds ? subtree_kernel<true, true, true>
                                       : subtree_kernel<true, false, true>)
                                : (lds ? subtree_kernel<false, true, true>
                                       : subtree_kernel<false, false, true>))
                        : (soft ? (lds ? subtree_kernel<true, true, false>
                                       : subtree_kernel<true, false, false>)
                                : (lds ? subtree_kernel<false, true, false>
                                       : subtree_kernel<false, false, false>));
    hipLaunchKernelGGL(kern, dim3(1), dim3(lds ? kSubTeam : kWG), 0, st, a, total, n, clevel, dt,
                       calc, acc_dev, n_dev, kick);
    HHS_LAUNCH_CHECK("subtree_kernel");
    return HHS_OK;
}

}  // namespace hhs

using namespace hhs;

extern "C" {

size_t hhs_ensemble_workspace_bytes(int n_systems, long long n_particles) {
    return layout(n_systems, n_particles).total;
}

int hhs_ensemble_run_f64(const hhs_ensemble_desc *d, const int *offsets_host, void *ws,
                         size_t ws_bytes, void *stream) {
    HHS_REQUIRE(d && offsets_host, "ensemble: null descriptor or offsets");
    const int nsys = d->n_systems;
    HHS_REQUIRE(nsys >= 0, "ensemble: negative system count");
    if (nsys == 0) return HHS_OK;
    HHS_REQUIRE(d->offsets && d->pos && d->vel && d->m && d->t_end && d->dt_root && d->results,
                "ensemble: null device array in descriptor");
    HHS_REQUIRE(d->eta > 0.0 && d->eps >= 0.0, "ensemble: need eta > 0 and eps >= 0");
    const int max_level = d->max_level > 0 ? d->max_level : 47;
    HHS_REQUIRE(max_level <= kMaxFrames - 3, "ensemble: max_level must be <= %d",
                kMaxFrames - 3);
    HHS_REQUIRE(offsets_host[0] == 0, "ensemble: offsets must start at 0");
    for (int s = 0; s < nsys; ++s)
        HHS_REQUIRE(offsets_host[s + 1] >= offsets_host[s], "ensemble: offsets not monotone");
    const long long np = offsets_host[nsys];
    const EnsLayout L = layout(nsys, np);
    HHS_REQUIRE(ws && ws_bytes >= L.total, "ensemble: workspace of %zu bytes, need %zu", ws_bytes,
                L.total);
    // One-wave teams up to 512 particles: the MI355X sweep of the
    // 4096-system workload gave 6.4-6.7 s against 7.8-9.4 s with 256.
    // Larger workgroup teams (512 or 1024 threads per big system) were
    // slower (7.1 s, 10.7 s): the recursion's per-call synchronisation, not
    // the lane count, bounds a big system.
    const int wave_max = d->wave_team_max_n > 0 ? d->wave_team_max_n : 512;

    // Whole-workgroup systems and one-wave systems, in index order (the
    // device orders the work by estimated cost).
    std::vector<int> big, small;
    for (int s = 0; s < nsys; ++s)
        (offsets_host[s + 1] - offsets_host[s] > wave_max ? big : small).push_back(s);

    char *w = static_cast<char *>(ws);
    hipStream_t st = as_stream(stream);
    // Ordered on `st`: the workspace may come from the stream-ordered
    // allocator on that stream.  The sources are pageable, so each copy has
    // read them before it returns.
    if (!big.empty())
        HHS_HIP(hipMemcpyAsync(w + L.big, big.data(), big.size() * 4, hipMemcpyHostToDevice, st));
    if (!small.empty())
        HHS_HIP(hipMemcpyAsync(w + L.small, small.data(), small.size() * 4, hipMemcpyHostToDevice,
                               st));
    unsigned int *counters = reinterpret_cast<unsigned int *>(w + L.counters);
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(256), 0, st, counters);
    HHS_LAUNCH_CHECK("zero_counters_kernel");

    EnsArgs a;
    a.offsets = d->offsets;
    a.pos = d->pos;
    a.vel = d->vel;
    a.m = d->m;
    a.t_end = d->t_end;
    a.dt_root = d->dt_root;
    a.eta = d->eta;
    a.c_ff = d->eta / TsConst<double>::sqrt2;
    a.eps2 = d->eps * d->eps;
    a.eps2f = static_cast<float>(a.eps2);
    a.max_level = max_level;
    a.max_calls = d->max_calls;
    a.wpos = reinterpret_cast<double *>(w + L.wpos);
    a.wvel = reinterpret_cast<double *>(w + L.wvel);
    a.wm = reinterpret_cast<double *>(w + L.wm);
    a.wts = reinterpret_cast<double *>(w + L.wts);
    a.wid = reinterpret_cast<int *>(w + L.wid);
    a.tpos = reinterpret_cast<double *>(w + L.tpos);
    a.tvel = reinterpret_cast<double *>(w + L.tvel);
    a.tm = reinterpret_cast<double *>(w + L.tm);
    a.tts = reinterpret_cast<double *>(w + L.tts);
    a.tid = reinterpret_cast<int *>(w + L.tid);
    a.big_list = reinterpret_cast<const int *>(w + L.big);
    a.n_big = static_cast<int>(big.size());
    a.small_list = reinterpret_cast<const int *>(w + L.small);
    a.n_small = static_cast<int>(small.size());
    a.est = reinterpret_cast<double *>(w + L.est);
    a.small_sorted = reinterpret_cast<int *>(w + L.small_sorted);
    a.queue = reinterpret_cast<int *>(w + L.queue);
    a.n_items = a.n_big + (a.n_small + 3) / 4;
    a.counters = counters;
    a.results = d->results;
    a.done = d->done;
    HHS_REQUIRE(d->timestep_mode == HHS_TIMESTEP_EXACT || d->timestep_mode == HHS_TIMESTEP_FAST ||
                    d->timestep_mode == HHS_TIMESTEP_FP32,
                "ensemble: bad timestep_mode %d", d->timestep_mode);
    a.ts_mode = d->timestep_mode;
    set_guard(a, d->approximate_decisions ? HHS_TIMESTEP_EXACT : a.ts_mode);

    // compute units of the device (the persistent grid's cap below)
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            v > 0)
            cus = v;
    }
    // SOFT: eps^2 > 0 in FP32, so r^2 > 0 in every FP32 pair formula; GUARD:
    // the decision guard (approximate timestep modes, unless the descriptor
    // asks for approximate decisions).  The guard is a template parameter:
    // compiled into the flattened engine behind a run-time test it slowed
    // every mode down (MI355X, 4096-system batch: 7.7-8.0 s against 6.1-6.4
    // s, also in the exact mode where it never runs); as a separate
    // instantiation both run at 6.2-6.6 s.
    auto kern = a.eps2f > 0.0f ? (a.guard ? ensemble_kernel<true, true> : ensemble_kernel<true, false>)
                               : (a.guard ? ensemble_kernel<false, true> : ensemble_kernel<false, false>);
    // Persistent grid: enough workgroups for every big system and every
    // group of four small ones, capped at 8 per CU.
    const int grid = std::max(1, std::min(a.n_items, 8 * cus));
    // Warm-up: the kernel spills to scratch, and the first dispatch that
    // needs scratch on a hardware queue ran the whole batch slowly (7.3-8.7
    // s against 5.2-5.3 s for later ones on the same queue; a one-workgroup
    // warm-up left 5.7-6.7 s).  So the first launch of each kernel variant
    // on a stream is preceded by the largest grid with no items: every
    // workgroup takes a ticket >= 0 = n_items and leaves, within
    // microseconds, and the tickets it drew are cleared again.  Streams
    // share a few hardware queues, so some warm-ups are redundant; a stream
    // created at a freed stream's address on a cold queue only runs its
    // first batch slowly.
    bool warm = false;
    {
        std::lock_guard<std::mutex> lk(g_warm_mu);
        if (g_warmed.size() >= 4096) g_warmed.clear();  // streams come and go: bound the record
        warm = g_warmed.insert({dev, st, reinterpret_cast<const void *>(kern)}).second;
    }
    if (warm) {
        EnsArgs z = a;
        z.n_items = 0;
        hipLaunchKernelGGL(kern, dim3(8 * cus), dim3(kWG), 0, st, z);
        HHS_LAUNCH_CHECK("ensemble_kernel (warm-up)");
        hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(256), 0, st, counters);
        HHS_LAUNCH_CHECK("zero_counters_kernel");
    }
    // Cost estimates, then the work queue in estimate order.
    hipLaunchKernelGGL(ensemble_estimate_kernel, dim3(nsys), dim3(64), 0, st, a);
    HHS_LAUNCH_CHECK("ensemble_estimate_kernel");
    if (a.n_small > 0) {
        hipLaunchKernelGGL(rank_small_kernel, dim3((a.n_small + 255) / 256), dim3(256), 0, st, a);
        HHS_LAUNCH_CHECK("rank_small_kernel");
    }
    hipLaunchKernelGGL(rank_items_kernel, dim3((a.n_items + 255) / 256), dim3(256), 0, st, a);
    HHS_LAUNCH_CHECK("rank_items_kernel");
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kWG), 0, st, a);
    HHS_LAUNCH_CHECK("ensemble_kernel");
    return HHS_OK;
}

}  // extern "C"
