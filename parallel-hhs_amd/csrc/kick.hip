// Direct-sum kick for gfx950.
//
// Reference: _kick_slow / _kick_sf and pp_force (src/kick.cuh:15-155),
// with the attractive sign of tests/test_nbforce.cu:14-18 and Plummer
// softening r^2 -> r^2 + eps^2.
//
// Design (CDNA4, wave64):
//  * one lane owns S sinks (register blocking: each source is reused S
//    times per lane);
//  * the source loop index is wave-uniform, so source coordinates and masses
//    are fetched with scalar loads into SGPRs and used directly as VALU
//    operands -- a wave-wide broadcast with no LDS traffic and no barriers;
//  * W waves of one workgroup share the same 64*S sinks and split the
//    source range; their partial sums are combined through LDS in a fixed
//    order (results do not depend on scheduling);
//  * separations come from FP64 positions (x_j - x_i in FP64, rounded once
//    to FP32), then r^2, v_rsq_f32, m r^-3 and the accumulation run in FP32;
//  * forces are summed in plain FP32 over blocks of kFold sources and each
//    block partial is folded into a (hi, lo) double-single accumulator with
//    Knuth's two-sum, so the rounding error of the long sum stays at the
//    level of one block.
#include <atomic>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hhs_internal.h"
#include "pair_math.h"

namespace hhs {
namespace {

constexpr int kFold = 64;  // sources per plain-FP32 block before a two-sum fold

// Per-sink force accumulators for the three precision modes.
//  P = 0: separations x_j - x_i in FP64 rounded once to FP32; r^2, rsqrt,
//         m r^-3 and the sum in FP32, summed plainly over blocks of kFold
//         sources and folded into a (hi, lo) two-sum accumulator.
//         15 VALU ops + 1 v_rsq_f32 per interaction (FP64 storage).
//  P = 1: separations, r^2 and the sum in FP64; v_rsq_f32 and the weight
//         m r^-3 in FP32.  14 VALU ops + 1 v_rsq_f32.
//  P = 2: FP64 throughout with the hardware reciprocal-square-root estimate
//         v_rsq_f64 (relative error <= 2^-23, FP32-level); 12 VALU ops +
//         1 v_rsq_f64.  The weight carries ~3 x 2^-23 relative error per
//         pair term, inside the 1e-6 force bar.
template <typename T, int S, int P, bool SOFT> struct KickAcc;

template <typename T, int S, bool SOFT> struct KickAcc<T, S, 0, SOFT> {
    float hx[S], lx[S], hy[S], ly[S], hz[S], lz[S];
    __device__ KickAcc() {
#pragma unroll
        for (int s = 0; s < S; ++s) hx[s] = lx[s] = hy[s] = ly[s] = hz[s] = lz[s] = 0.0f;
    }
    // mf: null, or the source masses converted to FP32 up front (the loop
    // then spends no VALU slot per source on the conversion); the values are
    // the same either way.
    __device__ void range(const T *__restrict__ sp, const T *__restrict__ sm, const float *mf,
                          int jb, int je, const T (&xi)[S], const T (&yi)[S], const T (&zi)[S],
                          float eps2f, double) {
        if (mf)
            fold<true>(sp, sm, mf, jb, je, xi, yi, zi, eps2f);
        else
            fold<false>(sp, sm, mf, jb, je, xi, yi, zi, eps2f);
    }
    template <bool PRE>
    __device__ void fold(const T *__restrict__ sp, const T *__restrict__ sm,
                         const float *__restrict__ mf, int jb, int je, const T (&xi)[S],
                         const T (&yi)[S], const T (&zi)[S], float eps2f) {
        for (int j0 = jb; j0 < je; j0 += kFold) {
            const int j1 = min(j0 + kFold, je);
            float px[S], py[S], pz[S];
#pragma unroll
            for (int s = 0; s < S; ++s) px[s] = py[s] = pz[s] = 0.0f;
#pragma unroll 4
            for (int j = j0; j < j1; ++j) {
                const T xj = sp[3 * j + 0], yj = sp[3 * j + 1], zj = sp[3 * j + 2];
                const float mj = PRE ? mf[j] : static_cast<float>(sm[j]);
#pragma unroll
                for (int s = 0; s < S; ++s)
                    pair_accel<SOFT>(xj, yj, zj, mj, xi[s], yi[s], zi[s], eps2f, px[s], py[s],
                                     pz[s]);
            }
#pragma unroll
            for (int s = 0; s < S; ++s) {
                two_sum_add(hx[s], lx[s], px[s]);
                two_sum_add(hy[s], ly[s], py[s]);
                two_sum_add(hz[s], lz[s], pz[s]);
            }
        }
    }
    // debiased for v_rsq_f32's mean error (kRsqDebias)
    __device__ void total(int s, double &ax, double &ay, double &az) const {
        ax = (static_cast<double>(hx[s]) + static_cast<double>(lx[s])) * kRsqDebias;
        ay = (static_cast<double>(hy[s]) + static_cast<double>(ly[s])) * kRsqDebias;
        az = (static_cast<double>(hz[s]) + static_cast<double>(lz[s])) * kRsqDebias;
    }
};

template <int S, int P, bool SOFT> struct KickAcc64 {
    double ax[S], ay[S], az[S];
    __device__ KickAcc64() {
#pragma unroll
        for (int s = 0; s < S; ++s) ax[s] = ay[s] = az[s] = 0.0;
    }
    __device__ void range(const double *__restrict__ sp, const double *__restrict__ sm,
                          const float *, int jb, int je, const double (&xi)[S],
                          const double (&yi)[S], const double (&zi)[S], float, double eps2) {
#pragma unroll 4
        for (int j = jb; j < je; ++j) {
            const double xj = sp[3 * j + 0], yj = sp[3 * j + 1], zj = sp[3 * j + 2];
            const double mj = sm[j];
            const float mjf = static_cast<float>(mj);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                const double dx = xj - xi[s], dy = yj - yi[s], dz = zj - zi[s];
                const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, __builtin_fma(dz, dz, eps2)));
                double w;
                if constexpr (P == 1) {
                    const float rinv = __builtin_amdgcn_rsqf(static_cast<float>(r2));
                    w = static_cast<double>((mjf * rinv) * (rinv * rinv));
                } else {
                    const double rinv = __builtin_amdgcn_rsq(r2);
                    w = (mj * rinv) * (rinv * rinv);
                }
                if (!SOFT) w = (r2 > 0.0) ? w : 0.0;  // src/kick.cuh:25 (skip d2 == 0)
                ax[s] = __builtin_fma(w, dx, ax[s]);
                ay[s] = __builtin_fma(w, dy, ay[s]);
                az[s] = __builtin_fma(w, dz, az[s]);
            }
        }
    }
    // P = 1 is debiased for v_rsq_f32's mean error (kRsqDebias)
    __device__ void total(int s, double &x, double &y, double &z) const {
        const double c = P == 1 ? kRsqDebias : 1.0;
        x = ax[s] * c;
        y = ay[s] * c;
        z = az[s] * c;
    }
};

template <int S, bool SOFT> struct KickAcc<double, S, 1, SOFT> : KickAcc64<S, 1, SOFT> {};
template <int S, bool SOFT> struct KickAcc<double, S, 2, SOFT> : KickAcc64<S, 2, SOFT> {};

//  P = 3: FP64 in the CPU direct sum's operation order with IEEE sqrt and
//         divide; launched with one wave per group so each sink sums its
//         sources in index order (bit-exact with the FP64 CPU sum).
template <int S, bool SOFT> struct KickAcc<double, S, 3, SOFT> {
    double ax[S], ay[S], az[S];
    __device__ KickAcc() {
#pragma unroll
        for (int s = 0; s < S; ++s) ax[s] = ay[s] = az[s] = 0.0;
    }
    __device__ void range(const double *__restrict__ sp, const double *__restrict__ sm,
                          const float *, int jb, int je, const double (&xi)[S],
                          const double (&yi)[S], const double (&zi)[S], float, double eps2) {
#pragma unroll 2
        for (int j = jb; j < je; ++j) {
            const double xj = sp[3 * j + 0], yj = sp[3 * j + 1], zj = sp[3 * j + 2];
            const double mj = sm[j];
#pragma unroll
            for (int s = 0; s < S; ++s)
                pair_accel_exact(xj, yj, zj, mj, xi[s], yi[s], zi[s], eps2, ax[s], ay[s], az[s]);
        }
    }
    __device__ void total(int s, double &x, double &y, double &z) const {
        x = ax[s];
        y = ay[s];
        z = az[s];
    }
};

// One workgroup: 64*S sinks (S per lane) against one source chunk of K
// (split-K: small sink counts get more waves over the same sources), W
// waves splitting the chunk (segment 0 then segment 1), per-wave totals
// merged through LDS in wave order (deterministic).  With K == 1 the
// epilogue writes acc <- a, vel += a * h; otherwise the chunk's FP64 partial
// sums go to partial[k][i][3] for kick_reduce_kernel.  Workgroups past the
// first group's tiles_a * K_a belong to the optional second group (sinks
// gb_*, segment 0 only) with its own chunk count K_b: in HOLD the two
// groups have opposite shapes (few slow sinks over all sources, many fast
// sinks over the few slow sources).  Each group's workgroups run chunk-major
// (those in flight together share a chunk); the second group's partials
// start after the first group's.
template <typename T, int S, int W, int P, bool SOFT>
__global__ __launch_bounds__(64 * W) void kick_kernel(
    const T *__restrict__ ga_pos, T *__restrict__ ga_vel, T *__restrict__ ga_acc, int n_sink,
    const T *__restrict__ gb_pos, T *__restrict__ gb_vel, T *__restrict__ gb_acc, int n_sink_b,
    const T *__restrict__ s0_pos, const T *__restrict__ s0_m, int n0,
    const T *__restrict__ s1_pos, const T *__restrict__ s1_m, int n1_a, T h, float eps2f,
    double eps2, int K_a, int K_b, double *__restrict__ partial, const float *__restrict__ mf) {
    __shared__ double red[(W > 1 ? W - 1 : 1) * 3 * S * 64];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tiles_a = (n_sink + 64 * S - 1) / (64 * S);
    const int tiles_b = (n_sink_b + 64 * S - 1) / (64 * S);
    const int wg_a = tiles_a * K_a;
    const bool grp_b = static_cast<int>(blockIdx.x) >= wg_a;
    const int bid = static_cast<int>(blockIdx.x) - (grp_b ? wg_a : 0);
    const int ntile = grp_b ? tiles_b : tiles_a;
    const int K = grp_b ? K_b : K_a;
    const int chunk_id = bid / ntile;
    const T *si_pos = grp_b ? gb_pos : ga_pos;
    T *si_vel = grp_b ? gb_vel : ga_vel;
    T *si_acc = grp_b ? gb_acc : ga_acc;
    const int n1 = grp_b ? 0 : n1_a;
    // this group's partials: [K][rows of the group][3]
    double *gpart = (partial && K > 1)
                        ? partial + (grp_b ? 3 * static_cast<size_t>(K_a > 1 ? K_a : 0) * n_sink : 0)
                        : nullptr;
    if (grp_b) n_sink = n_sink_b;
    const int tile0 = (bid - chunk_id * ntile) * (64 * S);

    T xi[S], yi[S], zi[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        int i = tile0 + s * 64 + lane;
        i = i < n_sink ? i : n_sink - 1;
        xi[s] = si_pos[3 * i + 0];
        yi[s] = si_pos[3 * i + 1];
        zi[s] = si_pos[3 * i + 2];
    }
    KickAcc<T, S, P, SOFT> acc;
    // This workgroup's chunk of the concatenated source list, then this
    // wave's contiguous share of it.
    const int total = n0 + n1;
    const int per = (total + K - 1) / K;
    const int kb = min(chunk_id * per, total);
    const int ke = min(kb + per, total);
    const int chunk = (ke - kb + W - 1) / W;
    const int jb = min(kb + wave * chunk, ke);
    const int je = min(jb + chunk, ke);
    // FP32 masses converted up front (segment 1 follows segment 0), or null.
    const float *mf0 = mf, *mf1 = mf ? mf + n0 : nullptr;
    if (jb < n0) acc.range(s0_pos, s0_m, mf0, jb, min(je, n0), xi, yi, zi, eps2f, eps2);
    if (je > n0)
        acc.range(s1_pos, s1_m, mf1, max(jb, n0) - n0, je - n0, xi, yi, zi, eps2f, eps2);

    double ax[S], ay[S], az[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc.total(s, ax[s], ay[s], az[s]);
    if constexpr (W > 1) {
        if (wave > 0) {
            double *r = red + (wave - 1) * 3 * S * 64;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                r[(0 * S + s) * 64 + lane] = ax[s];
                r[(1 * S + s) * 64 + lane] = ay[s];
                r[(2 * S + s) * 64 + lane] = az[s];
            }
        }
        __syncthreads();
        if (wave != 0) return;
        for (int w = 1; w < W; ++w) {  // fixed order: deterministic
            const double *r = red + (w - 1) * 3 * S * 64;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                ax[s] += r[(0 * S + s) * 64 + lane];
                ay[s] += r[(1 * S + s) * 64 + lane];
                az[s] += r[(2 * S + s) * 64 + lane];
            }
        }
    }
    if (gpart) {
        double *q = gpart + 3 * static_cast<size_t>(chunk_id) * n_sink;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int i = tile0 + s * 64 + lane;
            if (i < n_sink) {
                q[3 * i + 0] = ax[s];
                q[3 * i + 1] = ay[s];
                q[3 * i + 2] = az[s];
            }
        }
        return;
    }
    // Epilogue (src/kick.cuh:98-104): acc <- a, vel += a * (fac * dt), as a
    // multiply then an add (no contraction, like the FP64 CPU update).
    {
#pragma clang fp contract(off)
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int i = tile0 + s * 64 + lane;
        if (i < n_sink) {
            const T a0 = static_cast<T>(ax[s]), a1 = static_cast<T>(ay[s]),
                    a2 = static_cast<T>(az[s]);
            if (si_acc) {
                si_acc[3 * i + 0] = a0;
                si_acc[3 * i + 1] = a1;
                si_acc[3 * i + 2] = a2;
            }
            si_vel[3 * i + 0] = si_vel[3 * i + 0] + a0 * h;
            si_vel[3 * i + 1] = si_vel[3 * i + 1] + a1 * h;
            si_vel[3 * i + 2] = si_vel[3 * i + 2] + a2 * h;
        }
    }
    }
}

// Interaction count from which MIXED32 kicks read FP32 masses converted once
// per launch instead of converting each source mass in every tile.
constexpr double kConvMinWork = 536870912.0; // 2^29 interactions

// Source masses of both segments rounded once to FP32 (segment 1 after
// segment 0).
template <typename T>
__global__ __launch_bounds__(256) void masses_to_f32_kernel(const T *__restrict__ m0, int n0,
                                                            const T *__restrict__ m1, int n1,
                                                            float *__restrict__ out) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < n0)
        out[j] = static_cast<float>(m0[j]);
    else if (j < n0 + n1)
        out[j] = static_cast<float>(m1[j - n0]);
}

// Sum of the K chunk partials of each sink in chunk order (deterministic),
// then the epilogue of kick_kernel, for the groups split over chunks (K_a,
// K_b > 1; a group with one chunk was written by kick_kernel itself): rows
// [0, ra) are the first group's, the next rb the second's.
template <typename T>
__global__ __launch_bounds__(256) void kick_reduce_kernel(T *__restrict__ ga_vel,
                                                          T *__restrict__ ga_acc, int n_sink,
                                                          T *__restrict__ gb_vel,
                                                          T *__restrict__ gb_acc, int n_sink_b,
                                                          const double *__restrict__ partial,
                                                          int K_a, int K_b, T h) {
#pragma clang fp contract(off)
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int ra = K_a > 1 ? n_sink : 0, rb = K_b > 1 ? n_sink_b : 0;
    if (r >= ra + rb) return;
    const bool grp_b = r >= ra;
    const int i = grp_b ? r - ra : r;
    const int rows = grp_b ? n_sink_b : n_sink;
    const int K = grp_b ? K_b : K_a;
    const double *base = partial + (grp_b ? 3 * static_cast<size_t>(K_a > 1 ? K_a : 0) * n_sink : 0);
    double ax, ay, az;
    sum_kick_partials(base, K, rows, i, ax, ay, az);
    T *si_vel = grp_b ? gb_vel : ga_vel;
    T *si_acc = grp_b ? gb_acc : ga_acc;
    const T a0 = static_cast<T>(ax), a1 = static_cast<T>(ay), a2 = static_cast<T>(az);
    if (si_acc) {
        si_acc[3 * i + 0] = a0;
        si_acc[3 * i + 1] = a1;
        si_acc[3 * i + 2] = a2;
    }
    si_vel[3 * i + 0] = si_vel[3 * i + 0] + a0 * h;
    si_vel[3 * i + 1] = si_vel[3 * i + 1] + a1 * h;
    si_vel[3 * i + 2] = si_vel[3 * i + 2] + a2 * h;
}

// Chunk counts of the two sink groups (kick_kernel).
struct SplitK {
    int a, b;
};

template <typename T, int S, int W, int P>
int launch_swp(const KickArgs<T> &a, SplitK K, double *partial, const float *mf,
               hipStream_t st) {
    const int tile = 64 * S;
    const int tiles_a = (a.n_sink + tile - 1) / tile, tiles_b = (a.n_sink_b + tile - 1) / tile;
    dim3 grid(tiles_a * K.a + tiles_b * K.b), block(64 * W);
    const double eps2 = a.eps2;
    const float eps2f = static_cast<float>(a.eps2);
    const bool split = K.a > 1 || K.b > 1;
    double *part = split ? partial : nullptr;
    if (a.eps2 > 0.0)
        hipLaunchKernelGGL((kick_kernel<T, S, W, P, true>), grid, block, 0, st, a.si_pos,
                           a.si_vel, a.si_acc, a.n_sink, a.sb_pos, a.sb_vel, a.sb_acc, a.n_sink_b,
                           a.s0_pos, a.s0_m, a.n0, a.s1_pos, a.s1_m, a.n1, a.h, eps2f, eps2, K.a,
                           K.b, part, mf);
    else
        hipLaunchKernelGGL((kick_kernel<T, S, W, P, false>), grid, block, 0, st, a.si_pos,
                           a.si_vel, a.si_acc, a.n_sink, a.sb_pos, a.sb_vel, a.sb_acc, a.n_sink_b,
                           a.s0_pos, a.s0_m, a.n0, a.s1_pos, a.s1_m, a.n1, a.h, eps2f, eps2, K.a,
                           K.b, part, mf);
    HHS_LAUNCH_CHECK("kick_kernel");
    return HHS_OK;
}

template <typename T, int S, int W>
int launch_sw(const KickArgs<T> &a, int P, SplitK K, double *partial, const float *mf,
              hipStream_t st) {
    if constexpr (std::is_same<T, double>::value) {
        if (P == 1) return launch_swp<T, S, W, 1>(a, K, partial, mf, st);
        if (P == 2) return launch_swp<T, S, W, 2>(a, K, partial, mf, st);
        if (P == 3) return launch_swp<T, S, W, 3>(a, K, partial, mf, st);
    }
    return launch_swp<T, S, W, 0>(a, K, partial, mf, st);
}

template <typename T, int S>
int launch_s(const KickArgs<T> &a, int W, int P, SplitK K, double *partial, const float *mf,
             hipStream_t st) {
    switch (W) {
    case 1: return launch_sw<T, S, 1>(a, P, K, partial, mf, st);
    case 2: return launch_sw<T, S, 2>(a, P, K, partial, mf, st);
    case 4: return launch_sw<T, S, 4>(a, P, K, partial, mf, st);
    case 8: return launch_sw<T, S, 8>(a, P, K, partial, mf, st);
    }
    set_error("kick: unsupported waves per group %d", W);
    return HHS_ERR_INVALID;
}

std::atomic<int> g_force_s{0}, g_force_w{0};
std::atomic<int> g_precision{HHS_KICK_MIXED32};

bool valid_s(int s) { return s == 1 || s == 2 || s == 4; }
bool valid_w(int w) { return w == 1 || w == 2 || w == 4 || w == 8; }

// Autotuned configurations per sink-count size class (floor(log2 n)),
// encoded S << 8 | W; 0 = use the heuristic.
std::atomic<int> g_tuned[32];
int size_class(int n) { return n <= 1 ? 0 : 31 - __builtin_clz(static_cast<unsigned>(n)); }

}  // namespace

// Heuristic from the MI355X sweep (FP64 storage, MIXED32): at N = 131072,
// S = 1 / W = 8 gives 2596 Ginteractions/s, S = 2 / W = 8 2566, S = 1 /
// W = 4 2377, S = 4 / W = 4 2150; at N = 16384, S = 1 / W = 8 1413 and
// S = 2 / W = 8 895.  Latency is hidden by waves, not by deeper register
// blocking: one sink per lane and up to 8 waves per workgroup while each
// wave keeps >= 32 sources (64 before: on whole HOLD runs with the split
// below from 256 sources, N = 1024 to t = 1 0.344 -> 0.316 s, N = 4096 to
// t = 0.25 0.232 -> 0.207 s; 16 did no better, 128 worse).
KickConfig kick_pick_config(int n_sink, int n_src) {
    KickConfig c;
    const int fs = g_force_s.load(), fw = g_force_w.load();
    c.sinks_per_lane = 1;
    int w = 1;
    while (w < 8 && n_src / (2 * w) >= 32) w *= 2;
    c.waves = w;
    const int tuned = g_tuned[size_class(n_sink)].load();
    if (tuned) {
        c.sinks_per_lane = tuned >> 8;
        c.waves = tuned & 0xff;
    }
    if (valid_s(fs)) c.sinks_per_lane = fs;
    if (valid_w(fw)) c.waves = fw;
    return c;
}

// Split-K for small sink counts: more workgroups over chunks of the same
// source list while the grid stays under 8192 waves (8 per SIMD on 256
// CUs; see below for large groups), each wave keeps enough sources and the
// FP64 partials (24 B per sink and chunk) stay within kMaxPartialBytes.  At N = 4096 a plain launch is
// 512 waves, each running 512 sources back to back with nothing to hide
// its latency.  MI355X, square kicks: N = 4096 48 -> 21 us, 8192 96 ->
// 39 us, 16384 194 -> 122 us, 32768 570 -> 460 us.  From 256 sources up
// (1024, then 512 before): the HOLD calls of a few hundred to a thousand
// particles gain more than the reduce launch costs (1024 -> 512: N = 1024
// to t = 1 0.375 -> 0.361 s, N = 16384 in one root step 0.415 -> 0.402 s;
// 512 -> 256 together with the wave rule above: see there).
// Measured on whole HOLD runs, a sink group of at least 2^28 interactions
// (the skinny kicks of the upper levels of a large system: a few thousand
// slow sinks over every source) may fill up to 32768 waves, four rounds of
// the chip, and every wave keeps >= 16 sources (32 before): N = 131072 root
// step 1.436 -> 1.378-1.385 s, N = 16384 HOLD 0.361 -> 0.353-0.354 s,
// N = 4096 to t = 0.25 0.237 -> 0.232-0.235 s, N = 1024 to t = 1 0.357 ->
// 0.350-0.354 s.  A 16384-wave cap for every group, or >= 8 sources per
// wave, did worse.
constexpr size_t kMaxPartialBytes = size_t{4} << 20;
constexpr int kSplitMinSources = 256;
constexpr long long kSplitWaves = 8192, kSplitWavesBig = 32768;
constexpr double kSplitBigWork = 268435456.0;  // 2^28 interactions
constexpr int kSplitMinChunk = 16;             // sources per wave

int pick_split(int n_sink, int n_src, int S, int W) {
    const long long tiles = (n_sink + 64LL * S - 1) / (64LL * S);
    int K = 1;
    if (n_sink <= 0 || n_src < kSplitMinSources)  // the reduce launch would cost more
        return K;
    const long long cap = static_cast<double>(n_sink) * n_src >= kSplitBigWork ? kSplitWavesBig
                                                                                : kSplitWaves;
    while (K < 64 && tiles * W * K * 2 <= cap && n_src / (W * K * 2) >= kSplitMinChunk &&
           24ULL * n_sink * K * 2 <= kMaxPartialBytes)
        K *= 2;
    return K;
}

template <typename T> int launch_kick(const KickArgs<T> &a, hipStream_t st, KickDefer *defer) {
    if (defer) *defer = KickDefer{};
    HHS_REQUIRE(a.n_sink >= 0 && a.n_sink_b >= 0 && a.n0 >= 0 && a.n1 >= 0,
                "kick: negative size");
    const int n_rows = a.n_sink + a.n_sink_b;  // sinks of both groups
    if (n_rows == 0) return HHS_OK;
    HHS_REQUIRE(a.n_sink == 0 || (a.si_pos && a.si_vel), "kick: null sink pos/vel");
    HHS_REQUIRE(a.n_sink_b == 0 || (a.sb_pos && a.sb_vel), "kick: null second-group pos/vel");
    HHS_REQUIRE(a.n0 == 0 || (a.s0_pos && a.s0_m), "kick: null source segment 0");
    HHS_REQUIRE(a.n1 == 0 || (a.s1_pos && a.s1_m), "kick: null source segment 1");
    // the first group's sources (the most any sink sees) set the waves per
    // group and the split; the second group's waves take shorter chunks
    const int n_src_cfg = a.n_sink > 0 ? a.n0 + a.n1 : a.n0;
    KickConfig c = kick_pick_config(n_rows, n_src_cfg);
    const int P = a.precision >= 0 ? a.precision : g_precision.load();
    HHS_REQUIRE(P >= HHS_KICK_MIXED32 && P <= HHS_KICK_F64_EXACT, "kick: precision mode %d", P);
    HHS_REQUIRE(a.n_sink_b == 0 || P != HHS_KICK_F64_EXACT,
                "kick: the exact FP64 mode takes one sink group per launch");
    if (P == HHS_KICK_F64_EXACT && std::is_same<T, double>::value) {
        // One lane per sink, one wave per group: each sink sums its sources
        // in index order.  Segment 1 follows segment 0 in the same sum.
        c.sinks_per_lane = 1;
        c.waves = 1;
    }
    // Each group's split from its own shape; the exact mode keeps one
    // sequential sum per sink.
    SplitK K{1, 1};
    if (P != HHS_KICK_F64_EXACT) {
        K.a = pick_split(a.n_sink, a.n0 + a.n1, c.sinks_per_lane, c.waves);
        K.b = pick_split(a.n_sink_b, a.n0, c.sinks_per_lane, c.waves);
    }
    static thread_local ThreadScratch scratch, mass_scratch;
    double *partial = nullptr;
    if (K.a > 1 || K.b > 1) {
        const size_t rows_k = (K.a > 1 ? static_cast<size_t>(a.n_sink) * K.a : 0) +
                              (K.b > 1 ? static_cast<size_t>(a.n_sink_b) * K.b : 0);
        void *p = nullptr;
        int rc = scratch.acquire(24 * rows_k, st, &p);
        if (rc) return rc;
        partial = static_cast<double *>(p);
    }
    // MIXED32 on FP64 storage, large launches: the source masses are
    // converted to FP32 once up front (the same values the loop would convert
    // per source and wave, one VALU op in 17); below ~2^29 interactions the
    // extra launch costs more than it saves.
    float *mf = nullptr;
    const int n_src = a.n0 + a.n1;
    const double work = static_cast<double>(a.n_sink) * n_src +
                        static_cast<double>(a.n_sink_b) * a.n0;
    const bool conv = std::is_same<T, double>::value && P == HHS_KICK_MIXED32 &&
                      work >= kConvMinWork;
    if (conv) {
        void *p = nullptr;
        int rc = mass_scratch.acquire(4 * static_cast<size_t>(n_src), st, &p);
        if (rc) return rc;
        mf = static_cast<float *>(p);
        hipLaunchKernelGGL((masses_to_f32_kernel<T>), dim3((n_src + 255) / 256), dim3(256), 0, st,
                           a.s0_m, a.n0, a.s1_m, a.n1, mf);
        HHS_LAUNCH_CHECK("masses_to_f32_kernel");
    }
    int rc = HHS_ERR_INVALID;
    switch (c.sinks_per_lane) {
    case 1: rc = launch_s<T, 1>(a, c.waves, P, K, partial, mf, st); break;
    case 2: rc = launch_s<T, 2>(a, c.waves, P, K, partial, mf, st); break;
    case 4: rc = launch_s<T, 4>(a, c.waves, P, K, partial, mf, st); break;
    default: set_error("kick: unsupported sinks per lane %d", c.sinks_per_lane);
    }
    if (rc == HHS_OK && (K.a > 1 || K.b > 1)) {
        if (defer) {
            // the readers add the partials (the layout kick_reduce_kernel reads)
            if (K.a > 1) defer->a = KickPending{partial, K.a, a.n_sink, static_cast<double>(a.h)};
            if (K.b > 1)
                defer->b = KickPending{partial + (K.a > 1 ? 3 * static_cast<size_t>(K.a) * a.n_sink
                                                          : 0),
                                       K.b, a.n_sink_b, static_cast<double>(a.h)};
        } else {
            const int rows = (K.a > 1 ? a.n_sink : 0) + (K.b > 1 ? a.n_sink_b : 0);
            hipLaunchKernelGGL((kick_reduce_kernel<T>), dim3((rows + 255) / 256), dim3(256), 0,
                               st, a.si_vel, a.si_acc, a.n_sink, a.sb_vel, a.sb_acc, a.n_sink_b,
                               partial, K.a, K.b, a.h);
            HHS_LAUNCH_CHECK("kick_reduce_kernel");
        }
        rc = scratch.release(st);
    }
    if (rc == HHS_OK && conv) rc = mass_scratch.release(st);
    return rc;
}

template int launch_kick<double>(const KickArgs<double> &, hipStream_t, KickDefer *);
template int launch_kick<float>(const KickArgs<float> &, hipStream_t, KickDefer *);

}  // namespace hhs

using namespace hhs;

template <typename T, typename P>
static int kick_self(const P *sys, int n, T dt, T fac, T eps2, void *stream) {
    HHS_REQUIRE(sys, "kick_self: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "kick_self: n=%d outside [0, %d]", n, sys->n);
    KickArgs<T> a{sys->pos, sys->vel, sys->acc, n, sys->pos, sys->m, n, nullptr, nullptr, 0,
                  fac * dt, static_cast<double>(eps2)};
    return launch_kick<T>(a, as_stream(stream));
}

template <typename T, typename P>
static int kick_cross(const P *so, const P *si, int n_sink, int n_src, T dt, T fac, T eps2,
                      void *stream) {
    HHS_REQUIRE(so && si, "kick_cross: null system");
    HHS_REQUIRE(n_sink >= 0 && n_sink <= si->n, "kick_cross: n_sink=%d outside [0, %d]", n_sink,
                si->n);
    HHS_REQUIRE(n_src >= 0 && n_src <= so->n, "kick_cross: n_src=%d outside [0, %d]", n_src,
                so->n);
    KickArgs<T> a{si->pos, si->vel, si->acc, n_sink, so->pos, so->m, n_src, nullptr, nullptr, 0,
                  fac * dt, static_cast<double>(eps2)};
    return launch_kick<T>(a, as_stream(stream));
}

extern "C" {

int hhs_kick_self_f64(const hhs_particles_f64 *sys, int n, double dt, double fac, double eps2,
                      void *stream) {
    return kick_self<double>(sys, n, dt, fac, eps2, stream);
}

int hhs_kick_self_f32(const hhs_particles_f32 *sys, int n, float dt, float fac, float eps2,
                      void *stream) {
    return kick_self<float>(sys, n, dt, fac, eps2, stream);
}

int hhs_kick_cross_f64(const hhs_particles_f64 *so, const hhs_particles_f64 *si, int n_sink,
                       int n_src, double dt, double fac, double eps2, void *stream) {
    return kick_cross<double>(so, si, n_sink, n_src, dt, fac, eps2, stream);
}

int hhs_kick_cross_f32(const hhs_particles_f32 *so, const hhs_particles_f32 *si, int n_sink,
                       int n_src, float dt, float fac, float eps2, void *stream) {
    return kick_cross<float>(so, si, n_sink, n_src, dt, fac, eps2, stream);
}

int hhs_kick_set_config(int sinks_per_lane, int waves_per_group) {
    if (sinks_per_lane == 0 && waves_per_group == 0) {
        g_force_s.store(0);
        g_force_w.store(0);
        return HHS_OK;
    }
    HHS_REQUIRE(valid_s(sinks_per_lane), "kick config: sinks_per_lane must be 1, 2 or 4");
    HHS_REQUIRE(valid_w(waves_per_group), "kick config: waves_per_group must be 1, 2, 4 or 8");
    g_force_s.store(sinks_per_lane);
    g_force_w.store(waves_per_group);
    return HHS_OK;
}

int hhs_kick_set_precision(int mode) {
    HHS_REQUIRE(mode >= HHS_KICK_MIXED32 && mode <= HHS_KICK_F64_EXACT,
                "kick precision: mode must be HHS_KICK_MIXED32, HHS_KICK_F64_RSQ32, HHS_KICK_F64 "
                "or HHS_KICK_F64_EXACT");
    g_precision.store(mode);
    return HHS_OK;
}

int hhs_kick_get_precision(void) { return g_precision.load(); }

int hhs_autotune_kick(int n, int reps, const char *cache_path, int *best_s, int *best_w,
                      double *best_ms) {
    HHS_REQUIRE(n >= 64, "autotune_kick: n must be >= 64");
    if (reps <= 0) reps = 3;
    std::vector<double> p(3 * static_cast<size_t>(n)), v(3 * static_cast<size_t>(n)),
        m(static_cast<size_t>(n));
    int rc = hhs_ic_plummer(n, 12345, 0, 0, 0.0, p.data(), v.data(), m.data());
    if (rc) return rc;
    double *dev = nullptr;
    const size_t nn = static_cast<size_t>(n);
    HHS_HIP(hipMalloc(&dev, nn * 80));  // pos, vel, acc (3n each) and m
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemcpy(dev, p.data(), nn * 24, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dev + 3 * nn, v.data(), nn * 24, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dev + 9 * nn, m.data(), nn * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const int saved_s = g_force_s.load(), saved_w = g_force_w.load();
    double best = 1e300;
    int bs = 0, bw = 0;
    for (int s : {1, 2, 4}) {
        for (int w : {1, 2, 4, 8}) {
            if (e != hipSuccess || rc) break;
            g_force_s.store(s);
            g_force_w.store(w);
            KickArgs<double> a{dev,      dev + 3 * nn, dev + 6 * nn, n,   dev, dev + 9 * nn,
                               n,        nullptr,      nullptr,      0,   0.0, 1e-8};
            rc = launch_kick<double>(a, nullptr);  // warm-up
            if (!rc) e = hipEventRecord(e0, nullptr);
            for (int r = 0; r < reps && !rc && e == hipSuccess; ++r)
                rc = launch_kick<double>(a, nullptr);
            if (!rc && e == hipSuccess) e = hipEventRecord(e1, nullptr);
            if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (!rc && e == hipSuccess && ms / reps < best) {
                best = ms / reps;
                bs = s;
                bw = w;
            }
        }
    }
    g_force_s.store(saved_s);
    g_force_w.store(saved_w);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_status(e, "autotune_kick", __FILE__, __LINE__);
    if (rc) return rc;
    const int cls = size_class(n);
    g_tuned[cls].store(bs << 8 | bw);
    if (cache_path) {
        FILE *f = std::fopen(cache_path, "a");
        if (!f) {
            set_error("autotune_kick: cannot append to %s", cache_path);
            return HHS_ERR_INVALID;
        }
        std::fprintf(f, "kick %d %d %d %.6f\n", cls, bs, bw, best);
        std::fclose(f);
    }
    if (best_s) *best_s = bs;
    if (best_w) *best_w = bw;
    if (best_ms) *best_ms = best;
    return HHS_OK;
}

int hhs_autotune_load(const char *cache_path) {
    HHS_REQUIRE(cache_path, "autotune_load: null path");
    FILE *f = std::fopen(cache_path, "r");
    if (!f) {
        set_error("autotune_load: cannot open %s", cache_path);
        return HHS_ERR_INVALID;
    }
    // "kick <class> <S> <W> <ms>" and "ts <mode> <class> <W> <K> <ms>";
    // later lines win
    char line[256], kind[16];
    int a0, a1, a2, a3;
    double ms;
    int loaded = 0;
    while (std::fgets(line, sizeof(line), f)) {
        if (std::sscanf(line, "%15s", kind) != 1) continue;
        if (std::strcmp(kind, "kick") == 0 &&
            std::sscanf(line, "%*s %d %d %d %lf", &a0, &a1, &a2, &ms) == 4 && a0 >= 0 &&
            a0 < 32 && valid_s(a1) && valid_w(a2)) {
            g_tuned[a0].store(a1 << 8 | a2);
            ++loaded;
        } else if (std::strcmp(kind, "ts") == 0 &&
                   std::sscanf(line, "%*s %d %d %d %d %lf", &a0, &a1, &a2, &a3, &ms) == 5 &&
                   ts_tuned_valid(a0, a1, a2, a3)) {
            ts_tuned_install(a0, a1, a2, a3);
            ++loaded;
        }
    }
    std::fclose(f);
    return loaded > 0 ? HHS_OK : (set_error("autotune_load: no entries in %s", cache_path),
                                  HHS_ERR_INVALID);
}

int hhs_autotune_clear(void) {
    for (auto &t : g_tuned) t.store(0);
    ts_tuned_clear();
    return HHS_OK;
}

int hhs_kick_get_config(int n_sink, int n_src, int *sinks_per_lane, int *waves_per_group) {
    HHS_REQUIRE(sinks_per_lane && waves_per_group, "kick_get_config: null output");
    const KickConfig c = kick_pick_config(n_sink, n_src);
    *sinks_per_lane = c.sinks_per_lane;
    *waves_per_group = c.waves;
    return HHS_OK;
}

}  // extern "C"
