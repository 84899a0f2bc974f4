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

#include "hhs_internal.h"
#include "pair_math.h"

namespace hhs {
namespace {

constexpr int kFold = 64;  // sources per plain-FP32 block before a two-sum fold

template <typename T, int S, bool SOFT>
__device__ __forceinline__ void accumulate_range(const T *__restrict__ sp,
                                                 const T *__restrict__ sm, int jb, int je,
                                                 const T (&xi)[S], const T (&yi)[S],
                                                 const T (&zi)[S], float eps2, float (&hx)[S],
                                                 float (&lx)[S], float (&hy)[S], float (&ly)[S],
                                                 float (&hz)[S], float (&lz)[S]) {
    for (int j0 = jb; j0 < je; j0 += kFold) {
        const int j1 = min(j0 + kFold, je);
        float px[S], py[S], pz[S];
#pragma unroll
        for (int s = 0; s < S; ++s) px[s] = py[s] = pz[s] = 0.0f;
#pragma unroll 4
        for (int j = j0; j < j1; ++j) {
            const T xj = sp[3 * j + 0];
            const T yj = sp[3 * j + 1];
            const T zj = sp[3 * j + 2];
            const float mj = static_cast<float>(sm[j]);
#pragma unroll
            for (int s = 0; s < S; ++s)
                pair_accel<SOFT>(xj, yj, zj, mj, xi[s], yi[s], zi[s], eps2, px[s], py[s], pz[s]);
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            two_sum_add(hx[s], lx[s], px[s]);
            two_sum_add(hy[s], ly[s], py[s]);
            two_sum_add(hz[s], lz[s], pz[s]);
        }
    }
}

template <typename T, int S, int W, bool SOFT>
__global__ __launch_bounds__(64 * W) void kick_kernel(
    const T *__restrict__ si_pos, T *__restrict__ si_vel, T *__restrict__ si_acc, int n_sink,
    const T *__restrict__ s0_pos, const T *__restrict__ s0_m, int n0,
    const T *__restrict__ s1_pos, const T *__restrict__ s1_m, int n1, T h, float eps2) {
    __shared__ float red[(W > 1 ? W - 1 : 1) * 6 * S * 64];

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tile0 = blockIdx.x * (64 * S);

    T xi[S], yi[S], zi[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        int i = tile0 + s * 64 + lane;
        i = i < n_sink ? i : n_sink - 1;
        xi[s] = si_pos[3 * i + 0];
        yi[s] = si_pos[3 * i + 1];
        zi[s] = si_pos[3 * i + 2];
    }
    float hx[S], lx[S], hy[S], ly[S], hz[S], lz[S];
#pragma unroll
    for (int s = 0; s < S; ++s) hx[s] = lx[s] = hy[s] = ly[s] = hz[s] = lz[s] = 0.0f;

    // This wave's contiguous share of the concatenated source list.
    const int total = n0 + n1;
    const int chunk = (total + W - 1) / W;
    const int jb = min(wave * chunk, total);
    const int je = min(jb + chunk, total);
    if (jb < n0)
        accumulate_range<T, S, SOFT>(s0_pos, s0_m, jb, min(je, n0), xi, yi, zi, eps2, hx, lx, hy,
                                     ly, hz, lz);
    if (je > n0)
        accumulate_range<T, S, SOFT>(s1_pos, s1_m, max(jb, n0) - n0, je - n0, xi, yi, zi, eps2,
                                     hx, lx, hy, ly, hz, lz);

    double ax[S], ay[S], az[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        ax[s] = static_cast<double>(hx[s]) + static_cast<double>(lx[s]);
        ay[s] = static_cast<double>(hy[s]) + static_cast<double>(ly[s]);
        az[s] = static_cast<double>(hz[s]) + static_cast<double>(lz[s]);
    }
    if constexpr (W > 1) {
        if (wave > 0) {
            float *r = red + (wave - 1) * 6 * S * 64;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                r[(0 * S + s) * 64 + lane] = hx[s];
                r[(1 * S + s) * 64 + lane] = lx[s];
                r[(2 * S + s) * 64 + lane] = hy[s];
                r[(3 * S + s) * 64 + lane] = ly[s];
                r[(4 * S + s) * 64 + lane] = hz[s];
                r[(5 * S + s) * 64 + lane] = lz[s];
            }
        }
        __syncthreads();
        if (wave != 0) return;
        for (int w = 1; w < W; ++w) {  // fixed order: deterministic
            const float *r = red + (w - 1) * 6 * S * 64;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                ax[s] += static_cast<double>(r[(0 * S + s) * 64 + lane]) +
                         static_cast<double>(r[(1 * S + s) * 64 + lane]);
                ay[s] += static_cast<double>(r[(2 * S + s) * 64 + lane]) +
                         static_cast<double>(r[(3 * S + s) * 64 + lane]);
                az[s] += static_cast<double>(r[(4 * S + s) * 64 + lane]) +
                         static_cast<double>(r[(5 * S + s) * 64 + lane]);
            }
        }
    }
    // Epilogue (src/kick.cuh:98-104): acc <- a, vel += a * (fac * dt).
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

template <typename T, int S, int W>
int launch_sw(const KickArgs<T> &a, hipStream_t st) {
    const int tile = 64 * S;
    dim3 grid((a.n_sink + tile - 1) / tile), block(64 * W);
    if (a.eps2 > 0.0f)
        hipLaunchKernelGGL((kick_kernel<T, S, W, true>), grid, block, 0, st, a.si_pos, a.si_vel,
                           a.si_acc, a.n_sink, a.s0_pos, a.s0_m, a.n0, a.s1_pos, a.s1_m, a.n1,
                           a.h, a.eps2);
    else
        hipLaunchKernelGGL((kick_kernel<T, S, W, false>), grid, block, 0, st, a.si_pos, a.si_vel,
                           a.si_acc, a.n_sink, a.s0_pos, a.s0_m, a.n0, a.s1_pos, a.s1_m, a.n1,
                           a.h, a.eps2);
    HHS_LAUNCH_CHECK("kick_kernel");
    return HHS_OK;
}

template <typename T, int S>
int launch_s(const KickArgs<T> &a, int W, hipStream_t st) {
    switch (W) {
    case 1: return launch_sw<T, S, 1>(a, st);
    case 2: return launch_sw<T, S, 2>(a, st);
    case 4: return launch_sw<T, S, 4>(a, st);
    case 8: return launch_sw<T, S, 8>(a, st);
    }
    set_error("kick: unsupported waves per group %d", W);
    return HHS_ERR_INVALID;
}

std::atomic<int> g_force_s{0}, g_force_w{0};

bool valid_s(int s) { return s == 1 || s == 2 || s == 4; }
bool valid_w(int w) { return w == 1 || w == 2 || w == 4 || w == 8; }

}  // namespace

// Heuristic: enough sinks per lane to amortise each scalar source fetch,
// then split sources across waves until the chip holds ~2 waves per SIMD
// (256 CUs x 4 SIMDs), keeping >= 64 sources per wave.
KickConfig kick_pick_config(int n_sink, int n_src) {
    KickConfig c;
    const int fs = g_force_s.load(), fw = g_force_w.load();
    c.sinks_per_lane = n_sink >= 65536 ? 4 : (n_sink >= 8192 ? 2 : 1);
    const int tiles = (n_sink + 64 * c.sinks_per_lane - 1) / (64 * c.sinks_per_lane);
    int w = 1;
    while (w < 8 && tiles * w < 2048 && n_src / (2 * w) >= 64) w *= 2;
    c.waves = w;
    if (valid_s(fs)) c.sinks_per_lane = fs;
    if (valid_w(fw)) c.waves = fw;
    return c;
}

template <typename T> int launch_kick(const KickArgs<T> &a, hipStream_t st) {
    HHS_REQUIRE(a.n_sink >= 0 && a.n0 >= 0 && a.n1 >= 0, "kick: negative size");
    if (a.n_sink == 0) return HHS_OK;
    HHS_REQUIRE(a.si_pos && a.si_vel, "kick: null sink pos/vel");
    HHS_REQUIRE(a.n0 == 0 || (a.s0_pos && a.s0_m), "kick: null source segment 0");
    HHS_REQUIRE(a.n1 == 0 || (a.s1_pos && a.s1_m), "kick: null source segment 1");
    const KickConfig c = kick_pick_config(a.n_sink, a.n0 + a.n1);
    switch (c.sinks_per_lane) {
    case 1: return launch_s<T, 1>(a, c.waves, st);
    case 2: return launch_s<T, 2>(a, c.waves, st);
    case 4: return launch_s<T, 4>(a, c.waves, st);
    }
    set_error("kick: unsupported sinks per lane %d", c.sinks_per_lane);
    return HHS_ERR_INVALID;
}

template int launch_kick<double>(const KickArgs<double> &, hipStream_t);
template int launch_kick<float>(const KickArgs<float> &, hipStream_t);

}  // namespace hhs

using namespace hhs;

template <typename T, typename P>
static int kick_self(const P *sys, int n, T dt, T fac, T eps2, void *stream) {
    HHS_REQUIRE(sys, "kick_self: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "kick_self: n=%d outside [0, %d]", n, sys->n);
    KickArgs<T> a{sys->pos, sys->vel, sys->acc, n, sys->pos, sys->m, n, nullptr, nullptr, 0,
                  fac * dt, static_cast<float>(eps2)};
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
                  fac * dt, static_cast<float>(eps2)};
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

int hhs_kick_get_config(int n_sink, int n_src, int *sinks_per_lane, int *waves_per_group) {
    HHS_REQUIRE(sinks_per_lane && waves_per_group, "kick_get_config: null output");
    const KickConfig c = kick_pick_config(n_sink, n_src);
    *sinks_per_lane = c.sinks_per_lane;
    *waves_per_group = c.waves;
    return HHS_OK;
}

}  // extern "C"
