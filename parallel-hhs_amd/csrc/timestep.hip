// Pairwise timestep minimum for gfx950.
//
// Reference: _timesteps / _get_pp_timestep (src/timestep.cuh:10-118).
// tau_i = min over j != i of the free-fall and fly-by pair timesteps, start
// value 1e30 (src/timestep.cuh:4,:75), with r^2 -> r^2 + eps^2.
//
// The formula is evaluated in the reference's operation order with IEEE
// sqrt and divide and no FMA contraction, so the result is bit-exact
// against an FP64 CPU evaluation of the same expression.  NaN candidates
// never win (`tau < minval` is false for NaN), as in the reference.
//
// One sink per lane; the source index is wave-uniform, so source position,
// velocity and mass arrive through scalar loads (a wave-wide broadcast).
// The reference's early `return` before a barrier (src/timestep.cuh:70 vs
// :89,:95) has no counterpart: there are no barriers here, and lanes past
// n compute on a clamped sink and do not store.
#include <algorithm>
#include <atomic>
#include <cstdio>
#include <type_traits>
#include <vector>

#include "hhs_internal.h"
#include "pair_math.h"

namespace hhs {
namespace {

// Running minimum for the three evaluation modes (HHS_TIMESTEP_EXACT,
// _FAST, _FP32).  The exact form keeps minval; the others keep a fraction
// (num, den) compared by cross-multiplication and divide once at the end.
template <typename T, int MODE> struct TsAcc {
    using V = typename std::conditional<MODE == HHS_TIMESTEP_FP32, float, T>::type;
    V num = static_cast<V>(TsConst<T>::huge), den = V(1);
    __device__ void pair(T xi, T yi, T zi, T vxi, T vyi, T vzi, T mi, const T *pos, const T *vel,
                         const T *m, int j, T eta, T c_ff, T eps2, bool other) {
        if constexpr (MODE == HHS_TIMESTEP_FAST)
            pair_timestep_fast(xi, yi, zi, vxi, vyi, vzi, mi, pos[3 * j], pos[3 * j + 1],
                               pos[3 * j + 2], vel[3 * j], vel[3 * j + 1], vel[3 * j + 2], m[j],
                               eta, c_ff, eps2, other, num, den);
        else if constexpr (MODE == HHS_TIMESTEP_FP32)
            pair_timestep_f32(xi, yi, zi, vxi, vyi, vzi, static_cast<float>(mi), pos[3 * j],
                              pos[3 * j + 1], pos[3 * j + 2], vel[3 * j], vel[3 * j + 1],
                              vel[3 * j + 2], static_cast<float>(m[j]), static_cast<float>(eta),
                              static_cast<float>(c_ff), static_cast<float>(eps2), other, num, den);
        else
            pair_timestep<T>(xi, yi, zi, vxi, vyi, vzi, mi, pos[3 * j], pos[3 * j + 1],
                             pos[3 * j + 2], vel[3 * j], vel[3 * j + 1], vel[3 * j + 2], m[j], eta,
                             c_ff, eps2, other, num);
    }
    __device__ void merge(V onum, V oden) {
        if constexpr (MODE == HHS_TIMESTEP_FAST)
            fold_fraction(onum, oden, true, num, den);
        else if constexpr (MODE == HHS_TIMESTEP_FP32)
            fold_fraction32(onum, oden, true, num, den);
        else if (onum < num)
            num = onum;
    }
    __device__ T value() const {
        if constexpr (MODE == HHS_TIMESTEP_EXACT)
            return num;
        else
            return static_cast<T>(num) / static_cast<T>(den);
    }
};

// 64 sinks per workgroup, one per lane of every wave, against source chunk
// blockIdx.y of K (split-K for small n); the W waves split the chunk and
// their minima are merged through LDS in wave order.  With K > 1 each chunk
// writes its (num, den) to partial[k][i][2] and timesteps_merge_kernel folds
// the chunks in order.  The minimum does not depend on the order, so the
// exact form stays bit-exact.
template <typename T, int W, int MODE>
__global__ __launch_bounds__(64 * W) void timesteps_kernel(
    const T *__restrict__ pos, const T *__restrict__ vel, const T *__restrict__ m,
    T *__restrict__ ts, int n, T eta, T eps2, int K,
    typename TsAcc<T, MODE>::V *__restrict__ partial) {
    using V = typename TsAcc<T, MODE>::V;
    __shared__ V red[2][(W > 1 ? W - 1 : 1) * 64];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = blockIdx.x * 64 + lane;
    const int ic = i < n ? i : n - 1;
    const T xi = pos[3 * ic], yi = pos[3 * ic + 1], zi = pos[3 * ic + 2];
    const T vxi = vel[3 * ic], vyi = vel[3 * ic + 1], vzi = vel[3 * ic + 2];
    const T mi = m[ic];
    const T c_ff = eta / TsConst<T>::sqrt2;
    const int per = (n + K - 1) / K;
    const int kb = min(static_cast<int>(blockIdx.y) * per, n), ke = min(kb + per, n);
    const int chunk = (ke - kb + W - 1) / W;
    const int jb = min(kb + wave * chunk, ke), je = min(jb + chunk, ke);
    TsAcc<T, MODE> acc;
    for (int j = jb; j < je; ++j)
        acc.pair(xi, yi, zi, vxi, vyi, vzi, mi, pos, vel, m, j, eta, c_ff, eps2, j != i);
    if constexpr (W > 1) {
        if (wave > 0) {
            red[0][(wave - 1) * 64 + lane] = acc.num;
            red[1][(wave - 1) * 64 + lane] = acc.den;
        }
        __syncthreads();
        if (wave != 0) return;
        for (int w = 1; w < W; ++w) acc.merge(red[0][(w - 1) * 64 + lane], red[1][(w - 1) * 64 + lane]);
    }
    if (i >= n) return;
    if (partial) {
        V *q = partial + 2 * (static_cast<size_t>(blockIdx.y) * n + i);
        q[0] = acc.num;
        q[1] = acc.den;
    } else {
        ts[i] = acc.value();
    }
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void timesteps_merge_kernel(
    T *__restrict__ ts, int n, int K, const typename TsAcc<T, MODE>::V *__restrict__ partial) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    TsAcc<T, MODE> acc;
    // eight chunks' loads issued before their folds (in chunk order, as a
    // plain loop would fold them): one memory round trip per eight chunks
    constexpr int B = 8;
    using V = typename TsAcc<T, MODE>::V;
    for (int k = 0; k < K; k += B) {
        V num[B], den[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            if (k + j < K) {
                const V *q = partial + 2 * (static_cast<size_t>(k + j) * n + i);
                num[j] = q[0];
                den[j] = q[1];
            }
        }
#pragma unroll
        for (int j = 0; j < B; ++j)
            if (k + j < K) acc.merge(num[j], den[j]);
    }
    ts[i] = acc.value();
}

// Waves per workgroup: up to 8 waves per SIMD over the chip (2048 SIMDs x 4
// at 256 CUs) while every wave keeps >= 64 sources.
int pick_waves(int n) {
    const int tiles = (n + 63) / 64;
    int w = 1;
    while (w < 8 && tiles * w < 8192 && n / (2 * w) >= 64) w *= 2;
    return w;
}

// Split-K as for the kick: more workgroups over chunks of the sources
// while the grid stays within kTsMaxWaves waves, each wave keeps >=
// kTsMinChunk sources and the partials (two values per sink and chunk) stay
// within 4 MB.  The limits follow the autotuner's choices on MI355X (FP32:
// n = 8192 K 8 -> 16, 16384 4 -> 8, 32768 2 -> 4, 15-20 % faster).  The
// chunk floor of 8 sources (16 before) comes from whole HOLD runs, where
// the calls of a few hundred to a few thousand particles are latency-bound:
// N = 1024 to t = 1 0.351 -> 0.345 s, N = 4096 to t = 0.25 0.234 -> 0.229 s,
// N = 16384 in one root step 0.354 -> 0.349 s (4 sources: no better).
constexpr int kTsMinN = 256, kTsMinChunk = 8;
constexpr long long kTsMaxWaves = 16384;
template <typename V> int pick_ts_split(int n, int W) {
    const long long tiles = (n + 63) / 64;
    int K = 1;
    if (n < kTsMinN) return K;
    while (K < 64 && tiles * W * K * 2 <= kTsMaxWaves && n / (W * K * 2) >= kTsMinChunk &&
           2ULL * sizeof(V) * n * K * 2 <= (size_t{4} << 20))
        K *= 2;
    return K;
}

template <typename T, int W, int MODE>
void launch_ts_w(const T *pos, const T *vel, const T *m, T *ts, int n, T eta, T eps2, int K,
                 typename TsAcc<T, MODE>::V *part, hipStream_t s) {
    hipLaunchKernelGGL((timesteps_kernel<T, W, MODE>), dim3((n + 63) / 64, K), dim3(64 * W), 0,
                       s, pos, vel, m, ts, n, eta, eps2, K, K > 1 ? part : nullptr);
}

// Autotuned (W, K) per mode and sink-count size class (floor(log2 n)),
// encoded W << 8 | K (0 = the heuristic above), and the sweep's override.
std::atomic<int> g_ts_tuned[3][32];
std::atomic<int> g_ts_force_w{0}, g_ts_force_k{0};
int ts_size_class(int n) { return n <= 1 ? 0 : 31 - __builtin_clz(static_cast<unsigned>(n)); }
bool valid_ts_w(int w) { return w == 1 || w == 2 || w == 4 || w == 8; }
bool valid_ts_k(int k) { return k >= 1 && k <= 64 && (k & (k - 1)) == 0; }
// A tuned K is used while its partials stay within 64 MB.
constexpr size_t kTsMaxTunedPartial = size_t{64} << 20;

// The launch configuration: the heuristic, or the tuned (or forced) entry
// for FP64 storage while its partials stay within the cap.
template <typename T, int MODE> void ts_config(int n, int &W, int &K) {
    using V = typename TsAcc<T, MODE>::V;
    W = pick_waves(n);
    K = pick_ts_split<V>(n, W);
    if constexpr (std::is_same<T, double>::value) {
        const int t = g_ts_tuned[MODE][ts_size_class(n)].load();
        int w = t >> 8, k = t & 0xff;
        if (g_ts_force_w.load()) w = g_ts_force_w.load();
        if (g_ts_force_k.load()) k = g_ts_force_k.load();
        if (valid_ts_w(w) && valid_ts_k(k) &&
            2ULL * sizeof(V) * static_cast<size_t>(n) * k <= kTsMaxTunedPartial) {
            W = w;
            K = k;
        }
    }
}

template <typename T, int MODE>
int launch_ts(const T *pos, const T *vel, const T *m, T *ts, int n, T eta, T eps2,
              hipStream_t s) {
    using V = typename TsAcc<T, MODE>::V;
    HHS_REQUIRE(n >= 0, "timesteps: negative n");
    if (n == 0) return HHS_OK;
    HHS_REQUIRE(pos && vel && m && ts, "timesteps: null array");
    int W, K;
    ts_config<T, MODE>(n, W, K);
    static thread_local ThreadScratch scratch;
    V *part = nullptr;
    if (K > 1) {
        void *p = nullptr;
        int rc = scratch.acquire(2 * sizeof(V) * static_cast<size_t>(n) * K, s, &p);
        if (rc) return rc;
        part = static_cast<V *>(p);
    }
    switch (W) {
    case 8: launch_ts_w<T, 8, MODE>(pos, vel, m, ts, n, eta, eps2, K, part, s); break;
    case 4: launch_ts_w<T, 4, MODE>(pos, vel, m, ts, n, eta, eps2, K, part, s); break;
    case 2: launch_ts_w<T, 2, MODE>(pos, vel, m, ts, n, eta, eps2, K, part, s); break;
    default: launch_ts_w<T, 1, MODE>(pos, vel, m, ts, n, eta, eps2, K, part, s); break;
    }
    HHS_LAUNCH_CHECK("timesteps_kernel");
    if (K > 1) {
        hipLaunchKernelGGL((timesteps_merge_kernel<T, MODE>), dim3((n + 255) / 256), dim3(256), 0,
                           s, ts, n, K, part);
        HHS_LAUNCH_CHECK("timesteps_merge_kernel");
        return scratch.release(s);
    }
    return HHS_OK;
}

}  // namespace

void ts_tuned_install(int mode, int cls, int w, int k) { g_ts_tuned[mode][cls].store(w << 8 | k); }
void ts_tuned_clear() {
    for (auto &row : g_ts_tuned)
        for (auto &t : row) t.store(0);
}
bool ts_tuned_valid(int mode, int cls, int w, int k) {
    return mode >= HHS_TIMESTEP_EXACT && mode <= HHS_TIMESTEP_FP32 && cls >= 0 && cls < 32 &&
           valid_ts_w(w) && valid_ts_k(k);
}

template <typename T>
int launch_timesteps(const T *pos, const T *vel, const T *m, T *ts, int n, T eta, T eps2,
                     hipStream_t s) {
    return launch_ts<T, HHS_TIMESTEP_EXACT>(pos, vel, m, ts, n, eta, eps2, s);
}

// Fast FP64 form (pair_timestep_fast): rsqrt + Newton instead of IEEE
// sqrt/divide, one division per sink.  Agrees with the exact form to a few
// ulp; not bit-exact.
int launch_timesteps_mode(int mode, const double *pos, const double *vel, const double *m,
                          double *ts, int n, double eta, double eps2, hipStream_t s) {
    switch (mode) {
    case HHS_TIMESTEP_EXACT: return launch_ts<double, HHS_TIMESTEP_EXACT>(pos, vel, m, ts, n, eta, eps2, s);
    case HHS_TIMESTEP_FAST: return launch_ts<double, HHS_TIMESTEP_FAST>(pos, vel, m, ts, n, eta, eps2, s);
    case HHS_TIMESTEP_FP32: return launch_ts<double, HHS_TIMESTEP_FP32>(pos, vel, m, ts, n, eta, eps2, s);
    }
    set_error("timesteps: unknown mode %d", mode);
    return HHS_ERR_INVALID;
}

namespace {
// ---------------------------------------------------- symmetric evaluation
// tau_ij = tau_ji bit for bit (pair_tau, pair_tau_ds), so each unordered
// pair is evaluated once and feeds the minima of both particles.  A
// workgroup (4 waves) takes a block of sinks, S per lane, against a range
// of sources (one wave-uniform source at a time, scalar loads).  It keeps
// the row minima of its sinks in registers and reduces each source's column
// minimum across the wave (shuffles) and the 4 waves (LDS).  Each partial
// minimum is written exactly once and a second kernel takes the minimum of
// a particle's partials.  No atomics; the minimum of a set does not depend
// on the order, so the exact mode stays bit-exact.
//
// Exact mode: blocks of 1024 (4 sinks per lane); workgroup (I, J), I <= J,
// takes sink block I against source block J, and P[K][i] is the minimum of
// particle i over block K:
//   (I, J), I < J: P[J][i] for i in I (rows), P[I][j] for j in J (columns);
//   (I, I): P[I][i], all j != i of the block (rows only).
//
// FP32 mode (timesteps_sym32_kernel, below): double-single records
// (pair_math.h, packed once per launch by pack_ds_kernel), whose
// separations cost 3 FP32 operations instead of an FP64 subtract and a
// conversion; a source is one 64-byte scalar load, fetched one iteration
// ahead.  With 3 sinks per lane (67 VGPRs, 7 waves per SIMD) and 768-source
// blocks it measured 11.2 ms at n = 131072 on MI355X against 14.0 ms for
// FP64 separations with 4 sinks per lane (5 waves).  The column minimum
// stays on LDS permutes (__shfl_xor): the kernel is VALU-bound, and a DPP
// reduction, VALU only, measured 12.2 ms against 10.9.
constexpr int kSymSpl = 4;                      // exact mode: sinks per lane
constexpr int kSymBlock = 64 * 4 * kSymSpl;     // 1024 particles per block

template <typename V> __device__ __forceinline__ V vmin(V a, V b) { return b < a ? b : a; }

template <typename V> __device__ __forceinline__ V wave_min(V x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = vmin(x, __shfl_xor(x, d, 64));
    return x;
}

// Minimum over the wave of values in [+0, 1e30]: for non-negative floats the
// order of the bit patterns as unsigned integers is the order of the values,
// so each step is one v_min_u32 (a float minimum of the permuted value would
// first have to quiet a possible signalling NaN).
__device__ __forceinline__ float wave_min_nonneg(float x) {
    unsigned u = __float_as_uint(x);
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) u = min(u, static_cast<unsigned>(__shfl_xor(static_cast<int>(u), d, 64)));
    return __uint_as_float(u);
}


__global__ __launch_bounds__(256) void timesteps_sym_kernel(const double *__restrict__ pos,
                                                            const double *__restrict__ vel,
                                                            const double *__restrict__ m, int n,
                                                            double eta, double eps2,
                                                            double *__restrict__ part) {
    const int I = blockIdx.x, J = blockIdx.y;
    if (I > J) return;
    __shared__ double colmin[4][kSymBlock];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const double c_ff = eta / TsConst<double>::sqrt2;
    const double huge = TsConst<double>::huge;
    double xi[kSymSpl], yi[kSymSpl], zi[kSymSpl], vxi[kSymSpl], vyi[kSymSpl], vzi[kSymSpl],
        mi[kSymSpl], row[kSymSpl];
    int ii[kSymSpl];
#pragma unroll
    for (int s = 0; s < kSymSpl; ++s) {
        ii[s] = I * kSymBlock + s * 256 + wave * 64 + lane;
        const int ic = ii[s] < n ? ii[s] : n - 1;
        xi[s] = pos[3 * ic];
        yi[s] = pos[3 * ic + 1];
        zi[s] = pos[3 * ic + 2];
        vxi[s] = vel[3 * ic];
        vyi[s] = vel[3 * ic + 1];
        vzi[s] = vel[3 * ic + 2];
        mi[s] = m[ic];
        row[s] = huge;
    }
    const int j0 = J * kSymBlock;
    const int j1 = min(j0 + kSymBlock, n);
    if (I == J) {
        for (int j = j0; j < j1; ++j) {
            const double xj = pos[3 * j], yj = pos[3 * j + 1], zj = pos[3 * j + 2];
            const double vxj = vel[3 * j], vyj = vel[3 * j + 1], vzj = vel[3 * j + 2];
            const double mj = m[j];
#pragma unroll
            for (int s = 0; s < kSymSpl; ++s) {
                const double v = pair_tau<double>(xi[s], yi[s], zi[s], vxi[s], vyi[s], vzi[s],
                                                  mi[s], xj, yj, zj, vxj, vyj, vzj, mj, eta, c_ff,
                                                  eps2);
                row[s] = (j != ii[s]) ? vmin(row[s], v) : row[s];
            }
        }
    } else {
        for (int j = j0; j < j1; ++j) {
            const double xj = pos[3 * j], yj = pos[3 * j + 1], zj = pos[3 * j + 2];
            const double vxj = vel[3 * j], vyj = vel[3 * j + 1], vzj = vel[3 * j + 2];
            const double mj = m[j];
            double c = huge;
#pragma unroll
            for (int s = 0; s < kSymSpl; ++s) {
                const double v = pair_tau<double>(xi[s], yi[s], zi[s], vxi[s], vyi[s], vzi[s],
                                                  mi[s], xj, yj, zj, vxj, vyj, vzj, mj, eta, c_ff,
                                                  eps2);
                row[s] = vmin(row[s], v);
                c = vmin(c, v);  // I < J: block I is full, every ii[s] < n
            }
            c = wave_min(c);
            if (lane == 0) colmin[wave][j - j0] = c;
        }
        __syncthreads();
        for (int jj = threadIdx.x; jj < j1 - j0; jj += 256) {
            const double c =
                vmin(vmin(colmin[0][jj], colmin[1][jj]), vmin(colmin[2][jj], colmin[3][jj]));
            part[static_cast<size_t>(I) * n + j0 + jj] = c;  // P[I][j]
        }
    }
#pragma unroll
    for (int s = 0; s < kSymSpl; ++s)
        if (ii[s] < n) part[static_cast<size_t>(J) * n + ii[s]] = row[s];  // P[J][i]
}

// Double-single records of particles [0, n) (pair_math.h, kDsRec floats each).
// With a guard fix stage it also clears the stage's counters (gfix_hdr).
__global__ __launch_bounds__(256) void pack_ds_kernel(const double *__restrict__ pos,
                                                      const double *__restrict__ vel,
                                                      const double *__restrict__ m, int n,
                                                      float4 *__restrict__ rec, int *gfix_hdr) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i == 0 && gfix_hdr) {
        gfix_hdr[0] = 0;  // flagged timesteps
        gfix_hdr[1] = 0;  // more than kGuardFixCap of them
        gfix_hdr[2] = 0;  // guard_partial_kernel workgroups done
    }
    if (i >= n) return;
    float v[kDsRec];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double x = pos[3 * i + k], w = vel[3 * i + k];
        const float xh = static_cast<float>(x), wh = static_cast<float>(w);
        v[2 * k] = xh;
        v[2 * k + 1] = static_cast<float>(x - static_cast<double>(xh));
        v[6 + 2 * k] = wh;
        v[6 + 2 * k + 1] = static_cast<float>(w - static_cast<double>(wh));
    }
    v[12] = static_cast<float>(m[i]);
    v[13] = v[14] = v[15] = 0.0f;
    float4 *o = rec + 4 * static_cast<size_t>(i);
#pragma unroll
    for (int q = 0; q < 4; ++q) o[q] = make_float4(v[4 * q], v[4 * q + 1], v[4 * q + 2], v[4 * q + 3]);
}

// FP32 symmetric kernel.  Particles are cut into units of U sources and
// sink blocks of 256 S particles (S sinks per lane), R = 256 S / U units
// each; workgroup w takes sink block I against the U sources of unit
// u >= R I -- the triangle of blocks against units, enumerated row by row
// (row I holds nu - R I workgroups), so no workgroup is launched empty and
// the work comes in units small enough to keep the end of the launch full.
// Partial minima, each written exactly once:
//   u >= R (I + 1) (sources above the block): prow[u][i] for the block's
//     sinks i, and pcol[I][j] for the unit's sources j;
//   u in [R I, R I + R) (the block's own units): prow[u][i], j != i.
// Particle i of block I takes the minimum of prow[u][i] over u in [R I, nu)
// and of pcol[K][i] over K < I.  Each unordered pair's two timesteps are the
// same bits whichever particle is the sink (the separations only change
// sign), and a minimum does not depend on order, so U changes no bit.
__device__ __forceinline__ long long sym_row_start(int I, int nu, int S) {
    return static_cast<long long>(I) * nu - static_cast<long long>(S) * I * (I - 1) / 2;
}
template <int S>
__device__ __forceinline__ void sym_tile(int w, int nu, int nb, int &I, int &u) {
    const double h = nu + 0.5 * S;
    int r = static_cast<int>((h - sqrt(h * h - 2.0 * S * static_cast<double>(w))) / S);
    r = r < 0 ? 0 : (r >= nb ? nb - 1 : r);
    while (r + 1 < nb && sym_row_start(r + 1, nu, S) <= w) ++r;
    while (r > 0 && sym_row_start(r, nu, S) > w) --r;
    I = r;
    u = S * r + static_cast<int>(w - sym_row_start(r, nu, S));
}

template <int S, int U, bool SOFT>
__global__ __launch_bounds__(256) void timesteps_sym32_kernel(const float *__restrict__ rec, int n,
                                                              int nu, int nb, float eta,
                                                              float c_ff, float eps2,
                                                              float *__restrict__ prow,
                                                              float *__restrict__ pcol) {
    constexpr int R = 256 * S / U;  // units of U sources per sink block
    static_assert(R * U == 256 * S, "a sink block holds whole units");
    int I, u;
    sym_tile<R>(blockIdx.x, nu, nb, I, u);
    I = __builtin_amdgcn_readfirstlane(I);
    u = __builtin_amdgcn_readfirstlane(u);
    __shared__ float colmin[4][U];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const float huge = TsConst<float>::huge;
    float a[S][13], row[S];
    int ii[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        ii[s] = I * 256 * S + s * 256 + wave * 64 + lane;
        const int ic = ii[s] < n ? ii[s] : n - 1;
        const float *r = rec + kDsRec * static_cast<size_t>(ic);
#pragma unroll
        for (int k = 0; k < 13; ++k) a[s][k] = r[k];
        row[s] = huge;
    }
    const int j0 = u * U;
    const int j1 = min(j0 + U, n);
    if (u < R * (I + 1)) {
        for (int j = j0; j < j1; ++j) {
            const float *b = rec + kDsRec * static_cast<size_t>(j);
#pragma unroll
            for (int s = 0; s < S; ++s) {
                float t_ff, t_fb;
                pair_tau_ds<SOFT>(a[s], b, eta, c_ff, eps2, t_ff, t_fb);
                row[s] = (j != ii[s]) ? fold_tau_ds(row[s], t_ff, t_fb) : row[s];
            }
        }
    } else {
        // the source record one iteration ahead (j1 > j0: unit u is not empty)
        float nb_[13];
#pragma unroll
        for (int k = 0; k < 13; ++k) nb_[k] = rec[kDsRec * static_cast<size_t>(j0) + k];
        for (int j = j0; j < j1; ++j) {
            float b[13];
#pragma unroll
            for (int k = 0; k < 13; ++k) b[k] = nb_[k];
            const int jn = j + 1 < j1 ? j + 1 : j;
#pragma unroll
            for (int k = 0; k < 13; ++k) nb_[k] = rec[kDsRec * static_cast<size_t>(jn) + k];
            float c = huge;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                float t_ff, t_fb;
                pair_tau_ds<SOFT>(a[s], b, eta, c_ff, eps2, t_ff, t_fb);
                row[s] = fold_tau_ds(row[s], t_ff, t_fb);
                c = fold_tau_ds(c, t_ff, t_fb);  // the block lies below unit u: every ii[s] < n
            }
            c = wave_min_nonneg(c);
            if (lane == 0) colmin[wave][j - j0] = c;
        }
        __syncthreads();
        const int jj = threadIdx.x;
        if (jj < U && jj < j1 - j0) {
            const float c =
                vmin(vmin(colmin[0][jj], colmin[1][jj]), vmin(colmin[2][jj], colmin[3][jj]));
            pcol[static_cast<size_t>(I) * n + j0 + jj] = c;
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s)
        if (ii[s] < n) prow[static_cast<size_t>(u) * n + ii[s]] = row[s];
}

// ------------------------------------------------ decision guard stage
// The HOLD split's decision guard (pair_math.h) recomputes with the exact
// formula every timestep within its band of a pivot, each over all n
// sources on one workgroup of the split -- at n = 65536 ... 131072 about
// 190 us per call, on the look-back's critical path.  After the FP32
// symmetric kernel the flagged timesteps are instead listed by the reduce
// kernel and recomputed by the whole GPU: guard_partial_kernel gives each
// workgroup 256 sources and every flagged particle one exact pair per
// thread (workgroup minimum per particle), and its last workgroup to finish
// takes the minimum over the workgroups and writes it into ts.  A minimum
// does not depend on order, so the values are the guard's bit for bit.  More
// than kGuardFixCap flagged timesteps set the overflow word (in the reduce
// kernel), and the split then runs its own guard as before (on every flagged
// timestep; the fixed ones give the same values again).
constexpr int kGuardFixCap = 256;

struct GuardFix {
    int on;
    TsGuard g;
    double eta, c_ff, eps2;
    int *hdr;          // [0] flagged count, [1] overflow, [2] guard workgroups done
    int *list;         // flagged particles (n entries)
    double *partial;   // [kGuardFixCap][groups]
};

// tau_i from the FP32 kernel's partials (the layout above); with the guard
// stage, the timesteps within the band of a pivot are listed.
// 64 particles per workgroup, one per lane; its kSymRedSlices waves take
// every kSymRedSlices-th row of the partials (a particle has up to n / U of
// them, too many for one thread's serial loop) and are combined in LDS.
// The result is the minimum over the same values in any grouping.
constexpr int kSymRedSlices = 16;
template <int S, int U>
__global__ __launch_bounds__(64 * kSymRedSlices) void timesteps_sym32_reduce_kernel(
    const float *__restrict__ prow, const float *__restrict__ pcol, int n, int nu,
    double *__restrict__ ts, GuardFix gf) {
    constexpr int R = 256 * S / U;
    __shared__ float part[kSymRedSlices][64];
    const int lane = threadIdx.x & 63;
    const int slice = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int i = blockIdx.x * 64 + lane;
    const int ic = i < n ? i : n - 1;
    const int I = ic / (256 * S);
    float best = TsConst<float>::huge;
    for (int u = R * I + slice; u < nu; u += kSymRedSlices)
        best = vmin(best, prow[static_cast<size_t>(u) * n + ic]);
    for (int K = slice; K < I; K += kSymRedSlices)
        best = vmin(best, pcol[static_cast<size_t>(K) * n + ic]);
    part[slice][lane] = best;
    __syncthreads();
    if (slice != 0 || i >= n) return;
#pragma unroll
    for (int q = 1; q < kSymRedSlices; ++q) best = vmin(best, part[q][lane]);
    const double v = static_cast<double>(best);
    ts[i] = v;
    if (gf.on && guard_near_pivot(gf.g, v)) {
        const int k = atomicAdd(gf.hdr, 1);
        if (k < n) gf.list[k] = i;
        if (k >= kGuardFixCap) gf.hdr[1] = 1;  // the split runs its own guard
    }
}

// Workgroup g: sources [256 g, 256 g + 256) against every listed particle;
// the last workgroup to finish then takes each listed particle's minimum
// over the workgroups and writes it into ts.  No timestep is listed in most
// calls: every workgroup then returns at once, with no atomic and no fence.
// Otherwise each workgroup's partials are released (device scope) before
// it counts itself done, and the last one acquires before it reads them.
__global__ __launch_bounds__(256) void guard_partial_kernel(const double *__restrict__ pos,
                                                            const double *__restrict__ vel,
                                                            const double *__restrict__ m, int n,
                                                            double *__restrict__ ts, GuardFix gf) {
    __shared__ double s_min[4];
    __shared__ int s_last;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int cnt = min(__builtin_amdgcn_readfirstlane(gf.hdr[0]), kGuardFixCap);
    if (cnt == 0) return;
    const int j = blockIdx.x * 256 + t;
    for (int f = 0; f < cnt; ++f) {
        const int s = __builtin_amdgcn_readfirstlane(gf.list[f]);
        double best = TsConst<double>::huge;
        if (j < n)
            pair_timestep<double>(pos[3 * s], pos[3 * s + 1], pos[3 * s + 2], vel[3 * s],
                                  vel[3 * s + 1], vel[3 * s + 2], m[s], pos[3 * j],
                                  pos[3 * j + 1], pos[3 * j + 2], vel[3 * j], vel[3 * j + 1],
                                  vel[3 * j + 2], m[j], gf.eta, gf.c_ff, gf.eps2, j != s, best);
        best = wave_min(best);
        if (lane == 0) s_min[wave] = best;
        __syncthreads();
        if (t == 0)
            gf.partial[static_cast<size_t>(f) * gridDim.x + blockIdx.x] =
                vmin(vmin(s_min[0], s_min[1]), vmin(s_min[2], s_min[3]));
        __syncthreads();
    }
    if (t == 0) {
        __threadfence();  // this workgroup's partials, before it counts itself done
        s_last = atomicAdd(gf.hdr + 2, 1) == static_cast<int>(gridDim.x) - 1 ? 1 : 0;
    }
    __syncthreads();
    if (!s_last) return;
    __threadfence();  // every workgroup's partials
    const int groups = static_cast<int>(gridDim.x);
    for (int f = 0; f < cnt; ++f) {
        double best = TsConst<double>::huge;
        for (int c = t; c < groups; c += 256)
            best = vmin(best, gf.partial[static_cast<size_t>(f) * groups + c]);
        best = wave_min(best);
        if (lane == 0) s_min[wave] = best;
        __syncthreads();
        if (t == 0) ts[gf.list[f]] = vmin(vmin(s_min[0], s_min[1]), vmin(s_min[2], s_min[3]));
        __syncthreads();
    }
}

template <typename V>
__global__ __launch_bounds__(256) void timesteps_sym_reduce_kernel(const V *__restrict__ part,
                                                                   int nb, int n,
                                                                   double *__restrict__ ts) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    V best = part[i];
    for (int k = 1; k < nb; ++k) best = vmin(best, part[static_cast<size_t>(k) * n + i]);
    ts[i] = static_cast<double>(best);
}

// Sinks per lane of the FP32 kernel, measured on MI355X (ms, S = 1 / 2 /
// 3: n = 16384 0.299 / 0.310 / 0.290, 32768 0.942 / 0.901 / 0.856, 65536
// 3.36 / 3.20 / 3.16, 131072 12.25 / 11.29 / 10.86).  With compare-and-select
// minima four sinks per lane won from 65536 particles (131072: 10.97 against
// 10.87), since larger blocks spend less per pair on the column minima.
// With the minima as v_min3_f32 folds and v_min_u32 wave steps, three win
// at every size (S = 3 / 4: 16384 0.276 / 0.339, 32768 0.749 / 0.837, 65536
// 2.77 / 2.81, 131072 10.25 / 10.57; same bits).
// hhs_timesteps_sym_set_sinks_per_lane forces 3 or 4 (0: the default, 3);
// the workspace is sized for 3, which needs the most.
std::atomic<int> g_sym32_spl{0};
int sym32_spl(int) {
    const int v = g_sym32_spl.load();
    return v == 4 ? 4 : 3;
}

// Source units of the FP32 kernel.  Below ~32768 particles 256-source units
// leave too few workgroups, each a long serial pass (n = 8192: 132
// workgroups); units of 64 sources give 4x as many, shorter ones, for the
// same bits.  MI355X, kernel and reduce, us, U = 64 / 128 / 256 (one-sided
// FP32 kernel): n = 6144 62 / 112 / 205 (62), 8192 76 / 115 / 206 (101),
// 12288 141 / 163 / 220 (215), 16384 215 / 220 / 254 (358), 32768 688 /
// 698 / 729, 65536 2628 / 2643 / 2685, 131072 10362 / 10282 / 10240.
// Units narrower than 256 are used up to kSymNarrowMaxN particles (their row
// partials grow as n^2 / U).  hhs_timesteps_sym_set_unit forces 64, 128 or
// 256 (0: by size).
constexpr int kSymNarrowMaxN = 65536;
std::atomic<int> g_sym32_unit{0};
int sym32_unit(int n) {
    const int f = g_sym32_unit.load();
    if (f == 256 || ((f == 64 || f == 128) && n <= kSymNarrowMaxN)) return f;
    return n <= kSymNarrowMaxN ? 64 : 256;
}

// Workspace of the FP32 kernel: the packed records (64 B per particle),
// then prow (ceil(n / U) * n floats) and pcol (one row per sink block,
// ceil(n / (256 S)) * n floats; 358 MB in all at n = 131072).  Sized for
// the narrowest unit any count up to n may use, so that a workspace sized
// for n serves every smaller count.
size_t sym32_rec_bytes(int n) {
    return (static_cast<size_t>(n) * kDsRec * sizeof(float) + 255) & ~static_cast<size_t>(255);
}
size_t sym32_region_bytes(size_t n, size_t U) {
    const size_t S = 3;  // the largest workspace of the selectable sinks per lane
    const size_t nu = (n + U - 1) / U, nb = (n + 256 * S - 1) / (256 * S);
    return ((nu + nb) * n * sizeof(float) + 255) & ~static_cast<size_t>(255);
}
size_t sym32_core_bytes(int n) {
    const size_t nn = static_cast<size_t>(n);
    const size_t narrow = sym32_region_bytes(std::min<size_t>(nn, kSymNarrowMaxN), 64);
    const size_t wide = sym32_region_bytes(nn, 256);
    return sym32_rec_bytes(n) + std::max(narrow, wide);
}
// the guard stage: 256-byte header, the list (n ints), the partial minima
size_t sym32_guard_bytes(int n) {
    const size_t groups = (static_cast<size_t>(n) + 255) / 256;
    return 256 + ((static_cast<size_t>(n) * sizeof(int) + 255) & ~static_cast<size_t>(255)) +
           static_cast<size_t>(kGuardFixCap) * groups * sizeof(double);
}
size_t sym32_ws_bytes(int n) { return sym32_core_bytes(n) + sym32_guard_bytes(n); }

template <int S, int U>
void launch_sym32(const float *rec, int n, float eta, float cff, float e2, float *prow,
                  double *ts, const GuardFix &gf, hipStream_t s) {
    constexpr int R = 256 * S / U;
    const int nu = (n + U - 1) / U;
    float *pcol = prow + static_cast<size_t>(nu) * n;
    const int nb = (n + 256 * S - 1) / (256 * S);
    long long w = 0;
    for (int I = 0; I < nb; ++I) w += nu - R * I;
    if (e2 > 0.0f)
        hipLaunchKernelGGL((timesteps_sym32_kernel<S, U, true>), dim3(static_cast<unsigned>(w)),
                           dim3(256), 0, s, rec, n, nu, nb, eta, cff, e2, prow, pcol);
    else
        hipLaunchKernelGGL((timesteps_sym32_kernel<S, U, false>), dim3(static_cast<unsigned>(w)),
                           dim3(256), 0, s, rec, n, nu, nb, eta, cff, e2, prow, pcol);
    hipLaunchKernelGGL((timesteps_sym32_reduce_kernel<S, U>), dim3((n + 63) / 64),
                       dim3(64 * kSymRedSlices), 0, s, prow, pcol, n, nu, ts, gf);
}
template <int S>
void launch_sym32_u(int U, const float *rec, int n, float eta, float cff, float e2, float *prow,
                    double *ts, const GuardFix &gf, hipStream_t s) {
    switch (U) {
    case 64: launch_sym32<S, 64>(rec, n, eta, cff, e2, prow, ts, gf, s); break;
    case 128: launch_sym32<S, 128>(rec, n, eta, cff, e2, prow, ts, gf, s); break;
    default: launch_sym32<S, 256>(rec, n, eta, cff, e2, prow, ts, gf, s); break;
    }
}

}  // namespace

size_t timesteps_sym_ws_bytes(int n) {
    if (n <= 0) return 0;
    const size_t nb = (static_cast<size_t>(n) + kSymBlock - 1) / kSymBlock;
    const size_t exact = nb * static_cast<size_t>(n) * sizeof(double);
    const size_t fp32 = sym32_ws_bytes(n);
    return exact > fp32 ? exact : fp32;
}

int launch_timesteps_sym(int mode, const double *pos, const double *vel, const double *m,
                         double *ts, int n, double eta, double eps2, void *ws, size_t ws_bytes,
                         hipStream_t s, const TsGuard *guard, const int **guard_overflow) {
    if (guard_overflow) *guard_overflow = nullptr;
    HHS_REQUIRE(n >= 0, "timesteps: negative n");
    if (n == 0) return HHS_OK;
    if (mode == HHS_TIMESTEP_FAST)
        return launch_timesteps_mode(mode, pos, vel, m, ts, n, eta, eps2, s);
    HHS_REQUIRE(mode == HHS_TIMESTEP_EXACT || mode == HHS_TIMESTEP_FP32,
                "timesteps: unknown mode %d", mode);
    HHS_REQUIRE(pos && vel && m && ts, "timesteps: null array");
    HHS_REQUIRE(ws && ws_bytes >= timesteps_sym_ws_bytes(n),
                "timesteps: symmetric workspace of %zu bytes, need %zu", ws_bytes,
                timesteps_sym_ws_bytes(n));
    const dim3 red((n + 255) / 256);
    if (mode == HHS_TIMESTEP_FP32) {
        float *rec = static_cast<float *>(ws);
        float *prow = reinterpret_cast<float *>(static_cast<char *>(ws) + sym32_rec_bytes(n));
        // the guard stage's area follows the kernel's partial minima
        char *gbase = static_cast<char *>(ws) + sym32_core_bytes(n);
        GuardFix gf{};
        if (guard && guard_overflow) {
            gf.on = 1;
            gf.g = *guard;
            gf.eta = eta;
            gf.c_ff = eta / TsConst<double>::sqrt2;
            gf.eps2 = eps2;
            gf.hdr = reinterpret_cast<int *>(gbase);
            gf.list = reinterpret_cast<int *>(gbase + 256);
            gf.partial = reinterpret_cast<double *>(
                gbase + 256 + ((static_cast<size_t>(n) * sizeof(int) + 255) & ~size_t{255}));
        }
        hipLaunchKernelGGL(pack_ds_kernel, red, dim3(256), 0, s, pos, vel, m, n,
                           reinterpret_cast<float4 *>(rec), gf.on ? gf.hdr : nullptr);
        HHS_LAUNCH_CHECK("pack_ds_kernel");
        const float etaf = static_cast<float>(eta);
        const float cff = static_cast<float>(eta / TsConst<double>::sqrt2);
        const float e2 = static_cast<float>(eps2);
        const int U = sym32_unit(n);
        switch (sym32_spl(n)) {
        case 4: launch_sym32_u<4>(U, rec, n, etaf, cff, e2, prow, ts, gf, s); break;
        default: launch_sym32_u<3>(U, rec, n, etaf, cff, e2, prow, ts, gf, s); break;
        }
        HHS_LAUNCH_CHECK("timesteps_sym32_kernel");
        if (gf.on) {
            hipLaunchKernelGGL(guard_partial_kernel, red, dim3(256), 0, s, pos, vel, m, n, ts, gf);
            HHS_LAUNCH_CHECK("guard_partial_kernel");
            *guard_overflow = gf.hdr + 1;
        }
    } else {
        const int nb = (n + kSymBlock - 1) / kSymBlock;
        double *part = static_cast<double *>(ws);
        hipLaunchKernelGGL(timesteps_sym_kernel, dim3(nb, nb), dim3(256), 0, s, pos, vel, m, n,
                           eta, eps2, part);
        HHS_LAUNCH_CHECK("timesteps_sym_kernel");
        hipLaunchKernelGGL(timesteps_sym_reduce_kernel<double>, red, dim3(256), 0, s, part, nb, n,
                           ts);
    }
    HHS_LAUNCH_CHECK("timesteps_sym_reduce_kernel");
    return HHS_OK;
}

template int launch_timesteps<double>(const double *, const double *, const double *, double *,
                                      int, double, double, hipStream_t);
template int launch_timesteps<float>(const float *, const float *, const float *, float *, int,
                                     float, float, hipStream_t);

}  // namespace hhs

using namespace hhs;

extern "C" {

int hhs_timesteps_f64(const hhs_particles_f64 *sys, int n, double eta, double eps2,
                      void *stream) {
    HHS_REQUIRE(sys, "timesteps: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "timesteps: n=%d outside [0, %d]", n, sys->n);
    return launch_timesteps<double>(sys->pos, sys->vel, sys->m, sys->timestep, n, eta, eps2,
                                    as_stream(stream));
}

int hhs_timesteps_fast_f64(const hhs_particles_f64 *sys, int n, double eta, double eps2,
                           void *stream) {
    HHS_REQUIRE(sys, "timesteps: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "timesteps: n=%d outside [0, %d]", n, sys->n);
    return launch_timesteps_mode(HHS_TIMESTEP_FAST, sys->pos, sys->vel, sys->m, sys->timestep, n,
                                 eta, eps2, as_stream(stream));
}

int hhs_timesteps_mode_f64(const hhs_particles_f64 *sys, int n, double eta, double eps2, int mode,
                           void *stream) {
    HHS_REQUIRE(sys, "timesteps: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "timesteps: n=%d outside [0, %d]", n, sys->n);
    return launch_timesteps_mode(mode, sys->pos, sys->vel, sys->m, sys->timestep, n, eta, eps2,
                                 as_stream(stream));
}

size_t hhs_timesteps_sym_workspace_bytes(int n) { return timesteps_sym_ws_bytes(n); }

int hhs_timesteps_sym_set_unit(int u) {
    HHS_REQUIRE(u == 0 || u == 64 || u == 128 || u == 256,
                "timesteps_sym: unit must be 0, 64, 128 or 256");
    g_sym32_unit.store(u);
    return HHS_OK;
}

int hhs_timesteps_sym_set_sinks_per_lane(int s) {
    HHS_REQUIRE(s == 0 || s == 3 || s == 4, "timesteps_sym: sinks per lane must be 0, 3 or 4");
    g_sym32_spl.store(s);
    return HHS_OK;
}

int hhs_timesteps_sym_f64(const hhs_particles_f64 *sys, int n, double eta, double eps2, int mode,
                          void *ws, size_t ws_bytes, void *stream) {
    HHS_REQUIRE(sys, "timesteps: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "timesteps: n=%d outside [0, %d]", n, sys->n);
    return launch_timesteps_sym(mode, sys->pos, sys->vel, sys->m, sys->timestep, n, eta, eps2, ws,
                                ws_bytes, as_stream(stream));
}

int hhs_timesteps_f32(const hhs_particles_f32 *sys, int n, float eta, float eps2,
                      void *stream) {
    HHS_REQUIRE(sys, "timesteps: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "timesteps: n=%d outside [0, %d]", n, sys->n);
    return launch_timesteps<float>(sys->pos, sys->vel, sys->m, sys->timestep, n, eta, eps2,
                                   as_stream(stream));
}

int hhs_autotune_timesteps(int n, int mode, int reps, const char *cache_path, int *best_w,
                           int *best_k, double *best_ms) {
    HHS_REQUIRE(n >= 64, "autotune_timesteps: n must be >= 64");
    HHS_REQUIRE(mode >= HHS_TIMESTEP_EXACT && mode <= HHS_TIMESTEP_FP32,
                "autotune_timesteps: mode must be HHS_TIMESTEP_EXACT, _FAST or _FP32");
    if (reps <= 0) reps = 3;
    const size_t nn = static_cast<size_t>(n);
    std::vector<double> p(3 * nn), v(3 * nn), mm(nn);
    int rc = hhs_ic_plummer(n, 54321, 0, 1, 1e-8, p.data(), v.data(), mm.data());
    if (rc) return rc;
    double *dev = nullptr;
    HHS_HIP(hipMalloc(&dev, nn * 64));  // pos, vel (3n each), m, timestep
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMemcpy(dev, p.data(), nn * 24, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dev + 3 * nn, v.data(), nn * 24, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dev + 6 * nn, mm.data(), nn * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    const long long tiles = (n + 63) / 64;
    const size_t vbytes = mode == HHS_TIMESTEP_FP32 ? sizeof(float) : sizeof(double);
    double best = 1e300;
    int bw = 0, bk = 0;
    for (int w : {1, 2, 4, 8}) {
        for (int k : {1, 2, 4, 8, 16, 32, 64}) {
            if (e != hipSuccess || rc) break;
            // keep every wave >= 8 sources, the grid <= 16384 waves and the
            // partials within the tuned cap
            if (n / (w * k) < 8 || tiles * w * k > 16384 ||
                2 * vbytes * nn * k > kTsMaxTunedPartial)
                continue;
            g_ts_force_w.store(w);
            g_ts_force_k.store(k);
            auto run = [&]() {
                return launch_timesteps_mode(mode, dev, dev + 3 * nn, dev + 6 * nn, dev + 7 * nn,
                                             n, 0.01, 1e-8, nullptr);
            };
            rc = run();  // warm-up
            if (!rc) e = hipEventRecord(e0, nullptr);
            for (int r = 0; r < reps && !rc && e == hipSuccess; ++r) rc = run();
            if (!rc && e == hipSuccess) e = hipEventRecord(e1, nullptr);
            if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (!rc && e == hipSuccess && ms / reps < best) {
                best = ms / reps;
                bw = w;
                bk = k;
            }
        }
    }
    g_ts_force_w.store(0);
    g_ts_force_k.store(0);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipDeviceSynchronize();
    (void)hipFree(dev);
    if (e != hipSuccess) return hip_status(e, "autotune_timesteps", __FILE__, __LINE__);
    if (rc) return rc;
    const int cls = ts_size_class(n);
    ts_tuned_install(mode, cls, bw, bk);
    if (cache_path) {
        FILE *f = std::fopen(cache_path, "a");
        if (!f) {
            set_error("autotune_timesteps: cannot append to %s", cache_path);
            return HHS_ERR_INVALID;
        }
        std::fprintf(f, "ts %d %d %d %d %.6f\n", mode, cls, bw, bk, best);
        std::fclose(f);
    }
    if (best_w) *best_w = bw;
    if (best_k) *best_k = bk;
    if (best_ms) *best_ms = best;
    return HHS_OK;
}

int hhs_timesteps_get_config(int n, int mode, int *waves_per_group, int *chunks) {
    HHS_REQUIRE(waves_per_group && chunks, "timesteps_get_config: null output");
    HHS_REQUIRE(n > 0, "timesteps_get_config: n must be positive");
    switch (mode) {
    case HHS_TIMESTEP_EXACT: ts_config<double, HHS_TIMESTEP_EXACT>(n, *waves_per_group, *chunks); break;
    case HHS_TIMESTEP_FAST: ts_config<double, HHS_TIMESTEP_FAST>(n, *waves_per_group, *chunks); break;
    case HHS_TIMESTEP_FP32: ts_config<double, HHS_TIMESTEP_FP32>(n, *waves_per_group, *chunks); break;
    default: set_error("timesteps_get_config: unknown mode %d", mode); return HHS_ERR_INVALID;
    }
    return HHS_OK;
}

}  // extern "C"
