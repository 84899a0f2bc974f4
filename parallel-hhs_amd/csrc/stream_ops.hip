// Streaming kernels: drift, parent scatter-back, FP64 energy.
//
// Reference: _drift (src/drift.cuh:16-43), _update_system
// (src/update_system.cuh:10-43).  The energy diagnostic has no reference
// counterpart.  Drift and update are HBM-bound (72 B and 132 B per
// particle); they run without FMA contraction so FP64 results are bit-exact
// against the same expression on the CPU.
#include <cmath>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "hhs_internal.h"

namespace hhs {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void drift_scalar_kernel(T *__restrict__ pos,
                                                           const T *__restrict__ vel, int n3,
                                                           T h) {
#pragma clang fp contract(off)
    const int stride = gridDim.x * blockDim.x;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n3; k += stride)
        pos[k] = pos[k] + vel[k] * h;  // src/drift.cuh:25
}

// Each thread moves 16 bytes of pos and of vel per step (one component
// pair of FP64, or four of FP32, as one vector load), so that a wave keeps
// 1 KB of each array in flight per access; the element expression is the
// scalar one.  For 16-byte aligned arrays (drift_scalar_kernel otherwise);
// an odd tail of components is taken by the first thread.
template <typename T>
__global__ __launch_bounds__(256) void drift_kernel(T *__restrict__ pos,
                                                    const T *__restrict__ vel, int n3, T h) {
#pragma clang fp contract(off)
    constexpr int V = 16 / sizeof(T);  // components per 16-byte vector
    using Vec = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    const int nv = n3 / V;
    const int stride = gridDim.x * blockDim.x;
    Vec *pv = reinterpret_cast<Vec *>(pos);
    const Vec *vv = reinterpret_cast<const Vec *>(vel);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nv; k += stride) {
        Vec p = pv[k];
        const Vec v = vv[k];
        T *pe = reinterpret_cast<T *>(&p);
        const T *ve = reinterpret_cast<const T *>(&v);
#pragma unroll
        for (int c = 0; c < V; ++c) pe[c] = pe[c] + ve[c] * h;  // src/drift.cuh:25
        pv[k] = p;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        for (int k = nv * V; k < n3; ++k) pos[k] = pos[k] + vel[k] * h;
}

template <typename T>
__global__ __launch_bounds__(256) void update_kernel(Particles<T> prev, Particles<T> cur,
                                                     int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // src/update_system.cuh:14-22.  All loads first: prev and cur may not
    // be proven disjoint, and interleaved stores would serialise the loads.
    const int p = cur.parent_id[i];
    const T x = cur.pos[3 * i + 0], y = cur.pos[3 * i + 1], z = cur.pos[3 * i + 2];
    const T vx = cur.vel[3 * i + 0], vy = cur.vel[3 * i + 1], vz = cur.vel[3 * i + 2];
    const T m = cur.m[i], ts = cur.timestep[i];
    prev.pos[3 * p + 0] = x;
    prev.pos[3 * p + 1] = y;
    prev.pos[3 * p + 2] = z;
    prev.vel[3 * p + 0] = vx;
    prev.vel[3 * p + 1] = vy;
    prev.vel[3 * p + 2] = vz;
    prev.m[p] = m;
    prev.timestep[p] = ts;
}

// HOLD's scatter-back of one call (src/hold.cuh:177-178) with the second
// slow half-drift (:158-159) folded in: threads [0, nf) write fast record i
// to prev[fast.parent_id[i]]; threads [nf, nf + ns) write slow record
// i - nf, its position drifted as pos + vel * h (the drift kernel's
// expression).  The fast recursion that runs between the drift and the
// scatter-back never reads the slow set, so deferring the drift to here
// changes no value.
// kp: the slow set's pending kick, applied first as kick_reduce_kernel
// applies it (chunk partials added in order from 0.0, vel + T(a) * h).
template <typename T>
__global__ __launch_bounds__(256) void update_pair_kernel(Particles<T> prev, Particles<T> fast,
                                                          int nf, Particles<T> slow, int ns,
                                                          T h, KickPending kp) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf + ns) return;
    const bool is_fast = t < nf;
    const Particles<T> &c = is_fast ? fast : slow;
    const int i = is_fast ? t : t - nf;
    const int p = c.parent_id[i];
    T x = c.pos[3 * i + 0], y = c.pos[3 * i + 1], z = c.pos[3 * i + 2];
    T vx = c.vel[3 * i + 0], vy = c.vel[3 * i + 1], vz = c.vel[3 * i + 2];
    const T m = c.m[i], ts = c.timestep[i];
    if (!is_fast && kp.part) {
        double ax, ay, az;
        sum_kick_partials(kp.part, kp.K, kp.rows, i, ax, ay, az);
        const T hk = static_cast<T>(kp.h);
        vx = vx + static_cast<T>(ax) * hk;
        vy = vy + static_cast<T>(ay) * hk;
        vz = vz + static_cast<T>(az) * hk;
    }
    if (!is_fast) {
        x = x + vx * h;
        y = y + vy * h;
        z = z + vz * h;
    }
    prev.pos[3 * p + 0] = x;
    prev.pos[3 * p + 1] = y;
    prev.pos[3 * p + 2] = z;
    prev.vel[3 * p + 0] = vx;
    prev.vel[3 * p + 1] = vy;
    prev.vel[3 * p + 2] = vz;
    prev.m[p] = m;
    prev.timestep[p] = ts;
}

// The scatter-backs of a chain of calls in one launch (ScatterChain): thread
// t takes element t of the concatenation slow_0, slow_1, ..., slow_{d-1},
// fast_{d-1} (their sizes add up to the top call's set), applies its level's
// pending kick and half-drift as update_pair_kernel does, and writes it to
// the top call's set through the parent_id maps of the levels above it.
template <typename T>
__global__ __launch_bounds__(256) void update_chain_kernel(Particles<T> prev, ScatterChain<T> c,
                                                           int n) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int k = 0, base = 0;
    while (k < c.depth && t >= base + c.lv[k].ns) {
        base += c.lv[k].ns;
        ++k;
    }
    const bool slow = k < c.depth;
    const int lvl = slow ? k : c.depth - 1;
    const Particles<T> &src = slow ? c.lv[k].slow : c.bottom_fast;
    const int i = t - base;
    int p = src.parent_id[i];
    for (int m = lvl - 1; m >= 0; --m) p = c.lv[m].fast_pid[p];
    T x = src.pos[3 * i + 0], y = src.pos[3 * i + 1], z = src.pos[3 * i + 2];
    T vx = src.vel[3 * i + 0], vy = src.vel[3 * i + 1], vz = src.vel[3 * i + 2];
    if (slow) {
        const KickPending kp = c.lv[k].kick;
        if (kp.part) {
            double ax, ay, az;
            sum_kick_partials(kp.part, kp.K, kp.rows, i, ax, ay, az);
            const T hk = static_cast<T>(kp.h);
            vx = vx + static_cast<T>(ax) * hk;
            vy = vy + static_cast<T>(ay) * hk;
            vz = vz + static_cast<T>(az) * hk;
        }
        const T h = c.lv[k].h;
        x = x + vx * h;
        y = y + vy * h;
        z = z + vz * h;
    }
    prev.pos[3 * p + 0] = x;
    prev.pos[3 * p + 1] = y;
    prev.pos[3 * p + 2] = z;
    prev.vel[3 * p + 0] = vx;
    prev.vel[3 * p + 1] = vy;
    prev.vel[3 * p + 2] = vz;
    prev.m[p] = src.m[i];
    prev.timestep[p] = src.timestep[i];
}

// Row sums of the potential: row[i] = m_i * sum_{j>i} m_j / sqrt(r^2+eps2),
// j ascending (the order of a serial CPU loop), plus the kinetic term of i.
// The j loop is wave-uniform (scalar loads); lanes with j <= i skip.
__global__ __launch_bounds__(256) void energy_rows_kernel(const double *__restrict__ pos,
                                                          const double *__restrict__ vel,
                                                          const double *__restrict__ m, int n,
                                                          double eps2, double *__restrict__ row,
                                                          double *__restrict__ kin) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ic = i < n ? i : n - 1;
    const double xi = pos[3 * ic], yi = pos[3 * ic + 1], zi = pos[3 * ic + 2];
    const int wave_first = __builtin_amdgcn_readfirstlane(blockIdx.x * blockDim.x +
                                                          (threadIdx.x & ~63));
    double s = 0.0;
    for (int j = wave_first + 1; j < n; ++j) {
        const double dx = pos[3 * j] - xi;
        const double dy = pos[3 * j + 1] - yi;
        const double dz = pos[3 * j + 2] - zi;
        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
        const double t = m[j] / sqrt(r2);
        if (j > i && r2 != 0.0) s += t;
    }
    if (i < n) {
        // an infinite coordinate would only zero its pair terms: mark the
        // row so that the sum reports the state as non-finite
        const bool fin = isfinite(xi) && isfinite(yi) && isfinite(zi);
        row[i] = fin ? m[i] * s : __builtin_nan("");
        const double vx = vel[3 * i], vy = vel[3 * i + 1], vz = vel[3 * i + 2];
        kin[i] = 0.5 * m[i] * (vx * vx + vy * vy + vz * vz);
    }
}

// Fixed-order reduction of row[] and kin[] in one workgroup: lane t sums
// indices t, t+1024, ... serially, then a fixed tree combines the lanes.
__global__ __launch_bounds__(1024) void energy_reduce_kernel(const double *__restrict__ row,
                                                             const double *__restrict__ kin,
                                                             int n, double *__restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double sp[1024], sk[1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int i = t; i < n; i += 1024) {
        a += row[i];
        b += kin[i];
    }
    sp[t] = a;
    sk[t] = b;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (t < w) {
            sp[t] += sp[t + w];
            sk[t] += sk[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[0] = sk[0];
        out[1] = -sp[0];
    }
}

}  // namespace

template <typename T> int launch_drift(T *pos, const T *vel, int n, T h, hipStream_t s) {
    HHS_REQUIRE(n >= 0, "drift: negative n");
    if (n == 0) return HHS_OK;
    HHS_REQUIRE(pos && vel, "drift: null array");
    const int n3 = 3 * n;
    if (reinterpret_cast<uintptr_t>(pos) % 16 != 0 || reinterpret_cast<uintptr_t>(vel) % 16 != 0) {
        int blocks = (n3 + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL((drift_scalar_kernel<T>), dim3(blocks), dim3(256), 0, s, pos, vel, n3,
                           h);
        HHS_LAUNCH_CHECK("drift_scalar_kernel");
        return HHS_OK;
    }
    // one vector per thread: a capped grid left part of the last grid-stride
    // pass on half the threads (2^24 particles: 0.66 of HBM against 0.89 at
    // 2^22)
    const int nv = n3 / static_cast<int>(16 / sizeof(T));
    const int blocks = nv > 0 ? (nv + 255) / 256 : 1;
    hipLaunchKernelGGL((drift_kernel<T>), dim3(blocks), dim3(256), 0, s, pos, vel, n3, h);
    HHS_LAUNCH_CHECK("drift_kernel");
    return HHS_OK;
}

template <typename T>
int launch_update(const Particles<T> &prev, const Particles<T> &cur, int n, hipStream_t s) {
    HHS_REQUIRE(n >= 0, "update: negative n");
    if (n == 0) return HHS_OK;  // src/update_system.cuh:32
    HHS_REQUIRE(prev.pos && prev.vel && prev.m && prev.timestep, "update: null field in prev");
    HHS_REQUIRE(cur.pos && cur.vel && cur.m && cur.timestep && cur.parent_id,
                "update: null field in cur");
    hipLaunchKernelGGL((update_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, s, prev, cur, n);
    HHS_LAUNCH_CHECK("update_kernel");
    return HHS_OK;
}

template int launch_drift<double>(double *, const double *, int, double, hipStream_t);
template int launch_drift<float>(float *, const float *, int, float, hipStream_t);
template int launch_update<double>(const Particles<double> &, const Particles<double> &, int,
                                   hipStream_t);
template int launch_update<float>(const Particles<float> &, const Particles<float> &, int,
                                  hipStream_t);

template <typename T>
int launch_update_pair(const Particles<T> &prev, const Particles<T> &fast, int nf,
                       const Particles<T> &slow, int ns, T h_slow, hipStream_t s,
                       const KickPending &slow_kick) {
    HHS_REQUIRE(nf >= 0 && ns >= 0, "update_pair: negative size");
    const int n = nf + ns;
    if (n == 0) return HHS_OK;
    HHS_REQUIRE(!slow_kick.part || slow_kick.rows == ns, "update_pair: pending kick of %d rows",
                slow_kick.rows);
    hipLaunchKernelGGL((update_pair_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, s, prev,
                       fast, nf, slow, ns, h_slow, slow_kick);
    HHS_LAUNCH_CHECK("update_pair_kernel");
    return HHS_OK;
}
template <typename T>
int launch_update_chain(const Particles<T> &prev, const ScatterChain<T> &c, hipStream_t s) {
    HHS_REQUIRE(c.depth >= 1 && c.depth <= kScatterChainMax, "update_chain: depth %d", c.depth);
    long long n = 0;
    for (int k = 0; k < c.depth; ++k) {
        HHS_REQUIRE(c.lv[k].ns >= 0 && c.lv[k].nf >= 0, "update_chain: negative size");
        HHS_REQUIRE(!c.lv[k].kick.part || c.lv[k].kick.rows == c.lv[k].ns,
                    "update_chain: pending kick of %d rows", c.lv[k].kick.rows);
        // each level's set is the slow set plus the next level's set
        if (k + 1 < c.depth)
            HHS_REQUIRE(c.lv[k].nf == c.lv[k + 1].ns + c.lv[k + 1].nf,
                        "update_chain: level %d has %d fast records, level %d holds %d", k,
                        c.lv[k].nf, k + 1, c.lv[k + 1].ns + c.lv[k + 1].nf);
        n += c.lv[k].ns;
    }
    n += c.lv[c.depth - 1].nf;
    if (n == 0) return HHS_OK;
    hipLaunchKernelGGL((update_chain_kernel<T>), dim3(static_cast<unsigned>((n + 255) / 256)),
                       dim3(256), 0, s, prev, c, static_cast<int>(n));
    HHS_LAUNCH_CHECK("update_chain_kernel");
    return HHS_OK;
}
template int launch_update_chain<double>(const Particles<double> &,
                                         const ScatterChain<double> &, hipStream_t);
template int launch_update_chain<float>(const Particles<float> &, const ScatterChain<float> &,
                                        hipStream_t);
template int launch_update_pair<double>(const Particles<double> &, const Particles<double> &,
                                        int, const Particles<double> &, int, double,
                                        hipStream_t, const KickPending &);
template int launch_update_pair<float>(const Particles<float> &, const Particles<float> &, int,
                                       const Particles<float> &, int, float, hipStream_t,
                                       const KickPending &);

}  // namespace hhs

using namespace hhs;

extern "C" {

int hhs_drift_f64(double *pos, const double *vel, int n, double dt, double fac, void *stream) {
    return launch_drift<double>(pos, vel, n, dt * fac, as_stream(stream));
}

int hhs_drift_f32(float *pos, const float *vel, int n, float dt, float fac, void *stream) {
    return launch_drift<float>(pos, vel, n, dt * fac, as_stream(stream));
}

int hhs_update_f64(hhs_particles_f64 *prev, const hhs_particles_f64 *cur, int n, void *stream) {
    HHS_REQUIRE(prev && cur, "update: null system");
    HHS_REQUIRE(n >= 0 && n <= cur->n, "update: n=%d outside [0, %d]", n, cur->n);
    return launch_update<double>(from_c(*prev), from_c(*cur), n, as_stream(stream));
}

int hhs_update_f32(hhs_particles_f32 *prev, const hhs_particles_f32 *cur, int n, void *stream) {
    HHS_REQUIRE(prev && cur, "update: null system");
    HHS_REQUIRE(n >= 0 && n <= cur->n, "update: n=%d outside [0, %d]", n, cur->n);
    return launch_update<float>(from_c(*prev), from_c(*cur), n, as_stream(stream));
}

int hhs_energy_f64(const double *pos, const double *vel, const double *m, int n, double eps2,
                   double *kinetic, double *potential, void *stream) {
    HHS_REQUIRE(kinetic && potential, "energy: null output");
    HHS_REQUIRE(n >= 0, "energy: negative n");
    *kinetic = 0.0;
    *potential = 0.0;
    if (n == 0) return HHS_OK;
    HHS_REQUIRE(pos && vel && m, "energy: null array");
    hipStream_t st = as_stream(stream);
    double *buf = nullptr;
    // stream-ordered allocation: no device-wide synchronisation while other
    // streams are busy
    HHS_HIP(hipMallocAsync(reinterpret_cast<void **>(&buf),
                           (2 * static_cast<size_t>(n) + 2) * sizeof(double), st));
    double *row = buf, *kin = buf + n, *out = buf + 2 * static_cast<size_t>(n);
    int rc = HHS_OK;
    hipLaunchKernelGGL(energy_rows_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pos, vel, m, n,
                       eps2, row, kin);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) {
        hipLaunchKernelGGL(energy_reduce_kernel, dim3(1), dim3(1024), 0, st, row, kin, n, out);
        e = hipGetLastError();
    }
    double host[2] = {0.0, 0.0};
    if (e == hipSuccess) e = hipMemcpyAsync(host, out, sizeof(host), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = hip_status(e, "energy", __FILE__, __LINE__);
    (void)hipFreeAsync(buf, st);
    if (rc) return rc;
    *kinetic = host[0];
    *potential = host[1];
    // a NaN or Inf anywhere in the state (positions through the row marks,
    // velocities and masses through the sums) or an overflow reaches one of
    // the two sums
    if (!std::isfinite(host[0]) || !std::isfinite(host[1])) {
        set_error("energy: non-finite result (kinetic %g, potential %g) for n=%d", host[0],
                  host[1], n);
        return HHS_ERR_NONFINITE;
    }
    return HHS_OK;
}

}  // extern "C"
