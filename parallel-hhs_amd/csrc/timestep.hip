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
#include "hhs_internal.h"
#include "pair_math.h"

namespace hhs {
namespace {

template <typename T>
__global__ __launch_bounds__(256) void timesteps_kernel(const T *__restrict__ pos,
                                                        const T *__restrict__ vel,
                                                        const T *__restrict__ m,
                                                        T *__restrict__ ts, int n, T eta,
                                                        T eps2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ic = i < n ? i : n - 1;
    const T xi = pos[3 * ic], yi = pos[3 * ic + 1], zi = pos[3 * ic + 2];
    const T vxi = vel[3 * ic], vyi = vel[3 * ic + 1], vzi = vel[3 * ic + 2];
    const T mi = m[ic];
    const T c_ff = eta / TsConst<T>::sqrt2;
    T minval = TsConst<T>::huge;
    for (int j = 0; j < n; ++j)
        pair_timestep<T>(xi, yi, zi, vxi, vyi, vzi, mi, pos[3 * j], pos[3 * j + 1], pos[3 * j + 2],
                         vel[3 * j], vel[3 * j + 1], vel[3 * j + 2], m[j], eta, c_ff, eps2, j != i,
                         minval);
    if (i < n) ts[i] = minval;
}

// Fast FP64 form (pair_timestep_fast): rsqrt + Newton instead of IEEE
// sqrt/divide, one division per sink.  Agrees with timesteps_kernel to a
// few ulp; not bit-exact.
__global__ __launch_bounds__(256) void timesteps_fast_kernel(const double *__restrict__ pos,
                                                             const double *__restrict__ vel,
                                                             const double *__restrict__ m,
                                                             double *__restrict__ ts, int n,
                                                             double eta, double eps2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ic = i < n ? i : n - 1;
    const double xi = pos[3 * ic], yi = pos[3 * ic + 1], zi = pos[3 * ic + 2];
    const double vxi = vel[3 * ic], vyi = vel[3 * ic + 1], vzi = vel[3 * ic + 2];
    const double mi = m[ic];
    const double c_ff = eta / TsConst<double>::sqrt2;
    double bnum = TsConst<double>::huge, bden = 1.0;
    for (int j = 0; j < n; ++j)
        pair_timestep_fast(xi, yi, zi, vxi, vyi, vzi, mi, pos[3 * j], pos[3 * j + 1],
                           pos[3 * j + 2], vel[3 * j], vel[3 * j + 1], vel[3 * j + 2], m[j], eta,
                           c_ff, eps2, j != i, bnum, bden);
    if (i < n) ts[i] = bnum / bden;
}

}  // namespace

template <typename T>
int launch_timesteps(const T *pos, const T *vel, const T *m, T *ts, int n, T eta, T eps2,
                     hipStream_t s) {
    HHS_REQUIRE(n >= 0, "timesteps: negative n");
    if (n == 0) return HHS_OK;
    HHS_REQUIRE(pos && vel && m && ts, "timesteps: null array");
    dim3 block(256), grid((n + 255) / 256);
    hipLaunchKernelGGL((timesteps_kernel<T>), grid, block, 0, s, pos, vel, m, ts, n, eta, eps2);
    HHS_LAUNCH_CHECK("timesteps_kernel");
    return HHS_OK;
}

int launch_timesteps_fast(const double *pos, const double *vel, const double *m, double *ts,
                          int n, double eta, double eps2, hipStream_t s) {
    HHS_REQUIRE(n >= 0, "timesteps: negative n");
    if (n == 0) return HHS_OK;
    HHS_REQUIRE(pos && vel && m && ts, "timesteps: null array");
    dim3 block(256), grid((n + 255) / 256);
    hipLaunchKernelGGL(timesteps_fast_kernel, grid, block, 0, s, pos, vel, m, ts, n, eta, eps2);
    HHS_LAUNCH_CHECK("timesteps_fast_kernel");
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
    return launch_timesteps_fast(sys->pos, sys->vel, sys->m, sys->timestep, n, eta, eps2,
                                 as_stream(stream));
}

int hhs_timesteps_f32(const hhs_particles_f32 *sys, int n, float eta, float eps2,
                      void *stream) {
    HHS_REQUIRE(sys, "timesteps: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "timesteps: n=%d outside [0, %d]", n, sys->n);
    return launch_timesteps<float>(sys->pos, sys->vel, sys->m, sys->timestep, n, eta, eps2,
                                   as_stream(stream));
}

}  // extern "C"
