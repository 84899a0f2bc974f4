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

namespace hhs {
namespace {

template <typename T> struct TsConst;
template <> struct TsConst<double> {
    static constexpr double sqrt2 = 1.41421356237309504880;
    static constexpr double huge = 1.e30;
};
template <> struct TsConst<float> {
    static constexpr float sqrt2 = 1.41421356237309504880f;
    static constexpr float huge = 1.e30f;
};

template <typename T>
__global__ __launch_bounds__(256) void timesteps_kernel(const T *__restrict__ pos,
                                                        const T *__restrict__ vel,
                                                        const T *__restrict__ m,
                                                        T *__restrict__ ts, int n, T eta,
                                                        T eps2) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ic = i < n ? i : n - 1;
    const T xi = pos[3 * ic], yi = pos[3 * ic + 1], zi = pos[3 * ic + 2];
    const T vxi = vel[3 * ic], vyi = vel[3 * ic + 1], vzi = vel[3 * ic + 2];
    const T mi = m[ic];
    const T c_ff = eta / TsConst<T>::sqrt2;
    T minval = TsConst<T>::huge;
    for (int j = 0; j < n; ++j) {
        const T drx = xi - pos[3 * j], dry = yi - pos[3 * j + 1], drz = zi - pos[3 * j + 2];
        const T dvx = vxi - vel[3 * j], dvy = vyi - vel[3 * j + 1], dvz = vzi - vel[3 * j + 2];
        const T mj = m[j];
        const T dr2 = drx * drx + dry * dry + drz * drz + eps2;
        const T dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
        const T r = sqrt(dr2);
        const T vdotdr2 = (dvx * drx + dvy * dry + dvz * drz) / dr2;
        // free-fall (src/timestep.cuh:36-42)
        T tau = c_ff * r * sqrt(r / (mj + mi));
        T dtau = T(3) * tau * vdotdr2 / T(2);
        dtau = (dtau < T(1)) ? dtau : T(1);
        tau /= (T(1) - dtau / T(2));
        const bool other = (j != i);
        if (other && tau < minval) minval = tau;
        // fly-by (src/timestep.cuh:46-57)
        tau = eta * r / sqrt(dv2);
        dtau = tau * vdotdr2 * (T(1) + (mi + mj) / (dv2 * r));
        dtau = (dtau < T(1)) ? dtau : T(1);
        tau /= (T(1) - dtau / T(2));
        if (other && tau < minval) minval = tau;
    }
    if (i < n) ts[i] = minval;
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

int hhs_timesteps_f32(const hhs_particles_f32 *sys, int n, float eta, float eps2,
                      void *stream) {
    HHS_REQUIRE(sys, "timesteps: null system");
    HHS_REQUIRE(n >= 0 && n <= sys->n, "timesteps: n=%d outside [0, %d]", n, sys->n);
    return launch_timesteps<float>(sys->pos, sys->vel, sys->m, sys->timestep, n, eta, eps2,
                                   as_stream(stream));
}

}  // extern "C"
