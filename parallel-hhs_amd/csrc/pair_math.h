// Device-side pair formulas shared by the per-system kernels and the
// batched (ensemble) kernel, so both paths produce identical numbers.
#pragma once

#include <hip/hip_runtime.h>

namespace hhs {

template <typename T> struct TsConst;
template <> struct TsConst<double> {
    static constexpr double sqrt2 = 1.41421356237309504880;
    static constexpr double huge = 1.e30;
};
template <> struct TsConst<float> {
    static constexpr float sqrt2 = 1.41421356237309504880f;
    static constexpr float huge = 1.e30f;
};

// One (i, j) pair of the timestep criterion, folded into minval.
// Reference: _get_pp_timestep, src/timestep.cuh:17-60, in its operation
// order (dr = x_i - x_j, dot as dv.x*dr.x + dv.y*dr.y + dv.z*dr.z,
// src/vec.cuh:31,44), with r^2 -> r^2 + eps2.  `other` is j != i.  IEEE
// sqrt/divide and no contraction: bit-exact with an FP64 CPU evaluation.
template <typename T>
__device__ __forceinline__ void pair_timestep(T xi, T yi, T zi, T vxi, T vyi, T vzi, T mi, T xj,
                                              T yj, T zj, T vxj, T vyj, T vzj, T mj, T eta,
                                              T c_ff, T eps2, bool other, T &minval) {
#pragma clang fp contract(off)
    const T drx = xi - xj, dry = yi - yj, drz = zi - zj;
    const T dvx = vxi - vxj, dvy = vyi - vyj, dvz = vzi - vzj;
    const T dr2 = drx * drx + dry * dry + drz * drz + eps2;
    const T dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
    const T r = sqrt(dr2);
    const T vdotdr2 = (dvx * drx + dvy * dry + dvz * drz) / dr2;
    // free-fall (src/timestep.cuh:36-42); c_ff = eta / sqrt(2)
    T tau = c_ff * r * sqrt(r / (mj + mi));
    T dtau = T(3) * tau * vdotdr2 / T(2);
    dtau = (dtau < T(1)) ? dtau : T(1);
    tau /= (T(1) - dtau / T(2));
    if (other && tau < minval) minval = tau;
    // fly-by (src/timestep.cuh:46-57)
    tau = eta * r / sqrt(dv2);
    dtau = tau * vdotdr2 * (T(1) + (mi + mj) / (dv2 * r));
    dtau = (dtau < T(1)) ? dtau : T(1);
    tau /= (T(1) - dtau / T(2));
    if (other && tau < minval) minval = tau;
}

// 1/sqrt(x) in FP64: the hardware estimate (v_rsq_f64) refined by two
// Newton-Raphson steps y <- y (3 - x y^2) / 2, which takes the estimate to
// full double precision.  x = 0 gives +inf and x = inf gives 0, as 1/sqrt.
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    if (x > 0.0 && x < __builtin_inf()) {
        y = y * __builtin_fma(-hx * y, y, 1.5);
        y = y * __builtin_fma(-hx * y, y, 1.5);
    }
    return y;
}

// Fast form of pair_timestep (same criterion, same NaN/inf behaviour): the
// square roots come from rsqrt_nr and the two divisions by (1 - dtau/2) are
// never performed -- each candidate is kept as a fraction num/den with
// den >= 1/2 and the running minimum is compared by cross-multiplication;
// the caller divides once per sink at the end (bnum / bden).  Agrees with the
// IEEE form to a few ulp.
__device__ __forceinline__ void fold_fraction(double num, double den, bool other, double &bnum,
                                              double &bden) {
    if (other && num * bden < bnum * den) {
        bnum = num;
        bden = den;
    }
}

__device__ __forceinline__ void pair_timestep_fast(double xi, double yi, double zi, double vxi,
                                                   double vyi, double vzi, double mi, double xj,
                                                   double yj, double zj, double vxj, double vyj,
                                                   double vzj, double mj, double eta, double c_ff,
                                                   double eps2, bool other, double &bnum,
                                                   double &bden) {
    const double drx = xi - xj, dry = yi - yj, drz = zi - zj;
    const double dvx = vxi - vxj, dvy = vyi - vyj, dvz = vzi - vzj;
    const double dr2 = __builtin_fma(drx, drx, __builtin_fma(dry, dry, __builtin_fma(drz, drz, eps2)));
    const double dv2 = __builtin_fma(dvx, dvx, __builtin_fma(dvy, dvy, dvz * dvz));
    const double rv = __builtin_fma(dvx, drx, __builtin_fma(dvy, dry, dvz * drz));
    const double mu = mj + mi;
    const double rinv = rsqrt_nr(dr2);
    const double r = dr2 > 0.0 ? dr2 * rinv : 0.0;  // sqrt(dr2), exact 0 at 0
    const double u = rv * (rinv * rinv);               // (dr.dv) / r^2
    // free-fall: c_ff * r * sqrt(r / mu) = c_ff / sqrt(mu / r^3)
    const double tff = c_ff * rsqrt_nr(mu * rinv * (rinv * rinv));
    double d = 1.5 * tff * u;
    d = (d < 1.0) ? d : 1.0;
    fold_fraction(tff, __builtin_fma(-0.5, d, 1.0), other, bnum, bden);
    // fly-by: eta * r / |dv|
    const double vinv = rsqrt_nr(dv2);
    const double tfb = eta * r * vinv;
    d = tfb * u * __builtin_fma(mu * (vinv * vinv), rinv, 1.0);
    d = (d < 1.0) ? d : 1.0;
    fold_fraction(tfb, __builtin_fma(-0.5, d, 1.0), other, bnum, bden);
}

// FP32 form of the same criterion: x_i - x_j and v_i - v_j are formed in
// FP64 and rounded once to FP32, everything after runs in FP32 with
// v_rsq_f32, and the running minimum is a fraction as in the fast form.
// Relative error ~1e-7 (inside the 1e-6 timestep bar).
__device__ __forceinline__ void fold_fraction32(float num, float den, bool other, float &bnum,
                                                float &bden) {
    if (other && num * bden < bnum * den) {
        bnum = num;
        bden = den;
    }
}

__device__ __forceinline__ void pair_timestep_f32(double xi, double yi, double zi, double vxi,
                                                  double vyi, double vzi, float mi, double xj,
                                                  double yj, double zj, double vxj, double vyj,
                                                  double vzj, float mj, float eta, float c_ff,
                                                  float eps2, bool other, float &bnum,
                                                  float &bden) {
    const float drx = static_cast<float>(xi - xj), dry = static_cast<float>(yi - yj),
                drz = static_cast<float>(zi - zj);
    const float dvx = static_cast<float>(vxi - vxj), dvy = static_cast<float>(vyi - vyj),
                dvz = static_cast<float>(vzi - vzj);
    const float dr2 = __builtin_fmaf(drx, drx, __builtin_fmaf(dry, dry, __builtin_fmaf(drz, drz, eps2)));
    const float dv2 = __builtin_fmaf(dvx, dvx, __builtin_fmaf(dvy, dvy, dvz * dvz));
    const float rv = __builtin_fmaf(dvx, drx, __builtin_fmaf(dvy, dry, dvz * drz));
    const float mu = mj + mi;
    const float rinv = __builtin_amdgcn_rsqf(dr2);
    const float r = dr2 > 0.0f ? dr2 * rinv : 0.0f;
    const float rinv2 = rinv * rinv;
    const float u = rv * rinv2;
    const float tff = c_ff * __builtin_amdgcn_rsqf(mu * rinv * rinv2);
    float d = 1.5f * tff * u;
    d = (d < 1.0f) ? d : 1.0f;
    fold_fraction32(tff, __builtin_fmaf(-0.5f, d, 1.0f), other, bnum, bden);
    const float vinv = __builtin_amdgcn_rsqf(dv2);
    const float tfb = eta * r * vinv;
    d = tfb * u * __builtin_fmaf(mu * (vinv * vinv), rinv, 1.0f);
    d = (d < 1.0f) ? d : 1.0f;
    fold_fraction32(tfb, __builtin_fmaf(-0.5f, d, 1.0f), other, bnum, bden);
}

// The pair's own timestep min(tau_ff, tau_fb) (start 1e30, NaN candidates
// ignored), from the same arithmetic as pair_timestep.  Symmetric in (i, j)
// bit for bit: the formula sees the pair only through r^2, v^2, v.r (signs
// of both differences flip together, exactly) and m_i + m_j (commutative).
template <typename T>
__device__ __forceinline__ T pair_tau(T xi, T yi, T zi, T vxi, T vyi, T vzi, T mi, T xj, T yj,
                                      T zj, T vxj, T vyj, T vzj, T mj, T eta, T c_ff, T eps2) {
    T best = TsConst<T>::huge;
    pair_timestep<T>(xi, yi, zi, vxi, vyi, vzi, mi, xj, yj, zj, vxj, vyj, vzj, mj, eta, c_ff, eps2,
                     true, best);
    return best;
}

// Double-single (hi, lo) FP32 record of one particle for the symmetric FP32
// kernel: x, y, z, vx, vy, vz each as hi = (float)v, lo = (float)(v - hi),
// then m (FP32) -- 13 floats, padded to 16 (64 bytes).
constexpr int kDsRec = 16;

// x_i - x_j from double-single parts: (hi_i - hi_j) + (lo_i - lo_j), all in
// FP32.  Relative error a few 2^-24 of the difference (the hi difference is
// exact for nearby values, by Sterbenz), against half an ulp for the FP64
// difference rounded once -- 3 FP32 operations instead of an FP64 subtract
// and a conversion, which on MI355X take ~1.7 FP32 slots each.
__device__ __forceinline__ float ds_sep(float hi, float lo, float hj, float lj) {
    return (hi - hj) + (lo - lj);
}

// FP32 form of the pair's two candidates tau_ff and tau_fb on double-single
// records (a: the sink in registers, b: the source), for the symmetric
// kernel: the arithmetic of pair_timestep_f32 from the separations on, each
// candidate divided by (1 - dtau/2) through the hardware reciprocal (~1 ulp)
// instead of being folded as a fraction.  Either candidate may be NaN; the
// caller folds them into running minima that start at 1e30 with fminf
// (fold_tau_ds), which ignores NaN.  Every candidate is >= +0 or NaN (c_ff,
// eta > 0; the denominators lie in [1/2, inf]).  Symmetric in (i, j) bit
// for bit: a swap flips the sign of every separation exactly (the operations
// on hi and lo parts are negated operand for operand), and the formula sees
// them only through r^2, v^2, v.r and m_i + m_j.  SOFT: eps2 > 0 in FP32, so
// r^2 > 0 and the r = 0 case needs no test.
template <bool SOFT>
__device__ __forceinline__ void pair_tau_ds(const float (&a)[13], const float *b, float eta,
                                            float c_ff, float eps2, float &t_ff, float &t_fb) {
    const float drx = ds_sep(a[0], a[1], b[0], b[1]), dry = ds_sep(a[2], a[3], b[2], b[3]),
                drz = ds_sep(a[4], a[5], b[4], b[5]);
    const float dvx = ds_sep(a[6], a[7], b[6], b[7]), dvy = ds_sep(a[8], a[9], b[8], b[9]),
                dvz = ds_sep(a[10], a[11], b[10], b[11]);
    const float dr2 = __builtin_fmaf(drx, drx, __builtin_fmaf(dry, dry, __builtin_fmaf(drz, drz, eps2)));
    const float dv2 = __builtin_fmaf(dvx, dvx, __builtin_fmaf(dvy, dvy, dvz * dvz));
    const float rv = __builtin_fmaf(dvx, drx, __builtin_fmaf(dvy, dry, dvz * drz));
    const float mu = b[12] + a[12];
    const float rinv = __builtin_amdgcn_rsqf(dr2);
    const float r = (SOFT || dr2 > 0.0f) ? dr2 * rinv : 0.0f;
    const float rinv2 = rinv * rinv;
    const float u = rv * rinv2;
    // fminf(NaN, 1) = 1, as the (d < 1) ? d : 1 cap
    const float tff = c_ff * __builtin_amdgcn_rsqf(mu * rinv * rinv2);
    float d = __builtin_fminf(1.5f * tff * u, 1.0f);
    t_ff = tff * __builtin_amdgcn_rcpf(__builtin_fmaf(-0.5f, d, 1.0f));
    const float vinv = __builtin_amdgcn_rsqf(dv2);
    const float tfb = eta * r * vinv;
    d = __builtin_fminf(tfb * u * __builtin_fmaf(mu * (vinv * vinv), rinv, 1.0f), 1.0f);
    t_fb = tfb * __builtin_amdgcn_rcpf(__builtin_fmaf(-0.5f, d, 1.0f));
}

// acc <- min(acc, t_ff, t_fb), NaN candidates ignored (acc is never NaN: it
// starts at 1e30).  One v_min3_f32.  Gives the bits of folding the pair
// minimum clamped to 1e30, min(acc, min(min(t_ff, t_fb), 1e30)), since
// acc <= 1e30.
__device__ __forceinline__ float fold_tau_ds(float acc, float t_ff, float t_fb) {
    return __builtin_fminf(__builtin_fminf(acc, t_ff), t_fb);
}

// ------------------------------------------------------ decision guard
// HOLD compares each timestep with the pivots dt * 2^-k (k >= 0) of the call
// that computed it and of every level below (src/split.cuh:115,
// src/hold.cuh:79-83,116-121).  An approximate timestep (FP32 or FAST mode)
// lands on the other side of a pivot from the exact one only when the two
// straddle it, i.e. when the approximate value lies within its error bound
// of the pivot.  The guard recomputes exactly such timesteps -- those
// within `band` (relative) of a pivot -- with the exact formula over the
// same sources, so the slow/fast decisions of the approximate modes are
// those of the exact mode.  band: FP32 1e-5 (the FP32 criterion is within
// 1e-6 relative), FAST 1e-12 (within 1e-14).  About 2 band / ln 2 of the
// timesteps (3e-5 in FP32) are recomputed.
constexpr double kGuardBandFp32 = 1e-5;
constexpr double kGuardBandFast = 1e-12;

struct TsGuard {
    double inv_dt;               // 1 / the call's dt (the largest pivot)
    double qmax;                 // 1 + band
    unsigned long long lo, hi;   // band in units of the FP64 mantissa
};

inline TsGuard make_ts_guard(double dt, double band) {
    TsGuard g;
    g.inv_dt = 1.0 / dt;
    g.qmax = 1.0 + band;
    g.lo = static_cast<unsigned long long>(band * 4503599627370496.0);            // band * 2^52
    g.hi = static_cast<unsigned long long>((1.0 - 2.0 * band) * 4503599627370496.0);
    return g;
}

// True when q = ts / dt lies within the band of a power of two 2^-k, k >= 0:
// q = 2^e (1 + f) is within band of 2^e when f <= band and within band of
// 2^(e+1) when (1 - f) / 2 <= band.  NaN, zero and values above every pivot
// are never flagged.  q is formed as ts * (1 / dt): exact for the
// power-of-two dt of HOLD's levels below a power-of-two root step, and
// within an ulp (far inside the band) otherwise.
__device__ __forceinline__ bool guard_near_pivot(const TsGuard &g, double ts) {
    const double q = ts * g.inv_dt;
    if (!(q > 0.0) || q > g.qmax) return false;
    const unsigned long long f =
        static_cast<unsigned long long>(__double_as_longlong(q)) & 0xFFFFFFFFFFFFFull;
    return f <= g.lo || f >= g.hi;
}

// Knuth two-sum of hi + p; the rounding error is carried into lo.
__device__ __forceinline__ void two_sum_add(float &hi, float &lo, float p) {
    const float s = hi + p;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (p - bp);
    hi = s;
    lo += err;
}

__device__ __forceinline__ float separation(double xj, double xi) {
    return static_cast<float>(xj - xi);
}
__device__ __forceinline__ float separation(float xj, float xi) { return xj - xi; }

// One source's contribution in FP64, in the CPU direct sum's operation order
// (src/kick.cuh:19-30 with the attractive sign of tests/test_nbforce.cu:14-18):
// dx = x_j - x_i, d2 = dx*dx + dy*dy + dz*dz + eps2, skip d2 == 0,
// s = m_j * (1 / (d2 * sqrt(d2))), a += dx * s.  IEEE sqrt and divide, no
// contraction: summed in source order it is bit-exact with that sum.
__device__ __forceinline__ void pair_accel_exact(double xj, double yj, double zj, double mj,
                                                 double xi, double yi, double zi, double eps2,
                                                 double &ax, double &ay, double &az) {
#pragma clang fp contract(off)
    const double dx = xj - xi, dy = yj - yi, dz = zj - zi;
    const double d2 = dx * dx + dy * dy + dz * dz + eps2;
    const double s = mj * (1.0 / (d2 * sqrt(d2)));
    const bool skip = d2 == 0.0;
    ax = skip ? ax : ax + dx * s;
    ay = skip ? ay : ay + dy * s;
    az = skip ? az : az + dz * s;
}

// Mean bias of v_rsq_f32.  Its relative error has a nonzero mean: -7.371e-9
// over inputs spread log-uniformly over whole octaves (MI355X, 2^22 inputs
// each in [1, 4), [0.5, 2) and [1e-8, 1): -7.371e-9, -7.371e-9, -7.342e-9;
// rms 2.79e-8, max 9.4e-8).  The pair weight m * rsq^3 therefore runs
// 3 * 7.371e-9 = 2.211e-8 low on average, and a direct sum over many
// sources turns that into a systematic force-scale error (mean relative
// error along a of -2.2e-8 on the N = 16384 and 131072 spheres, most of the
// rms error of 3e-8) -- an energy drift that FP64 forces do not have.  The
// force paths that use v_rsq_f32 multiply their FP64 totals by kRsqDebias,
// which removes the mean and leaves the zero-mean part.
constexpr double kRsqDebias = 1.0 + 2.2113e-8;

// One source's contribution to a sink's acceleration (FP32 from FP64-derived
// separations); SOFT = (eps2 > 0), else pairs with r^2 == 0 are skipped
// (src/kick.cuh:25).
template <bool SOFT, typename T>
__device__ __forceinline__ void pair_accel(T xj, T yj, T zj, float mj, T xi, T yi, T zi,
                                           float eps2, float &px, float &py, float &pz) {
    const float dx = separation(xj, xi);
    const float dy = separation(yj, yi);
    const float dz = separation(zj, zi);
    const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, __builtin_fmaf(dz, dz, eps2)));
    const float rinv = __builtin_amdgcn_rsqf(r2);
    const float rinv2 = rinv * rinv;
    float w = (mj * rinv) * rinv2;
    if (!SOFT) w = (r2 > 0.0f) ? w : 0.0f;
    px = __builtin_fmaf(w, dx, px);
    py = __builtin_fmaf(w, dy, py);
    pz = __builtin_fmaf(w, dz, pz);
}

}  // namespace hhs
