// Drop-in replacement for the reference's kernel and integrator headers
// (src/hold.cuh, src/kick.cuh, src/timestep.cuh, src/split.cuh,
// src/drift.cuh, src/update_system.cuh), implemented over the C ABI.
//
// Template parameters BLOCKSIZE and UNROLL are accepted for source
// compatibility and ignored: launch shapes are chosen by the library.
// All calls go to the default stream, like the reference's launches.
//
// Semantics follow the reference with its defects fixed (attractive kick,
// race-free lone-fast-particle rule, no leaks, no per-call allocation in
// hold_step); the kick wrappers without an eps argument are unsoftened, as
// in the reference (src/kick.cuh:15-32).
#pragma once

#include <cstdint>
#include <type_traits>

#include "hhs.h"
#include "hhs/particle_system.hpp"

namespace hhs_detail {

// Grow-only scan workspace of the calling thread (the reference's scans take
// no workspace and allocate scratch on every call, src/split.cuh:513-514).
inline void *scan_ws(int n, size_t *bytes) {
    void *ws = nullptr;
    gpuErrchk(hhs_scan_workspace_thread(n, &ws, bytes));
    return ws;
}

}  // namespace hhs_detail

// ---------------------------------------------------------------- kicks
// kick_slow (src/kick.cuh:158-168)
template <typename T, int BLOCKSIZE>
void kick_slow(particle_system<T> *sys, T dt, T fac, const int N, T eps = T(0)) {
    auto v = sys->c_view();
    if constexpr (std::is_same<T, double>::value)
        gpuErrchk(hhs_kick_self_f64(&v, N, dt, fac, eps * eps, nullptr));
    else
        gpuErrchk(hhs_kick_self_f32(&v, N, dt, fac, eps * eps, nullptr));
}

// kick_sf (src/kick.cuh:170-181): N sinks of si_sys, M sources of so_sys.
template <typename T, int BLOCKSIZE>
void kick_sf(particle_system<T> *so_sys, particle_system<T> *si_sys, T dt, T fac, const int N,
             const int M, T eps = T(0)) {
    auto so = so_sys->c_view();
    auto si = si_sys->c_view();
    if constexpr (std::is_same<T, double>::value)
        gpuErrchk(hhs_kick_cross_f64(&so, &si, N, M, dt, fac, eps * eps, nullptr));
    else
        gpuErrchk(hhs_kick_cross_f32(&so, &si, N, M, dt, fac, eps * eps, nullptr));
}

// ------------------------------------------------------------ timesteps
// timesteps (src/timestep.cuh:103-118); eta and eps from globals->params.
template <typename T, int BLOCKSIZE>
void timesteps(particle_system<T> *sys, nbodysystem_globals<T> *globals, const int N) {
    auto v = sys->c_view();
    const T eps2 = globals->params->eps * globals->params->eps;
    if constexpr (std::is_same<T, double>::value)
        gpuErrchk(hhs_timesteps_f64(&v, N, globals->params->eta, eps2, nullptr));
    else
        gpuErrchk(hhs_timesteps_f32(&v, N, globals->params->eta, eps2, nullptr));
}

// ------------------------------------------------------------ scan/split
// scan_predicate_small / _large / _large_ev (src/split.cuh:437-565): all
// three are the same exclusive scan here (the library picks its form by n).
template <int UNROLL>
void scan_predicate_small(int *predicate, int *scanned_predicate, const int N) {
    size_t bytes = 0;
    void *ws = hhs_detail::scan_ws(N, &bytes);
    gpuErrchk(hhs_scan_exclusive_i32(predicate, scanned_predicate, N, ws, bytes, nullptr));
}
template <int UNROLL>
void scan_predicate_large(int *predicate, int *scanned_predicate, const int N, int) {
    scan_predicate_small<UNROLL>(predicate, scanned_predicate, N);
}
template <int UNROLL>
void scan_predicate_large_ev(int *predicate, int *scanned_predicate, const int N, int) {
    scan_predicate_small<UNROLL>(predicate, scanned_predicate, N);
}

// split (src/split.cuh:573-610).  `part` must be memory the device can
// write (device, managed or nbodysystem_buffers::part); like the
// reference, the call returns after the device has finished, so a
// host-visible `part` can be read at once.
template <typename T, int BLOCKSIZE, int UNROLL>
void split(particle_system<T> *total, particle_system<T> *slow, particle_system<T> *fast,
           int *predicate, int *scanned_predicate, const T pivot, const int N,
           partition *part) {
    size_t bytes = 0;
    void *ws = hhs_detail::scan_ws(N, &bytes);
    auto t = total->c_view(), s = slow->c_view(), f = fast->c_view();
    auto *p = reinterpret_cast<hhs_partition *>(part);
    if constexpr (std::is_same<T, double>::value)
        gpuErrchk(hhs_split_f64(&t, &s, &f, predicate, scanned_predicate, pivot, N, p, ws, bytes,
                                nullptr));
    else
        gpuErrchk(hhs_split_f32(&t, &s, &f, predicate, scanned_predicate, pivot, N, p, ws, bytes,
                                nullptr));
    gpuErrchk(hhs_stream_synchronize(nullptr));  // src/split.cuh:430
}

// Scalar-array split of the scan prototypes' tests
// (src/scan_optimized_3.cuh:485-515), with split.cuh naming: values
// <= pivot are fast.  predicate/scanned_predicate are unused scratch.
template <typename T, int BLOCKSIZE, int UNROLL>
void split(T *total, T *slow, T *fast, int *, int *, const T pivot, const int N,
           partition *sizes) {
    size_t bytes = 0;
    void *ws = hhs_detail::scan_ws(N, &bytes);
    auto *p = reinterpret_cast<hhs_partition *>(sizes);
    if constexpr (std::is_same<T, double>::value)
        gpuErrchk(hhs_split_values_f64(total, slow, fast, pivot, N, p, ws, bytes, nullptr));
    else
        gpuErrchk(hhs_split_values_f32(total, slow, fast, pivot, N, p, ws, bytes, nullptr));
    gpuErrchk(hhs_stream_synchronize(nullptr));
}

// -------------------------------------------------------- drift / update
// drift (src/drift.cuh:30-43)
template <typename T, int BLOCKSIZE>
void drift(particle_system<T> *sys, const T dt, const T fac, const int N) {
    if constexpr (std::is_same<T, double>::value)
        gpuErrchk(hhs_drift_f64(reinterpret_cast<double *>(sys->pos),
                                reinterpret_cast<const double *>(sys->vel), N, dt, fac, nullptr));
    else
        gpuErrchk(hhs_drift_f32(reinterpret_cast<float *>(sys->pos),
                                reinterpret_cast<const float *>(sys->vel), N, dt, fac, nullptr));
}

// update_system (src/update_system.cuh:29-43)
template <typename T, int BLOCKSIZE>
void update_system(particle_system<T> *prev_level_sys, particle_system<T> *cur_level_sys, int N) {
    if (N == 0) return;
    auto p = prev_level_sys->c_view(), c = cur_level_sys->c_view();
    if constexpr (std::is_same<T, double>::value)
        gpuErrchk(hhs_update_f64(&p, &c, N, nullptr));
    else
        gpuErrchk(hhs_update_f32(&p, &c, N, nullptr));
}

// ------------------------------------------------------------------ HOLD
// hold_step (src/hold.cuh:51-181), FP64 and FP32 particle systems.
// Per-level buffers are owned by `globals` and reused; globals->simtime
// tracks the reference's host-side simtime (src/hold.cuh:106-114).
template <typename T, int BLOCKSIZE>
void hold_step(int clevel, nbodysystem_globals<T> *globals, particle_system<T> *total,
               const int N_total, T stime, T etime, T dt, bool calc_timestep) {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "hold_step: FP64 or FP32 particle systems");
    auto v = total->c_view();
    hhs_globals_set_simtime(globals->handle, static_cast<double>(globals->simtime));
    if constexpr (std::is_same<T, double>::value)
        gpuErrchk(hhs_hold_step_f64(globals->handle, &v, N_total, clevel, stime, etime, dt,
                                    calc_timestep ? 1 : 0, nullptr));
    else
        gpuErrchk(hhs_hold_step_f32(globals->handle, &v, N_total, clevel, stime, etime, dt,
                                    calc_timestep ? 1 : 0, nullptr));
    globals->simtime = static_cast<T>(hhs_globals_simtime(globals->handle));
}

// hold, with the signature the reference declares and never defines
// (src/hold.cuh:20-23): root steps of dt from stime while t < etime, each a
// hold_step at level clevel.  calc_timestep applies to the first root step
// (false: sys->timestep already holds the timesteps of the current state);
// every later root step recomputes them, since the positions have moved.
// eta and eps come from buffers->params; the HOLD state lives in
// buffers->globals (created on the first call).
template <typename T>
void hold(int clevel, particle_system<T> *sys, nbodysystem_buffers<T> *buffers,
          const int N_total, T stime, T etime, T dt, bool calc_timestep) {
    nbodysystem_globals<T> *g = buffers->hold_globals();
    bool calc = calc_timestep;
    for (T t = stime; t < etime; t += dt) {
        hold_step<T, 0>(clevel, g, sys, N_total, t, t + dt, dt, calc);
        calc = true;
    }
}

// hold over an explicit globals object (eta and eps from globals->params):
// the same root-step loop, timesteps computed at every root step.
template <typename T>
void hold(int clevel, nbodysystem_globals<T> *globals, particle_system<T> *sys,
          const int N_total, T stime, T etime, T dt) {
    for (T t = stime; t < etime; t += dt)
        hold_step<T, 0>(clevel, globals, sys, N_total, t, t + dt, dt, true);
}
