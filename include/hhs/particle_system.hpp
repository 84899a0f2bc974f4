// Drop-in replacement for the reference's data model headers
// (src/vec.cuh, src/utils.cuh, src/particle_system.cuh), implemented over
// the C ABI in hhs.h.  Plain C++17: no HIP headers are needed to use it;
// link against libhhs.so.
//
// Same type names, fields and methods as the reference.  Differences:
//  * errors go through hhs_check(): by default it prints the message and
//    calls exit(code) like gpuErrchk (src/utils.cuh:4-12); define
//    HHS_THROW_ON_ERROR to get std::runtime_error instead;
//  * nbody_params and partition live in host memory the device can also
//    write (the reference uses cudaMallocManaged, src/particle_system.cuh:81,
//    :123);
//  * cpu_free also frees parent_id (the reference leaks it, :203-209).
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "hhs.h"

inline void hhs_check(int code, const char *file, int line) {
    if (code == HHS_OK) return;
#ifdef HHS_THROW_ON_ERROR
    throw std::runtime_error(std::string("hhs error ") + std::to_string(code) + ": " +
                             hhs_last_error() + " (" + file + ":" + std::to_string(line) + ")");
#else
    std::fprintf(stderr, "GPUassert: %s %s %d\n", hhs_last_error(), file, line);
    std::exit(code);
#endif
}

// Same spelling as src/utils.cuh:4; wraps an hhs_* status instead of a
// cudaError_t.
#define gpuErrchk(ans) hhs_check((ans), __FILE__, __LINE__)

// src/vec.cuh:5-53 (host side; same layout: three T, no padding).
template <typename T> struct vec3 {
    T x, y, z;
    vec3() : x(0.0), y(0.0), z(0.0) {}
    vec3(const T &xi, const T &yi, const T &zi) : x(xi), y(yi), z(zi) {}
    vec3 operator+(vec3 const &v) const { return vec3(v.x + x, v.y + y, v.z + z); }
    vec3 operator-(vec3 const &v) const { return vec3(x - v.x, y - v.y, z - v.z); }
    T operator*(vec3 const &v) const { return v.x * x + v.y * y + v.z * z; }
    vec3 operator*(T const &v) const { return vec3(x * v, y * v, z * v); }
    vec3 operator-() const { return vec3(-x, -y, -z); }
    T norm2() const { return x * x + y * y + z * z; }
    T norm() const { return std::sqrt(norm2()); }
};

// src/particle_system.cuh:8-11
struct partition {
    int N_slow;
    int N_fast;
};

// src/particle_system.cuh:13-17
template <typename T> struct nbody_params {
    T eta;
    T eps;
};

namespace hhs_detail {
template <typename T> struct c_particles;
template <> struct c_particles<double> {
    using type = hhs_particles_f64;
};
template <> struct c_particles<float> {
    using type = hhs_particles_f32;
};
}  // namespace hhs_detail

// src/particle_system.cuh:20-54
template <typename T> class particle_system {
  public:
    vec3<T> *pos;
    vec3<T> *vel;
    vec3<T> *acc;
    T *timestep;
    T *m;
    int N;
    int *id;
    int *parent_id;
    char store;  // 'g' (device) or 'c' (host)
    particle_system<T> *d_ptr;

    particle_system(int N_init, char s) : N(N_init), store(s) {
        pos = vel = acc = nullptr;
        timestep = m = nullptr;
        id = parent_id = nullptr;
        d_ptr = nullptr;
    }

    void gpu_alloc();
    void cpu_alloc();
    void gpu_free();
    void cpu_free();
    void host_to_gpu(particle_system *host_system);
    void gpu_to_host(particle_system *gpu_system);

    // View of this system for the C ABI.
    typename hhs_detail::c_particles<T>::type c_view() const {
        typename hhs_detail::c_particles<T>::type v;
        v.pos = reinterpret_cast<T *>(pos);
        v.vel = reinterpret_cast<T *>(vel);
        v.acc = reinterpret_cast<T *>(acc);
        v.timestep = timestep;
        v.m = m;
        v.id = id;
        v.parent_id = parent_id;
        v.n = N;
        return v;
    }

  private:
    void require_store(char want, const char *what) const {
        if (store != want) {
            std::cerr << "Particle system store type is " << (store == 'g' ? "GPU" : "CPU")
                      << " but trying to " << what << "! Exiting" << std::endl;
            std::exit(1);  // src/particle_system.cuh:138-141 and friends
        }
    }
};

template <typename T> inline void particle_system<T>::gpu_alloc() {
    require_store('g', "allocate on the GPU");
    const size_t nv = static_cast<size_t>(N) * sizeof(vec3<T>);
    const size_t ns = static_cast<size_t>(N) * sizeof(T);
    gpuErrchk(hhs_malloc(reinterpret_cast<void **>(&pos), nv));
    gpuErrchk(hhs_malloc(reinterpret_cast<void **>(&vel), nv));
    gpuErrchk(hhs_malloc(reinterpret_cast<void **>(&acc), nv));
    gpuErrchk(hhs_malloc(reinterpret_cast<void **>(&timestep), ns));
    gpuErrchk(hhs_malloc(reinterpret_cast<void **>(&m), ns));
    gpuErrchk(hhs_malloc(reinterpret_cast<void **>(&id), static_cast<size_t>(N) * sizeof(int)));
    gpuErrchk(hhs_malloc(reinterpret_cast<void **>(&parent_id),
                         static_cast<size_t>(N) * sizeof(int)));
    // device mirror of the struct, as src/particle_system.cuh:150-157
    gpuErrchk(hhs_malloc(reinterpret_cast<void **>(&d_ptr), sizeof(particle_system<T>)));
    gpuErrchk(hhs_memcpy_h2d(d_ptr, this, sizeof(particle_system<T>)));
}

template <typename T> inline void particle_system<T>::gpu_free() {
    require_store('g', "free on the GPU");
    gpuErrchk(hhs_free(pos));
    gpuErrchk(hhs_free(vel));
    gpuErrchk(hhs_free(acc));
    gpuErrchk(hhs_free(timestep));
    gpuErrchk(hhs_free(m));
    gpuErrchk(hhs_free(id));
    gpuErrchk(hhs_free(parent_id));
    gpuErrchk(hhs_free(d_ptr));
    pos = vel = acc = nullptr;
    timestep = m = nullptr;
    id = parent_id = nullptr;
    d_ptr = nullptr;
}

template <typename T> inline void particle_system<T>::cpu_alloc() {
    require_store('c', "allocate on the CPU");
    pos = new vec3<T>[N];
    vel = new vec3<T>[N];
    acc = new vec3<T>[N];
    timestep = new T[N];
    m = new T[N];
    id = new int[N];
    parent_id = new int[N];
}

template <typename T> inline void particle_system<T>::cpu_free() {
    require_store('c', "free on the CPU");
    delete[] pos;
    delete[] vel;
    delete[] acc;
    delete[] timestep;
    delete[] m;
    delete[] id;
    delete[] parent_id;
    pos = vel = acc = nullptr;
    timestep = m = nullptr;
    id = parent_id = nullptr;
}

// src/particle_system.cuh:212-228 (copies pos, vel, acc, m, timestep, id)
template <typename T> inline void particle_system<T>::host_to_gpu(particle_system<T> *h) {
    require_store('g', "transfer from the CPU");
    const size_t nv = static_cast<size_t>(N) * sizeof(vec3<T>);
    const size_t ns = static_cast<size_t>(N) * sizeof(T);
    gpuErrchk(hhs_memcpy_h2d(pos, h->pos, nv));
    gpuErrchk(hhs_memcpy_h2d(vel, h->vel, nv));
    gpuErrchk(hhs_memcpy_h2d(acc, h->acc, nv));
    gpuErrchk(hhs_memcpy_h2d(m, h->m, ns));
    gpuErrchk(hhs_memcpy_h2d(timestep, h->timestep, ns));
    gpuErrchk(hhs_memcpy_h2d(id, h->id, static_cast<size_t>(N) * sizeof(int)));
}

// src/particle_system.cuh:230-247
template <typename T> inline void particle_system<T>::gpu_to_host(particle_system<T> *g) {
    require_store('c', "transfer from the GPU");
    const size_t nv = static_cast<size_t>(N) * sizeof(vec3<T>);
    const size_t ns = static_cast<size_t>(N) * sizeof(T);
    gpuErrchk(hhs_memcpy_d2h(pos, g->pos, nv));
    gpuErrchk(hhs_memcpy_d2h(vel, g->vel, nv));
    gpuErrchk(hhs_memcpy_d2h(acc, g->acc, nv));
    gpuErrchk(hhs_memcpy_d2h(m, g->m, ns));
    gpuErrchk(hhs_memcpy_d2h(timestep, g->timestep, ns));
    gpuErrchk(hhs_memcpy_d2h(id, g->id, static_cast<size_t>(N) * sizeof(int)));
}

// src/particle_system.cuh:65-104.  Also owns the HOLD state of the C ABI
// (per-level buffers, scan workspace, counters).
template <typename T> class nbodysystem_globals {
  public:
    int *predicate;
    int *scanned_predicate;
    T simtime;
    int N_total;
    nbody_params<T> *params;
    hhs_globals *handle;

    nbodysystem_globals(int N_init, T eta, T eps, int max_level = 0) : N_total(N_init) {
        gpuErrchk(hhs_globals_create(&handle, N_init, static_cast<double>(eta),
                                     static_cast<double>(eps), max_level));
        predicate = hhs_globals_predicate(handle);
        scanned_predicate = hhs_globals_scanned_predicate(handle);
        params = new nbody_params<T>{eta, eps};
        simtime = static_cast<T>(0.0);
    }
    ~nbodysystem_globals() {
        hhs_globals_destroy(handle);
        delete params;
    }
    nbodysystem_globals(const nbodysystem_globals &) = delete;
    nbodysystem_globals &operator=(const nbodysystem_globals &) = delete;

    void flush_simtime() {
        simtime = static_cast<T>(0.0);
        hhs_globals_set_simtime(handle, 0.0);
    }
    void print_simtime() { std::cout << "simtime: " << simtime << std::endl; }
};

// src/particle_system.cuh:107-132.  The reference's destructor frees `part`
// and the arrays but leaks the two particle_system objects; this one frees
// them all.
//
// The reference declares hold<T>(clevel, sys, buffers, ...) with the
// buffers as its only state (src/hold.cuh:20-23), so the buffers here also
// carry the integrator parameters (`params`: eta, eps; by default 0.01, the
// value of every benchmark configuration, and 0, the reference kick's
// softening) and, from the first hold() call on, the HOLD state of the C ABI
// (`globals`, whose simtime tracks the integration).
template <typename T> class nbodysystem_buffers {
  public:
    particle_system<T> *slow;
    particle_system<T> *fast;
    int N_total;
    partition *part;
    nbody_params<T> *params;
    nbodysystem_globals<T> *globals = nullptr;

    explicit nbodysystem_buffers(int N_init, T eta = T(0.01), T eps = T(0)) : N_total(N_init) {
        slow = new particle_system<T>(N_init, 'g');
        fast = new particle_system<T>(N_init, 'g');
        slow->gpu_alloc();
        fast->gpu_alloc();
        gpuErrchk(hhs_host_alloc_mapped(reinterpret_cast<void **>(&part), sizeof(partition)));
        part->N_slow = part->N_fast = 0;
        params = new nbody_params<T>{eta, eps};
    }
    ~nbodysystem_buffers() {
        delete globals;
        delete params;
        hhs_host_free_mapped(part);
        slow->gpu_free();
        fast->gpu_free();
        delete slow;
        delete fast;
    }
    // The HOLD state for hold(): created with params on first use.
    nbodysystem_globals<T> *hold_globals() {
        if (!globals) globals = new nbodysystem_globals<T>(N_total, params->eta, params->eps);
        return globals;
    }
    nbodysystem_buffers(const nbodysystem_buffers &) = delete;
    nbodysystem_buffers &operator=(const nbodysystem_buffers &) = delete;
};
