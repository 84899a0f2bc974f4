// Drop-in API test: reference-style host code (particle_system<T>,
// split<T,BS,U>, kick_slow, timesteps, drift, update_system, hold_step)
// compiled with g++ against include/hhs/*.hpp and libhhs.so only.
//
// Mirrors the reference's tests -- the split size/block matrix of
// tests/test_scan_unrolled.cu:191-215 and the CPU-vs-GPU force comparison of
// tests/test_nbforce.cu:64-121 -- with real assertions and an exit code.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "hhs/hold.hpp"

static int g_fail = 0;
#define CHECK(cond, ...)                                                    \
    do {                                                                    \
        if (!(cond)) {                                                      \
            std::printf("[FAIL] " __VA_ARGS__);                             \
            std::printf("\n");                                              \
            ++g_fail;                                                       \
        }                                                                   \
    } while (0)

// tests/test_scan_unrolled.cu: values U[0,10), pivot 5; split.cuh naming
// (values <= pivot are fast).
template <int NSIZE, int BLOCKSIZE, int UNROLL> void test_split_values() {
    const int N = NSIZE;
    std::vector<float> total(N), slow(N), fast(N);
    for (int i = 0; i < N; i++) total[i] = 10.0f * float(std::rand()) / float(RAND_MAX);
    const float pivot = 5.0f;
    float *d_total, *d_slow, *d_fast;
    int *d_pred, *d_scan;
    partition *d_sizes;
    gpuErrchk(hhs_malloc((void **)&d_total, sizeof(float) * N));
    gpuErrchk(hhs_malloc((void **)&d_slow, sizeof(float) * N));
    gpuErrchk(hhs_malloc((void **)&d_fast, sizeof(float) * N));
    gpuErrchk(hhs_malloc((void **)&d_pred, sizeof(int) * N));
    gpuErrchk(hhs_malloc((void **)&d_scan, sizeof(int) * N));
    gpuErrchk(hhs_malloc((void **)&d_sizes, sizeof(partition)));
    gpuErrchk(hhs_memcpy_h2d(d_total, total.data(), sizeof(float) * N));
    split<float, BLOCKSIZE, UNROLL>(d_total, d_slow, d_fast, d_pred, d_scan, pivot, N, d_sizes);
    partition sizes;
    gpuErrchk(hhs_memcpy_d2h(&sizes, d_sizes, sizeof(partition)));
    gpuErrchk(hhs_memcpy_d2h(slow.data(), d_slow, sizeof(float) * N));
    gpuErrchk(hhs_memcpy_d2h(fast.data(), d_fast, sizeof(float) * N));
    CHECK(sizes.N_slow + sizes.N_fast == N, "split N=%d: sizes do not sum to N", N);
    int nf = 0;
    for (int i = 0; i < N; i++) {
        if (total[i] <= pivot) {
            CHECK(nf < sizes.N_fast && fast[nf] == total[i], "split N=%d: fast[%d]", N, nf);
            ++nf;
        } else {
            const int s = i - nf;
            CHECK(s < sizes.N_slow && slow[s] == total[i], "split N=%d: slow[%d]", N, s);
        }
    }
    CHECK(nf == sizes.N_fast, "split N=%d: N_fast %d vs %d", N, sizes.N_fast, nf);
    hhs_free(d_total);
    hhs_free(d_slow);
    hhs_free(d_fast);
    hhs_free(d_pred);
    hhs_free(d_scan);
    hhs_free(d_sizes);
    std::printf("[PASS?] split N=%d BS=%d UNROLL=%d (N_slow=%d N_fast=%d)\n", N, BLOCKSIZE, UNROLL,
                sizes.N_slow, sizes.N_fast);
}

// tests/test_nbforce.cu: unit masses at U[0,1)^3, GPU kick vs an FP64 CPU
// direct sum (attractive sign, tests/test_nbforce.cu:14-18).
void test_nbforce() {
    const int N = 1024;
    particle_system<double> h(N, 'c'), d(N, 'g');
    h.cpu_alloc();
    d.gpu_alloc();
    for (int i = 0; i < N; i++) {
        h.pos[i] = vec3<double>(double(std::rand()) / RAND_MAX, double(std::rand()) / RAND_MAX,
                                double(std::rand()) / RAND_MAX);
        h.vel[i] = vec3<double>(0, 0, 0);
        h.acc[i] = vec3<double>(0, 0, 0);
        h.m[i] = 1.0;
        h.timestep[i] = 0.0;
        h.id[i] = i;
    }
    d.host_to_gpu(&h);
    kick_slow<double, 256>(&d, 0.5, 0.5, N);
    particle_system<double> out(N, 'c');
    out.cpu_alloc();
    out.gpu_to_host(&d);
    double worst = 0.0;
    for (int i = 0; i < N; i++) {
        vec3<double> a(0, 0, 0);
        double scale = 0.0;
        for (int j = 0; j < N; j++) {
            if (i == j) continue;
            vec3<double> dx = h.pos[i] - h.pos[j];
            double d2 = dx.norm2();
            double r = std::sqrt(d2);
            a = a + dx * (-h.m[j] / (d2 * r));
            scale += h.m[j] / d2;
        }
        vec3<double> e = out.acc[i] - a;
        worst = std::max(worst, e.norm() / scale);
        vec3<double> dv = out.vel[i] - a * 0.25;
        CHECK(dv.norm() <= 1e-6 * 0.25 * scale, "nbforce: vel[%d] not acc*dt*fac", i);
    }
    CHECK(worst < 1e-6, "nbforce: normalised force error %.3e", worst);
    std::printf("[PASS?] nbforce N=%d max normalised error %.3e\n", N, worst);
    h.cpu_free();
    out.cpu_free();
    d.gpu_free();
}

// hold_step through the drop-in globals on a Plummer sphere: energy is
// conserved to the integrator's error and simtime advances by dt per root
// step.
void test_hold_step() {
    const int N = 256;
    std::vector<double> p(3 * N), v(3 * N), m(N);
    gpuErrchk(hhs_ic_plummer(N, 5, 0, 1, 1e-4, p.data(), v.data(), m.data()));
    particle_system<double> h(N, 'c'), d(N, 'g');
    h.cpu_alloc();
    d.gpu_alloc();
    for (int i = 0; i < N; i++) {
        h.pos[i] = vec3<double>(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
        h.vel[i] = vec3<double>(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
        h.acc[i] = vec3<double>();
        h.m[i] = m[i];
        h.timestep[i] = 0.0;
        h.id[i] = i;
    }
    d.host_to_gpu(&h);
    nbodysystem_globals<double> globals(N, 0.01, 0.01);
    double k0, u0, k1, u1;
    gpuErrchk(hhs_energy_f64((double *)d.pos, (double *)d.vel, d.m, N, 1e-4, &k0, &u0, nullptr));
    hold<double>(0, &globals, &d, N, 0.0, 0.125, 1.0 / 32);
    gpuErrchk(hhs_energy_f64((double *)d.pos, (double *)d.vel, d.m, N, 1e-4, &k1, &u1, nullptr));
    const double de = std::fabs((k1 + u1 - k0 - u0) / (k0 + u0));
    CHECK(de < 1e-6, "hold: |dE/E| = %.3e", de);
    CHECK(std::fabs(globals.simtime - 0.125) < 1e-15, "hold: simtime %.17g", globals.simtime);
    std::printf("[PASS?] hold N=%d t=0.125 |dE/E|=%.3e simtime=%.6f\n", N, de, globals.simtime);
    h.cpu_free();
    d.gpu_free();
}

// split on particle systems through nbodysystem_buffers (src/hold.cuh:79-83),
// then update_system scatters both halves back (src/hold.cuh:177-178).
void test_split_update_roundtrip() {
    const int N = 3000;
    particle_system<double> h(N, 'c'), d(N, 'g');
    h.cpu_alloc();
    d.gpu_alloc();
    for (int i = 0; i < N; i++) {
        h.pos[i] = vec3<double>(i, 2.0 * i, 3.0 * i);
        h.vel[i] = vec3<double>(-i, 0.5, 0.25);
        h.acc[i] = vec3<double>();
        h.m[i] = 1.0 + i;
        h.timestep[i] = double((i * 7919) % 1000) / 1000.0;
        h.id[i] = i;
    }
    d.host_to_gpu(&h);
    nbodysystem_globals<double> globals(N, 0.01, 0.0);
    nbodysystem_buffers<double> buf(N);
    split<double, 256, 4>(&d, buf.slow, buf.fast, globals.predicate, globals.scanned_predicate,
                          0.25, N, buf.part);
    int expect_fast = 0;
    for (int i = 0; i < N; i++) expect_fast += h.timestep[i] <= 0.25;
    CHECK(buf.part->N_fast == expect_fast && buf.part->N_slow == N - expect_fast,
          "split counts %d/%d", buf.part->N_slow, buf.part->N_fast);
    std::vector<int> scan(N);
    gpuErrchk(hhs_memcpy_d2h(scan.data(), globals.scanned_predicate, sizeof(int) * N));
    int run = 0;
    for (int i = 0; i < N; i++) {
        CHECK(scan[i] == run, "scanned_predicate[%d]", i);
        run += h.timestep[i] <= 0.25;
    }
    drift<double, 256>(buf.slow, 1.0, 0.5, buf.part->N_slow);
    particle_system<double> back(N, 'g');
    back.gpu_alloc();
    update_system<double, 256>(&back, buf.fast, buf.part->N_fast);
    update_system<double, 256>(&back, buf.slow, buf.part->N_slow);
    particle_system<double> out(N, 'c');
    out.cpu_alloc();
    out.gpu_to_host(&back);
    for (int i = 0; i < N; i++) {
        const bool fast = h.timestep[i] <= 0.25;
        vec3<double> want = fast ? h.pos[i] : h.pos[i] + h.vel[i] * (1.0 * 0.5);
        CHECK((out.pos[i] - want).norm2() == 0.0, "roundtrip pos[%d]", i);
        CHECK(out.m[i] == h.m[i], "roundtrip m[%d]", i);
    }
    std::printf("[PASS?] split/drift/update N=%d fast=%d\n", N, expect_fast);
    h.cpu_free();
    out.cpu_free();
    d.gpu_free();
    back.gpu_free();
}

// Host FP64 energy of a downloaded system (any T), softened with eps2.
template <typename T> double host_energy(const particle_system<T> &h, int N, double eps2) {
    double k = 0.0, u = 0.0;
    for (int i = 0; i < N; i++) {
        const double vx = h.vel[i].x, vy = h.vel[i].y, vz = h.vel[i].z;
        k += 0.5 * double(h.m[i]) * (vx * vx + vy * vy + vz * vz);
        for (int j = i + 1; j < N; j++) {
            const double dx = double(h.pos[i].x) - h.pos[j].x, dy = double(h.pos[i].y) - h.pos[j].y,
                         dz = double(h.pos[i].z) - h.pos[j].z;
            u -= double(h.m[i]) * h.m[j] / std::sqrt(dx * dx + dy * dy + dz * dz + eps2);
        }
    }
    return k + u;
}

template <typename T>
void load_plummer(particle_system<T> &h, int N, uint64_t seed, double eps2) {
    std::vector<double> p(3 * N), v(3 * N), m(N);
    gpuErrchk(hhs_ic_plummer(N, seed, 0, 1, eps2, p.data(), v.data(), m.data()));
    for (int i = 0; i < N; i++) {
        h.pos[i] = vec3<T>(T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2]));
        h.vel[i] = vec3<T>(T(v[3 * i]), T(v[3 * i + 1]), T(v[3 * i + 2]));
        h.acc[i] = vec3<T>();
        h.m[i] = T(m[i]);
        h.timestep[i] = T(0);
        h.id[i] = i;
    }
}

// hold_step<float, BS> (src/hold.cuh:51-53 with T = float): the FP32
// integration conserves energy to the integrator's error, advances simtime
// by dt per root step, and stays within 1e-5 of the rms of the FP64 run.
void test_hold_step_f32() {
    const int N = 256;
    const double eps = 0.01;
    particle_system<float> h(N, 'c'), d(N, 'g'), out(N, 'c');
    particle_system<double> h64(N, 'c'), d64(N, 'g'), out64(N, 'c');
    h.cpu_alloc();
    out.cpu_alloc();
    h64.cpu_alloc();
    out64.cpu_alloc();
    d.gpu_alloc();
    d64.gpu_alloc();
    load_plummer(h, N, 6, eps * eps);
    load_plummer(h64, N, 6, eps * eps);
    // the FP64 run starts from the FP32-rounded state
    for (int i = 0; i < N; i++) {
        h64.pos[i] = vec3<double>(h.pos[i].x, h.pos[i].y, h.pos[i].z);
        h64.vel[i] = vec3<double>(h.vel[i].x, h.vel[i].y, h.vel[i].z);
        h64.m[i] = h.m[i];
    }
    d.host_to_gpu(&h);
    d64.host_to_gpu(&h64);
    const double e0 = host_energy(h, N, eps * eps);
    nbodysystem_globals<float> g32(N, 0.01f, float(eps));
    nbodysystem_globals<double> g64(N, 0.01, eps);
    for (int k = 0; k < 4; k++) {
        hold_step<float, 256>(0, &g32, &d, N, k / 32.0f, (k + 1) / 32.0f, 1.0f / 32, true);
        hold_step<double, 256>(0, &g64, &d64, N, k / 32.0, (k + 1) / 32.0, 1.0 / 32, true);
    }
    out.gpu_to_host(&d);
    out64.gpu_to_host(&d64);
    const double de = std::fabs(host_energy(out, N, eps * eps) / e0 - 1.0);
    double dp = 0.0, rms = 0.0;
    for (int i = 0; i < N; i++) {
        const double dx = out.pos[i].x - out64.pos[i].x, dy = out.pos[i].y - out64.pos[i].y,
                     dz = out.pos[i].z - out64.pos[i].z;
        dp = std::max(dp, std::sqrt(dx * dx + dy * dy + dz * dz));
        rms += out64.pos[i].norm2();
    }
    rms = std::sqrt(rms / N);
    CHECK(de < 1e-5, "hold_step<float>: |dE/E| = %.3e", de);
    CHECK(g32.simtime == 0.125f, "hold_step<float>: simtime %.9g", double(g32.simtime));
    CHECK(dp <= 1e-5 * rms, "hold_step<float>: max |x32 - x64| = %.3e (rms %.3e)", dp, rms);
    std::printf("[PASS?] hold_step<float> N=%d |dE/E|=%.3e max|x32-x64|/rms=%.3e\n", N, de, dp / rms);
    h.cpu_free();
    out.cpu_free();
    h64.cpu_free();
    out64.cpu_free();
    d.gpu_free();
    d64.gpu_free();
}

// hold<T> with the signature the reference declares (src/hold.cuh:20-23):
// eta and eps from the buffers, the same bits as the globals form, and
// calc_timestep = false on the first root step reuses the timesteps held.
void test_hold_declared_signature() {
    const int N = 300;
    const double eps = 0.01;
    particle_system<double> h(N, 'c'), a(N, 'g'), b(N, 'g'), c(N, 'g'), oa(N, 'c'), ob(N, 'c'),
        oc(N, 'c');
    h.cpu_alloc();
    oa.cpu_alloc();
    ob.cpu_alloc();
    oc.cpu_alloc();
    a.gpu_alloc();
    b.gpu_alloc();
    c.gpu_alloc();
    load_plummer(h, N, 8, eps * eps);
    a.host_to_gpu(&h);
    b.host_to_gpu(&h);
    c.host_to_gpu(&h);
    nbodysystem_buffers<double> buf(N, 0.01, eps);
    hold<double>(0, &a, &buf, N, 0.0, 0.125, 1.0 / 32, true);
    nbodysystem_globals<double> g(N, 0.01, eps);
    hold<double>(0, &g, &b, N, 0.0, 0.125, 1.0 / 32);
    // timesteps computed up front, then hold without recomputing them first
    nbodysystem_buffers<double> buf2(N, 0.01, eps);
    timesteps<double, 256>(&c, buf2.hold_globals(), N);
    hold<double>(0, &c, &buf2, N, 0.0, 0.125, 1.0 / 32, false);
    oa.gpu_to_host(&a);
    ob.gpu_to_host(&b);
    oc.gpu_to_host(&c);
    int diff = 0, diffc = 0;
    for (int i = 0; i < N; i++) {
        diff += !(oa.pos[i].x == ob.pos[i].x && oa.pos[i].y == ob.pos[i].y &&
                  oa.pos[i].z == ob.pos[i].z && oa.vel[i].x == ob.vel[i].x &&
                  oa.vel[i].y == ob.vel[i].y && oa.vel[i].z == ob.vel[i].z);
        diffc += !(oa.pos[i].x == oc.pos[i].x && oa.vel[i].z == oc.vel[i].z);
    }
    CHECK(diff == 0, "hold(buffers): %d particles differ from hold(globals)", diff);
    CHECK(diffc == 0, "hold(buffers, calc=false after timesteps): %d particles differ", diffc);
    CHECK(buf.globals && buf.globals->simtime == 0.125, "hold(buffers): simtime");
    std::printf("[PASS?] hold(clevel, sys, buffers, N, stime, etime, dt, calc) N=%d\n", N);
    h.cpu_free();
    oa.cpu_free();
    ob.cpu_free();
    oc.cpu_free();
    a.gpu_free();
    b.gpu_free();
    c.gpu_free();
}

int main() {
    int ndev = 0;
    if (hhs_device_count(&ndev) != HHS_OK || ndev == 0) {
        std::printf("no GPU\n");
        return 77;
    }
    test_split_values<512, 256, 2>();
    test_split_values<4090, 1024, 4>();
    test_split_values<2048, 512, 2>();
    test_split_values<2151, 512, 2>();
    test_split_values<524288, 128, 4>();
    test_split_values<588021, 128, 4>();
    test_split_values<1048576, 128, 4>();
    test_split_values<2097152, 128, 4>();
    test_nbforce();
    test_split_update_roundtrip();
    test_hold_step();
    test_hold_step_f32();
    test_hold_declared_signature();
    std::printf("%s: %d failure(s)\n", g_fail ? "FAILED" : "OK", g_fail);
    return g_fail ? 1 : 0;
}
