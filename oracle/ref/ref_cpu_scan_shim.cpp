// Thin extern "C" entry points over the reference's own CPU partition
// (src/cpu_scan.h:20-50) and vec3 arithmetic (src/vec.cuh:5-53), compiled
// straight from /root/reference by oracle/Makefile into oracle/_ref/.
// Test infrastructure only: used to pin the oracle's restatement.
#include <cstdlib>
#include <cstdio>
#include "cpu_scan.h"
#include "vec.cuh"

extern "C" {

int ref_cpu_split_f64(double* total, double pivot, int n, int* n_slow, int* n_fast) {
    partition_c sizes{0, 0};
    cpu_split<double>(total, nullptr, nullptr, pivot, &sizes, n);
    *n_slow = sizes.N_slow;
    *n_fast = sizes.N_fast;
    return 0;
}

int ref_cpu_split_f32(float* total, float pivot, int n, int* n_slow, int* n_fast) {
    partition_c sizes{0, 0};
    cpu_split<float>(total, nullptr, nullptr, pivot, &sizes, n);
    *n_slow = sizes.N_slow;
    *n_fast = sizes.N_fast;
    return 0;
}

double ref_vec3_dot(const double* a, const double* b) {
    vec3<double> va(a[0], a[1], a[2]), vb(b[0], b[1], b[2]);
    return va * vb;
}

double ref_vec3_norm2(const double* a) {
    vec3<double> va(a[0], a[1], a[2]);
    return va.norm2();
}

void ref_vec3_sub(const double* a, const double* b, double* out) {
    vec3<double> va(a[0], a[1], a[2]), vb(b[0], b[1], b[2]);
    vec3<double> r = va - vb;
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}

}
