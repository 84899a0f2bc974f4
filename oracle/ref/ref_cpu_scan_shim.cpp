// Thin extern "C" entry points over the reference's own CPU partition
// (src/cpu_scan.h:20-50), compiled straight from /root/reference by
// oracle/Makefile into oracle/_ref/ with no stand-ins (the header needs only
// the C++ standard library; <cstdlib> first, as the reference's own drivers
// include it, perf/perf_scan_multi_block.cu:2).  Test infrastructure only:
// used to pin the oracle's restatement in this container.
#include <cstdlib>
#include <cstdio>
#include "cpu_scan.h"

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

}
