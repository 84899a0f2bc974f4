/* Drives every entry point of the FP64 oracle (oracle/hhs_oracle.c) on
 * small inputs and edge cases, for a build under AddressSanitizer and
 * UndefinedBehaviorSanitizer (tests/test_sanitize.py compiles it together
 * with the oracle's source).  Test infrastructure: it checks the oracle's
 * memory and arithmetic hygiene, not its numbers, which tests/test_oracle.py
 * pins.  Exit status 0 when every call returned as documented.  -DPLANT_OOB
 * plants a write one past an array, which the sanitizer build must catch. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hhs_oracle.h"

static unsigned long long rng_state = 88172645463325252ull;
static double urand(void) {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (double)(rng_state >> 11) * (1.0 / 9007199254740992.0);
}


static oracle_particles alloc_set(int n) {
    oracle_particles p;
    const size_t k = n > 0 ? (size_t)n : 1;
    p.pos = calloc(3 * k, sizeof(double));
    p.vel = calloc(3 * k, sizeof(double));
    p.acc = calloc(3 * k, sizeof(double));
    p.timestep = calloc(k, sizeof(double));
    p.m = calloc(k, sizeof(double));
    p.id = calloc(k, sizeof(int));
    p.parent_id = calloc(k, sizeof(int));
    p.n = n;
    return p;
}

static void free_set(oracle_particles *p) {
    free(p->pos);
    free(p->vel);
    free(p->acc);
    free(p->timestep);
    free(p->m);
    free(p->id);
    free(p->parent_id);
}

static void fill(oracle_particles *p, double spread) {
    for (int i = 0; i < p->n; ++i) {
        for (int k = 0; k < 3; ++k) {
            p->pos[3 * i + k] = spread * (2.0 * urand() - 1.0);
            p->vel[3 * i + k] = 0.3 * (2.0 * urand() - 1.0);
        }
        p->m[i] = 1.0 / (p->n > 0 ? p->n : 1);
        p->id[i] = i;
    }
}

static int failures = 0;
#define CHECK(c, ...)                                              \
    do {                                                           \
        if (!(c)) {                                                \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);   \
            fprintf(stderr, __VA_ARGS__);                          \
            fprintf(stderr, "\n");                                 \
            ++failures;                                            \
        }                                                          \
    } while (0)

int main(void) {
    oracle_set_num_threads(2);
    CHECK(oracle_num_threads() >= 1, "threads");
    const int sizes[] = {0, 1, 2, 3, 17, 64, 257};
    for (size_t q = 0; q < sizeof(sizes) / sizeof(sizes[0]); ++q) {
        const int n = sizes[q];
        oracle_particles a = alloc_set(n);
        fill(&a, 1.0);
        const double eps2 = (q % 2) ? 1e-4 : 0.0;
        /* forces, kicks, timesteps */
        double *acc = calloc(3 * (size_t)(n > 0 ? n : 1), sizeof(double));
        oracle_accel(a.pos, n, a.pos, a.m, n, eps2, acc);
#ifdef PLANT_OOB
        acc[3 * (n > 0 ? n : 1)] = 0.0; /* planted: one past the end (must be caught) */
#endif
        int idx[3] = {0, n > 1 ? n - 1 : 0, n > 2 ? n / 2 : 0};
        const int k = n < 3 ? n : 3;
        oracle_accel_subset(a.pos, idx, k, a.pos, a.m, n, eps2, acc);
        oracle_kick(a.pos, a.vel, a.acc, n, a.pos, a.m, n, 0.01, 1.0, eps2);
        oracle_timesteps(a.pos, a.vel, a.m, n, 0.01, eps2, a.timestep);
        for (int i = 0; i < n; ++i) CHECK(a.timestep[i] > 0.0, "timestep %d", i);
        /* scan */
        int *in = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        int *out = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) in[i] = (int)(urand() * 3.0);
        oracle_exclusive_scan_i32(in, out, n);
        int run = 0;
        for (int i = 0; i < n; ++i) {
            CHECK(out[i] == run, "scan %d", i);
            run += in[i];
        }
        /* splits at pivots that give none, some, all, and one fast */
        oracle_particles slow = alloc_set(n), fast = alloc_set(n);
        int *pred = calloc((size_t)(n > 0 ? n : 1), sizeof(int));
        int *scanned = calloc((size_t)(n > 0 ? n : 1), sizeof(int));
        double tmin = 1e300, tmax = -1e300;
        for (int i = 0; i < n; ++i) {
            tmin = fmin(tmin, a.timestep[i]);
            tmax = fmax(tmax, a.timestep[i]);
        }
        const double pivots[] = {-1.0, 0.5 * (tmin + tmax), tmax, tmin};
        for (int pv = 0; pv < 4; ++pv) {
            int ns = -1, nf = -1;
            oracle_split(&a, &slow, &fast, pred, scanned, pivots[pv], n, &ns, &nf);
            CHECK(ns >= 0 && nf >= 0 && ns + nf == n, "split counts %d %d", ns, nf);
            CHECK(nf != 1, "lone fast particle reported");
            oracle_split(&a, &slow, &fast, NULL, NULL, pivots[pv], n, &ns, &nf);
            /* scatter back: the parent's set is restored */
            oracle_particles back = alloc_set(n);
            oracle_update(&back, &slow, ns);
            oracle_update(&back, &fast, nf);
            for (int i = 0; i < n; ++i) CHECK(back.m[i] == a.m[i], "update %d", i);
            free_set(&back);
        }
        /* value splits and the in-place partition */
        double *vals = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        double *vs = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        double *vf = malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        float *fv = malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) {
            vals[i] = 10.0 * urand();
            fv[i] = (float)vals[i];
        }
        int ns = 0, nf = 0;
        oracle_split_values_f64(vals, vs, vf, 5.0, n, &ns, &nf);
        CHECK(ns + nf == n, "value split");
        CHECK(oracle_cpu_split_f64(vals, 5.0, n, &ns, &nf) == 0 && ns + nf == n, "cpu_split f64");
        CHECK(oracle_cpu_split_f32(fv, 5.0f, n, &ns, &nf) == 0 && ns + nf == n, "cpu_split f32");
        /* drift, energy */
        oracle_drift(a.pos, a.vel, n, 0.01, 0.5);
        const double e = oracle_energy(a.pos, a.vel, a.m, n, eps2);
        CHECK(isfinite(e) || n < 2, "energy");
        CHECK(oracle_kinetic(a.vel, a.m, n) >= 0.0, "kinetic");
        (void)oracle_potential(a.pos, a.m, n, eps2);
        /* HOLD: a few root steps, and the depth guard */
        if (n >= 1 && n <= 64) {
            oracle_hold_stats st;
            memset(&st, 0, sizeof(st));
            CHECK(oracle_hold(&a, 1.0 / 32, 1.0 / 64, 0.05, 1e-4, 40, &st) == ORACLE_OK,
                  "hold n=%d", n);
            memset(&st, 0, sizeof(st));
            const int rc = oracle_hold_step(&a, 1.0, 1, 0.05, 1e-4, 1, &st);
            CHECK(rc == ORACLE_OK || rc == ORACLE_ERR_DEPTH, "hold_step guard rc %d", rc);
        }
        free(vals);
        free(vs);
        free(vf);
        free(fv);
        free(pred);
        free(scanned);
        free(in);
        free(out);
        free(acc);
        free_set(&slow);
        free_set(&fast);
        free_set(&a);
    }
    /* a coincident pair with eps = 0 drives the recursion to its guard */
    {
        oracle_particles c = alloc_set(2);
        c.m[0] = c.m[1] = 0.5;
        c.vel[0] = 0.1;
        oracle_hold_stats st;
        memset(&st, 0, sizeof(st));
        const int rc = oracle_hold_step(&c, 1.0, 1, 0.01, 0.0, 12, &st);
        CHECK(rc == ORACLE_ERR_DEPTH || rc == ORACLE_OK, "coincident pair rc %d", rc);
        free_set(&c);
    }
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("oracle sanitizer driver: ok\n");
    return 0;
}
