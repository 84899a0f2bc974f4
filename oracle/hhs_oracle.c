/*
 * hhs_oracle.c -- FP64 CPU restatement of the HOLD hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Built with -ffp-contract=off so
 * every expression rounds exactly as written; the HIP kernels that claim
 * bit-exact parity (drift, split, scan, update, exact-mode timesteps) use
 * the same operation order.
 */
#include "hhs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_HUGE 1.e30 /* src/timestep.cuh:4 */

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ---------------------------------------------------------------- kick -- */

/* One sink against all sources.  Loop body restates src/kick.cuh:19-30:
 * dx = x_j - x_i, d2 = |dx|^2 (+eps2), skip d2 == 0, d = sqrt(d2),
 * 1/(d2*d), and the attractive sign of tests/test_nbforce.cu:14-18
 * (there dx = x_i - x_j times -m_j; here dx = x_j - x_i times +m_j). */
static void accel_one(const double *xi, const double *so_pos, const double *so_m,
                      int n_src, double eps2, double *out) {
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int j = 0; j < n_src; j++) {
        double dx = so_pos[3 * j + 0] - xi[0];
        double dy = so_pos[3 * j + 1] - xi[1];
        double dz = so_pos[3 * j + 2] - xi[2];
        double d2 = dx * dx + dy * dy + dz * dz + eps2;
        if (d2 == 0.0) continue;
        double d = sqrt(d2);
        double one_over_d3 = 1.0 / (d2 * d);
        double s = so_m[j] * one_over_d3;
        ax = ax + dx * s;
        ay = ay + dy * s;
        az = az + dz * s;
    }
    out[0] = ax;
    out[1] = ay;
    out[2] = az;
}

void oracle_accel(const double *si_pos, int n_sink, const double *so_pos,
                  const double *so_m, int n_src, double eps2, double *acc_out) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n_sink; i++)
        accel_one(&si_pos[3 * i], so_pos, so_m, n_src, eps2, &acc_out[3 * i]);
}

void oracle_accel_subset(const double *si_pos, const int *idx, int k,
                         const double *so_pos, const double *so_m, int n_src,
                         double eps2, double *acc_out) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int q = 0; q < k; q++)
        accel_one(&si_pos[3 * idx[q]], so_pos, so_m, n_src, eps2, &acc_out[3 * q]);
}

/* src/kick.cuh:98-104 (vel += acc * (fac * dt)); acc holds the whole sum of
 * this call (the reference keeps only the last partial, defect D18). */
void oracle_kick(const double *si_pos, double *si_vel, double *si_acc, int n_sink,
                 const double *so_pos, const double *so_m, int n_src,
                 double dt, double fac, double eps2) {
    double h = fac * dt;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n_sink; i++) {
        double a[3];
        accel_one(&si_pos[3 * i], so_pos, so_m, n_src, eps2, a);
        if (si_acc) {
            si_acc[3 * i + 0] = a[0];
            si_acc[3 * i + 1] = a[1];
            si_acc[3 * i + 2] = a[2];
        }
        si_vel[3 * i + 0] = si_vel[3 * i + 0] + a[0] * h;
        si_vel[3 * i + 1] = si_vel[3 * i + 1] + a[1] * h;
        si_vel[3 * i + 2] = si_vel[3 * i + 2] + a[2] * h;
    }
}

/* ----------------------------------------------------------- timesteps -- */

/* Restates _get_pp_timestep, src/timestep.cuh:17-60, in its operation
 * order: dr = x_i - x_j, dv = v_i - v_j, dot(dr,dv) summed as
 * dv.x*dr.x + dv.y*dr.y + dv.z*dr.z (src/vec.cuh:31), norm2 as
 * x*x + y*y + z*z (src/vec.cuh:44), plus eps2 on r^2. */
static double timestep_one(int i, const double *pos, const double *vel,
                           const double *m, int n, double eta, double eps2) {
    const double M_SQRT2_ = 1.41421356237309504880;
    double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
    double vxi = vel[3 * i], vyi = vel[3 * i + 1], vzi = vel[3 * i + 2];
    double mi = m[i];
    double minval = ORACLE_HUGE;
    for (int j = 0; j < n; j++) {
        if (j == i) continue;
        double drx = xi - pos[3 * j], dry = yi - pos[3 * j + 1], drz = zi - pos[3 * j + 2];
        double dvx = vxi - vel[3 * j], dvy = vyi - vel[3 * j + 1], dvz = vzi - vel[3 * j + 2];
        double dr2 = drx * drx + dry * dry + drz * drz + eps2;
        double dv2 = dvx * dvx + dvy * dvy + dvz * dvz;
        double r = sqrt(dr2);
        double vdotdr2 = (dvx * drx + dvy * dry + dvz * drz) / dr2;
        double tau = eta / M_SQRT2_ * r * sqrt(r / (m[j] + mi));
        double dtau = 3 * tau * vdotdr2 / 2;
        dtau = (dtau < 1) ? dtau : 1;
        tau /= (1 - dtau / 2);
        if (tau < minval) minval = tau;
        tau = eta * r / sqrt(dv2);
        dtau = tau * vdotdr2 * (1 + (mi + m[j]) / (dv2 * r));
        dtau = (dtau < 1) ? dtau : 1;
        tau /= (1 - dtau / 2);
        if (tau < minval) minval = tau;
    }
    return minval;
}

void oracle_timesteps(const double *pos, const double *vel, const double *m,
                      int n, double eta, double eps2, double *timestep) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++)
        timestep[i] = timestep_one(i, pos, vel, m, n, eta, eps2);
}

/* ---------------------------------------------------------- scan/split -- */

void oracle_exclusive_scan_i32(const int *in, int *out, int n) {
    int run = 0;
    for (int i = 0; i < n; i++) {
        int v = in[i];
        out[i] = run;
        run += v;
    }
}

static void copy_record(oracle_particles *dst, int d, const oracle_particles *src, int s,
                        int parent) {
    for (int k = 0; k < 3; k++) {
        dst->pos[3 * d + k] = src->pos[3 * s + k];
        dst->vel[3 * d + k] = src->vel[3 * s + k];
    }
    dst->m[d] = src->m[s];
    dst->timestep[d] = src->timestep[s];
    dst->id[d] = src->id[s];
    dst->parent_id[d] = parent;
}

void oracle_split(const oracle_particles *total, oracle_particles *slow,
                  oracle_particles *fast, int *predicate, int *scanned,
                  double pivot, int n, int *n_slow, int *n_fast) {
    int nf = 0;
    /* predicate and exclusive scan (src/split.cuh:107-117, :437-565) */
    for (int i = 0; i < n; i++) {
        int p = total->timestep[i] <= pivot;
        if (predicate) predicate[i] = p;
        if (scanned) scanned[i] = nf;
        nf += p;
    }
    if (nf == 1) {
        /* src/split.cuh:321-324: a lone fast particle stays slow */
        for (int i = 0; i < n; i++) copy_record(slow, i, total, i, i);
        *n_fast = 0;
        *n_slow = n;
        return;
    }
    int f = 0, s = 0;
    for (int i = 0; i < n; i++) {
        if (total->timestep[i] <= pivot)
            copy_record(fast, f++, total, i, i);
        else
            copy_record(slow, s++, total, i, i);
    }
    *n_fast = f;
    *n_slow = s;
}

void oracle_split_values_f64(const double *values, double *slow, double *fast,
                             double pivot, int n, int *n_slow, int *n_fast) {
    int f = 0, s = 0;
    for (int i = 0; i < n; i++) {
        if (values[i] <= pivot)
            fast[f++] = values[i];
        else
            slow[s++] = values[i];
    }
    *n_fast = f;
    *n_slow = s;
}

/* src/cpu_scan.h:20-50, restated line by line (pointer walk, SWAP, the
 * i > N guard and the final counts). */
#define ORACLE_CPU_SPLIT_BODY(T)                                              \
    if (n <= 0) {                                                             \
        *n_slow = 0;                                                          \
        *n_fast = 0;                                                          \
        return 0;                                                             \
    }                                                                         \
    T *left = total, *right = &total[n - 1];                                  \
    int i = 0;                                                                \
    for (;;) {                                                                \
        if (i > n) return 1;                                                  \
        i++;                                                                  \
        while (*left <= pivot && left < right) left++;                        \
        while (*right > pivot && left < right) right--;                       \
        if (left < right) {                                                   \
            T t = *left;                                                      \
            *left = *right;                                                   \
            *right = t;                                                       \
        } else {                                                              \
            break;                                                            \
        }                                                                     \
    }                                                                         \
    if (*left <= pivot) left++;                                               \
    *n_slow = (int)(&total[n - 1] - left + 1);                                \
    *n_fast = (int)(left - total);                                            \
    return 0;

int oracle_cpu_split_f64(double *total, double pivot, int n, int *n_slow, int *n_fast) {
    ORACLE_CPU_SPLIT_BODY(double)
}

int oracle_cpu_split_f32(float *total, float pivot, int n, int *n_slow, int *n_fast) {
    ORACLE_CPU_SPLIT_BODY(float)
}

/* ------------------------------------------------------- drift, update -- */

void oracle_drift(double *pos, const double *vel, int n, double dt, double fac) {
    double h = dt * fac; /* src/drift.cuh:25: vel * (dt * fac) */
    for (int i = 0; i < 3 * n; i++) pos[i] = pos[i] + vel[i] * h;
}

void oracle_update(oracle_particles *prev, const oracle_particles *cur, int n) {
    for (int i = 0; i < n; i++) {
        int p = cur->parent_id[i];
        for (int k = 0; k < 3; k++) {
            prev->pos[3 * p + k] = cur->pos[3 * i + k];
            prev->vel[3 * p + k] = cur->vel[3 * i + k];
        }
        prev->m[p] = cur->m[i];
        prev->timestep[p] = cur->timestep[i];
    }
}

/* ---------------------------------------------------------------- HOLD -- */

static int alloc_particles(oracle_particles *p, int n) {
    size_t nn = n > 0 ? (size_t)n : 1;
    p->n = n;
    p->pos = (double *)malloc(3 * nn * sizeof(double));
    p->vel = (double *)malloc(3 * nn * sizeof(double));
    p->acc = (double *)calloc(3 * nn, sizeof(double));
    p->timestep = (double *)malloc(nn * sizeof(double));
    p->m = (double *)malloc(nn * sizeof(double));
    p->id = (int *)malloc(nn * sizeof(int));
    p->parent_id = (int *)malloc(nn * sizeof(int));
    return p->pos && p->vel && p->acc && p->timestep && p->m && p->id && p->parent_id;
}

static void free_particles(oracle_particles *p) {
    free(p->pos);
    free(p->vel);
    free(p->acc);
    free(p->timestep);
    free(p->m);
    free(p->id);
    free(p->parent_id);
}

static int hold_rec(int clevel, oracle_particles *total, int n, double dt, int calc,
                    double eta, double eps2, int max_level, oracle_hold_stats *st) {
    if (clevel > max_level) return ORACLE_ERR_DEPTH;
    if (st) {
        st->calls++;
        if (clevel < 64) st->calls_per_level[clevel]++;
        if (clevel > st->max_level) st->max_level = clevel;
    }
    /* src/hold.cuh:74-76 */
    if (calc) {
        oracle_timesteps(total->pos, total->vel, total->m, n, eta, eps2, total->timestep);
        if (st) st->ts_pairs += (long long)n * (n - 1);
    }
    oracle_particles slow, fast;
    if (!alloc_particles(&slow, n) || !alloc_particles(&fast, n)) {
        free_particles(&slow);
        free_particles(&fast);
        return ORACLE_ERR_ALLOC;
    }
    int ns = 0, nf = 0;
    /* src/hold.cuh:79-83 */
    oracle_split(total, &slow, &fast, NULL, NULL, dt, n, &ns, &nf);
    int rc = ORACLE_OK;
    /* src/hold.cuh:106-114 */
    if (nf == 0 && st) st->simtime += dt;
    /* src/hold.cuh:116-121 */
    if (nf > 0) rc = hold_rec(clevel + 1, &fast, nf, dt / 2, 0, eta, eps2, max_level, st);
    if (rc == ORACLE_OK) {
        /* src/hold.cuh:123-124 */
        if (ns > 0) oracle_drift(slow.pos, slow.vel, ns, dt, 0.5);
        /* src/hold.cuh:137-145: S<-S, S<-F, F<-S, each with fac = 1 */
        if (ns > 0) {
            oracle_kick(slow.pos, slow.vel, slow.acc, ns, slow.pos, slow.m, ns, dt, 1.0, eps2);
            if (st) st->interactions += (long long)ns * ns;
            if (nf > 0) {
                oracle_kick(slow.pos, slow.vel, slow.acc, ns, fast.pos, fast.m, nf, dt, 1.0, eps2);
                oracle_kick(fast.pos, fast.vel, fast.acc, nf, slow.pos, slow.m, ns, dt, 1.0, eps2);
                if (st) st->interactions += 2LL * ns * nf;
            }
        }
        /* src/hold.cuh:158-159 */
        if (ns > 0) oracle_drift(slow.pos, slow.vel, ns, dt, 0.5);
        /* src/hold.cuh:168-173 */
        if (nf > 0) rc = hold_rec(clevel + 1, &fast, nf, dt / 2, 1, eta, eps2, max_level, st);
        /* src/hold.cuh:177-178 */
        if (rc == ORACLE_OK) {
            oracle_update(total, &fast, nf);
            oracle_update(total, &slow, ns);
        }
    }
    free_particles(&slow);
    free_particles(&fast);
    return rc;
}

int oracle_hold_step(oracle_particles *total, double dt, int calc_timestep,
                     double eta, double eps2, int max_level, oracle_hold_stats *stats) {
    return hold_rec(0, total, total->n, dt, calc_timestep, eta, eps2, max_level, stats);
}

int oracle_hold(oracle_particles *sys, double t_end, double dt_root, double eta,
                double eps2, int max_level, oracle_hold_stats *stats) {
    if (!(dt_root > 0.0)) return 1;
    double t = 0.0;
    while (t < t_end) {
        int rc = oracle_hold_step(sys, dt_root, 1, eta, eps2, max_level, stats);
        if (rc != ORACLE_OK) return rc;
        t += dt_root;
    }
    return ORACLE_OK;
}

/* -------------------------------------------------------------- energy -- */

double oracle_kinetic(const double *vel, const double *m, int n) {
    double ke = 0.0;
    for (int i = 0; i < n; i++) {
        double v2 = vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] +
                    vel[3 * i + 2] * vel[3 * i + 2];
        ke += 0.5 * m[i] * v2;
    }
    return ke;
}

double oracle_potential(const double *pos, const double *m, int n, double eps2) {
    double *row = (double *)calloc(n > 0 ? (size_t)n : 1, sizeof(double));
    if (!row) return NAN;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = i + 1; j < n; j++) {
            double dx = pos[3 * j] - pos[3 * i];
            double dy = pos[3 * j + 1] - pos[3 * i + 1];
            double dz = pos[3 * j + 2] - pos[3 * i + 2];
            double r2 = dx * dx + dy * dy + dz * dz + eps2;
            if (r2 == 0.0) continue;
            s += m[j] / sqrt(r2);
        }
        row[i] = m[i] * s;
    }
    double pe = 0.0;
    for (int i = 0; i < n; i++) pe -= row[i];
    free(row);
    return pe;
}

double oracle_energy(const double *pos, const double *vel, const double *m, int n,
                     double eps2) {
    return oracle_kinetic(vel, m, n) + oracle_potential(pos, m, n, eps2);
}
