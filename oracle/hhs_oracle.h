/*
 * hhs_oracle.h -- FP64 CPU restatement of the HOLD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in the product links or calls it.
 *
 * Every function cites the reference file:line it restates.  Reference
 * defects (repulsive kick sign, N_fast==1 race, missing softening) are NOT
 * reproduced; the intended behaviour is restated instead.
 *
 * Arrays follow the reference layout: positions, velocities and
 * accelerations are AoS vec3<double> (x,y,z interleaved, 3*n doubles),
 * scalars are plain arrays of n entries.
 *
 * Parity status of each function (how it is pinned, see tests/test_oracle_*):
 *   oracle_accel / oracle_kick      closed forms (two-body, symmetric shells),
 *                                   momentum invariant, brute force in numpy
 *   oracle_timesteps                closed form (circular pair), symmetry,
 *                                   independent numpy restatement
 *   oracle_exclusive_scan_i32       serial prefix sum (numpy cumsum)
 *   oracle_split                    numpy stable partition; counts and
 *                                   multisets against the compiled reference
 *                                   cpu_split (oracle/_ref)
 *   oracle_cpu_split                bit-exact against the compiled reference
 *                                   cpu_split (oracle/_ref)
 *   oracle_drift / oracle_update    closed forms
 *   oracle_hold_step / oracle_hold  Kepler two-body period, momentum
 *                                   conservation, second-order energy scaling
 *   oracle_energy                   two-body closed form
 */
#ifndef HHS_ORACLE_H
#define HHS_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One particle set, mirroring particle_system<double>
 * (reference src/particle_system.cuh:20-30). */
typedef struct oracle_particles {
    double *pos;      /* 3n */
    double *vel;      /* 3n */
    double *acc;      /* 3n */
    double *timestep; /* n  */
    double *m;        /* n  */
    int *id;          /* n  */
    int *parent_id;   /* n  */
    int n;
} oracle_particles;

typedef struct oracle_hold_stats {
    double simtime;        /* sum of dt over calls with N_fast == 0 (src/hold.cuh:106-114) */
    long long calls;       /* hold_step invocations */
    long long interactions;/* kick pair evaluations (N_sink * N_source) */
    long long ts_pairs;    /* ordered pair evaluations in timesteps */
    int max_level;         /* deepest clevel reached */
    long long calls_per_level[64];
} oracle_hold_stats;

/* Status codes shared with the product's C ABI. */
#define ORACLE_OK 0
#define ORACLE_ERR_DEPTH 5
#define ORACLE_ERR_ALLOC 2

int oracle_num_threads(void);
void oracle_set_num_threads(int n);

/* acc[i] = sum_j m_j (x_j - x_i) / (|x_j - x_i|^2 + eps2)^{3/2}, skipping
 * pairs whose softened d2 is exactly zero.
 * Restates src/kick.cuh:15-32 with the attractive sign of
 * tests/test_nbforce.cu:14-18. */
void oracle_accel(const double *si_pos, int n_sink,
                  const double *so_pos, const double *so_m, int n_src,
                  double eps2, double *acc_out);

/* Sink-subset variant: only sinks listed in idx[0..k) are evaluated
 * (used for bounded CPU timing samples and large-N spot checks). */
void oracle_accel_subset(const double *si_pos, const int *idx, int k,
                         const double *so_pos, const double *so_m, int n_src,
                         double eps2, double *acc_out /* 3k */);

/* kick: acc = accel(sinks <- sources); vel += acc * (fac * dt).
 * Restates src/kick.cuh:59-106 (self) and :111-155 (cross). */
void oracle_kick(const double *si_pos, double *si_vel, double *si_acc, int n_sink,
                 const double *so_pos, const double *so_m, int n_src,
                 double dt, double fac, double eps2);

/* tau_i = min_{j != i} min(tau_ff, tau_fb), start value 1e30.
 * Restates src/timestep.cuh:17-60 (formula and operation order) and
 * :65-99 (j != i by index), with r^2 -> r^2 + eps2. */
void oracle_timesteps(const double *pos, const double *vel, const double *m,
                      int n, double eta, double eps2, double *timestep);

/* Exclusive prefix sum.  Restates the scan of src/split.cuh:437-565. */
void oracle_exclusive_scan_i32(const int *in, int *out, int n);

/* Fused predicate + scan + stable compaction + counts.
 * predicate[i] = timestep[i] <= pivot (src/split.cuh:107-117); fast particles
 * go to fast[scan[i]], slow ones to slow[i - scan[i]], each with
 * parent_id = i (src/split.cuh:335-357); N_fast == 1 collapses to
 * N_fast = 0, slow = total in order (src/split.cuh:321-324, intended
 * behaviour of the racy :361-368).
 * predicate / scanned may be NULL.  Returns N_fast via *n_fast. */
void oracle_split(const oracle_particles *total, oracle_particles *slow,
                  oracle_particles *fast, int *predicate, int *scanned,
                  double pivot, int n, int *n_slow, int *n_fast);

/* Scalar-array split used by the reference's scan tests
 * (src/scan_optimized_3.cuh:485-515 semantics with split.cuh naming):
 * values <= pivot are fast, others slow; both outputs stable. */
void oracle_split_values_f64(const double *values, double *slow, double *fast,
                             double pivot, int n, int *n_slow, int *n_fast);

/* In-place unstable Hoare partition (src/cpu_scan.h:20-50).
 * N_fast = #(<= pivot) on the left, N_slow on the right.  n == 0 gives 0/0
 * (the reference forms &total[-1] there). Returns 0, or 1 on the
 * reference's "catastrophic failure" guard. */
int oracle_cpu_split_f64(double *total, double pivot, int n, int *n_slow, int *n_fast);
int oracle_cpu_split_f32(float *total, float pivot, int n, int *n_slow, int *n_fast);

/* pos += vel * (dt * fac)  (src/drift.cuh:16-28) */
void oracle_drift(double *pos, const double *vel, int n, double dt, double fac);

/* prev[parent_id[i]] <- cur[i] for pos, vel, m, timestep
 * (src/update_system.cuh:10-25) */
void oracle_update(oracle_particles *prev, const oracle_particles *cur, int n);

/* One HOLD step on `total` (src/hold.cuh:51-181, Appendix-A order).
 * max_level bounds the recursion (the reference has no guard). */
int oracle_hold_step(oracle_particles *total, double dt, int calc_timestep,
                     double eta, double eps2, int max_level,
                     oracle_hold_stats *stats);

/* Top-level driver the reference declares but never defines
 * (src/hold.cuh:20-35): root steps of dt_root from 0 to t_end. */
int oracle_hold(oracle_particles *sys, double t_end, double dt_root,
                double eta, double eps2, int max_level, oracle_hold_stats *stats);

/* E = sum 1/2 m v^2 - sum_{i<j} m_i m_j / sqrt(r^2 + eps2) */
double oracle_energy(const double *pos, const double *vel, const double *m,
                     int n, double eps2);
double oracle_kinetic(const double *vel, const double *m, int n);
double oracle_potential(const double *pos, const double *m, int n, double eps2);

#ifdef __cplusplus
}
#endif
#endif
