/*
 * hhs.h -- C ABI of the MI355X HOLD hot path (libhhs.so).
 *
 * Plain C: device pointers, sizes and status codes.  No HIP or torch types
 * appear here; a stream is passed as `void *` (a hipStream_t, NULL for the
 * default stream).  Every function returns HHS_OK (0) or a positive status
 * code and never calls exit(); hhs_last_error() gives the message of the
 * last failure on the calling thread.
 *
 * Each entry point names the reference interface it replaces
 * (dipto4/parallel-hhs, file:line).  include/hhs/hold.hpp re-exposes the
 * reference's C++ templates (particle_system<T>, hold_step<T,BS>, ...) on
 * top of these functions.
 *
 * Layout: positions, velocities and accelerations are AoS vec3 (x,y,z
 * interleaved, 3*n values) exactly as vec3<T> arrays in the reference
 * (src/vec.cuh:5-53, src/particle_system.cuh:23-25); timestep and m are n
 * values; id and parent_id are n ints.
 */
#ifndef HHS_H
#define HHS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status -- */
#define HHS_OK 0
#define HHS_ERR_INVALID 1   /* bad argument (null pointer, negative size)   */
#define HHS_ERR_ALLOC 2     /* device or host allocation failed             */
#define HHS_ERR_HIP 3       /* a HIP runtime call or launch failed          */
#define HHS_ERR_NO_DEVICE 4 /* no usable GPU                                */
#define HHS_ERR_DEPTH 5     /* HOLD recursion exceeded max_level            */
#define HHS_ERR_BUDGET 6    /* request exceeds the device-memory budget     */
#define HHS_ERR_NONFINITE 7 /* NaN/Inf detected in a diagnostic             */
#define HHS_ERR_WORKSPACE 8 /* workspace too small / not initialised        */
#define HHS_ERR_IO 9        /* a snapshot file could not be opened/written  */
#define HHS_ERR_FORMAT 10   /* not a snapshot, or a damaged one             */

/* Message of the last failure on this thread ("" if none). */
const char *hhs_last_error(void);
/* Library version string. */
const char *hhs_version(void);
/* HHS_OK when one HIP runtime is mapped into the process; HHS_ERR_HIP, with
 * both paths in hhs_last_error(), when two are (PyTorch-ROCm bundles a
 * libamdhip64.so.7 of its own: the first device call of a process holding
 * both crashes).  A host that uses PyTorch loads torch's runtime before this
 * library, so that libamdhip64.so.7 resolves to it.  No reference
 * counterpart. */
int hhs_runtime_check(void);

/* ------------------------------------------------------- data model -- */

/* Device-side particle set; mirrors the fields of particle_system<T>
 * (src/particle_system.cuh:20-30).  Any pointer may be NULL when the
 * called function does not touch that field. */
typedef struct hhs_particles_f64 {
    double *pos;      /* 3n */
    double *vel;      /* 3n */
    double *acc;      /* 3n */
    double *timestep; /* n  */
    double *m;        /* n  */
    int *id;          /* n  */
    int *parent_id;   /* n  */
    int n;
} hhs_particles_f64;

typedef struct hhs_particles_f32 {
    float *pos;
    float *vel;
    float *acc;
    float *timestep;
    float *m;
    int *id;
    int *parent_id;
    int n;
} hhs_particles_f32;

/* Slow/fast counts; same layout as `struct partition`
 * (src/particle_system.cuh:8-11). */
typedef struct hhs_partition {
    int N_slow;
    int N_fast;
} hhs_partition;

/* ---------------------------------------------- device memory helpers -- */
/* Used by the C++ drop-in header for particle_system<T>::gpu_alloc and
 * friends (src/particle_system.cuh:135-247). */
int hhs_device_count(int *count);
int hhs_malloc(void **ptr, size_t bytes);
int hhs_free(void *ptr);
int hhs_memcpy_h2d(void *dst, const void *src, size_t bytes);
int hhs_memcpy_d2h(void *dst, const void *src, size_t bytes);
int hhs_memcpy_d2d(void *dst, const void *src, size_t bytes);
int hhs_memset(void *dst, int value, size_t bytes);
int hhs_device_synchronize(void);
int hhs_stream_synchronize(void *stream);
/* Pinned host memory that kernels can write directly; stands in for the
 * reference's cudaMallocManaged partition/params (src/particle_system.cuh:81,
 * :123) so the host can read counts after a synchronisation. */
int hhs_host_alloc_mapped(void **ptr, size_t bytes);
int hhs_host_free_mapped(void *ptr);

/* ------------------------------------------------------------- kicks -- */
/* Direct-sum kick with Plummer softening eps2 = eps^2 and the attractive
 * sign (the reference GPU kick is repulsive, src/kick.cuh:23,28; the sign
 * here follows tests/test_nbforce.cu:14-18).  Pairs whose softened r^2 is
 * exactly 0 are skipped (src/kick.cuh:25).  acc (if non-NULL) receives the
 * acceleration of this call; vel += acc * (fac * dt).
 * FP64 storage: separations are formed in FP64 and rounded to FP32, forces
 * are summed in FP32 with double-single (two-sum) compensation. */

/* kick_slow (src/kick.cuh:158-168): sinks = sources = sys[0..n). */
int hhs_kick_self_f64(const hhs_particles_f64 *sys, int n, double dt, double fac,
                      double eps2, void *stream);
int hhs_kick_self_f32(const hhs_particles_f32 *sys, int n, float dt, float fac,
                      float eps2, void *stream);

/* kick_sf (src/kick.cuh:170-181): sinks si[0..n_sink) kicked by sources
 * so[0..n_src). */
int hhs_kick_cross_f64(const hhs_particles_f64 *so, const hhs_particles_f64 *si,
                       int n_sink, int n_src, double dt, double fac, double eps2,
                       void *stream);
int hhs_kick_cross_f32(const hhs_particles_f32 *so, const hhs_particles_f32 *si,
                       int n_sink, int n_src, float dt, float fac, float eps2,
                       void *stream);

/* Kick tuning: sinks per lane (1,2,4), waves per workgroup sharing one sink
 * tile (1,2,4,8).  0/0 restores the built-in heuristic.  Applies to calls
 * issued afterwards from any thread. */
int hhs_kick_set_config(int sinks_per_lane, int waves_per_group);

/* Kick arithmetic for FP64 particle sets (FP32 sets always use MIXED32):
 * MIXED32:   x_j - x_i in FP64 rounded to FP32; r^2, rsqrt, m r^-3 and the
 *            sum in FP32 with two-sum compensation (15 VALU ops/interaction)
 * F64_RSQ32: separations, r^2 and the sum in FP64; rsqrt and m r^-3 in FP32
 *            (14 ops)
 * F64:       FP64 throughout with the hardware 1/sqrt estimate (relative
 *            error <= 2^-23) (12 ops)
 * F64_EXACT: FP64 in the CPU direct sum's operation order with IEEE sqrt
 *            and divide, each sink summing its sources in index order:
 *            bit-exact with the FP64 CPU kick (src/kick.cuh:19-30 with the
 *            attractive sign); one wave per group, several times slower.
 * All four meet the 1e-6 normalised force bar. */
#define HHS_KICK_MIXED32 0
#define HHS_KICK_F64_RSQ32 1
#define HHS_KICK_F64 2
#define HHS_KICK_F64_EXACT 3
int hhs_kick_set_precision(int mode);
int hhs_kick_get_precision(void);

/* Kick autotuner (no reference counterpart).  Times every (sinks per lane,
 * waves per group) configuration on a synthetic n-particle FP64 system
 * (reps timed launches each, default stream, blocking), installs the
 * fastest for sink counts in n's power-of-two size class, and appends a
 * line "kick <class> <S> <W> <ms>" to cache_path when it is non-NULL.
 * hhs_autotune_load installs the entries of such a file (later lines
 * win); hhs_autotune_clear returns to the built-in heuristic.  An explicit
 * hhs_kick_set_config overrides tuned entries. */
int hhs_autotune_kick(int n, int reps, const char *cache_path, int *best_s, int *best_w,
                      double *best_ms);
/* Timestep autotuner (no reference counterpart).  Times the one-sided
 * timestep kernel of `mode` (HHS_TIMESTEP_*) at every (waves per group W,
 * source chunks K) that keeps each wave >= 8 sources, on a synthetic
 * n-particle FP64 system (reps timed launches each, default stream,
 * blocking), installs the fastest for that mode and n's power-of-two size
 * class (FP64 storage), and appends "ts <mode> <class> <W> <K> <ms>" to
 * cache_path when it is non-NULL.  Minima do not depend on the split, so
 * EXACT stays bit-exact under any tuned configuration. */
int hhs_autotune_timesteps(int n, int mode, int reps, const char *cache_path, int *best_w,
                           int *best_k, double *best_ms);
/* The (waves per group, source chunks) the one-sided timestep kernel of
 * `mode` uses for n particles in FP64 storage: tuned, or the heuristic. */
int hhs_timesteps_get_config(int n, int mode, int *waves_per_group, int *chunks);
/* Installs the "kick" and "ts" entries of a cache file (later lines win);
 * hhs_autotune_clear returns both to the built-in heuristics. */
int hhs_autotune_load(const char *cache_path);
int hhs_autotune_clear(void);
int hhs_kick_get_config(int n_sink, int n_src, int *sinks_per_lane, int *waves_per_group);

/* --------------------------------------------------------- timesteps -- */
/* timesteps (src/timestep.cuh:103-118): timestep[i] = min over j != i of
 * the free-fall and fly-by pair timesteps (src/timestep.cuh:17-60), with
 * r^2 -> r^2 + eps2, start value 1e30.  FP64, same operation order and
 * IEEE sqrt/divide as the formula, so results are bit-exact against an
 * FP64 CPU evaluation in that order. */
int hhs_timesteps_f64(const hhs_particles_f64 *sys, int n, double eta, double eps2,
                      void *stream);
/* Symmetric evaluation of the same minimum: tau_ij = tau_ji bit for bit, so
 * each unordered pair is evaluated once and feeds both particles (blocks of
 * particles, partial minima in `ws`, then a min-reduction; no atomics).
 * HHS_TIMESTEP_EXACT stays bit-exact with hhs_timesteps_f64; FP32 (from
 * double-single FP32 copies of the positions and velocities, packed into
 * `ws`) is within 1e-6; HHS_TIMESTEP_FAST uses the one-sided kernel.  ws must
 * hold hhs_timesteps_sym_workspace_bytes(n) bytes: the larger of
 * ceil(n/1024) * n doubles (exact) and, for FP32, 64 n bytes + the larger of
 * (ceil(n/256) + ceil(n/768)) * n and (ceil(k/64) + ceil(k/768)) * k floats
 * with k = min(n, 65536), + the decision guard's area.
 * No reference counterpart (the reference evaluates every ordered pair). */
size_t hhs_timesteps_sym_workspace_bytes(int n);
/* Sinks per lane of the FP32 symmetric kernel: 0 (default: 3, the faster
 * at every size measured) or 3 or 4; the workspace size above does not
 * change.  Process-wide; the timesteps are the same bits either way (the
 * minimum over a particle's pairs does not depend on the blocking).  No
 * reference counterpart. */
int hhs_timesteps_sym_set_sinks_per_lane(int s);
/* Source unit of the FP32 symmetric kernel: 0 (default: 64 sources up to
 * 65536 particles, 256 above) or 64, 128 or 256 (64 and 128 only up to
 * 65536 particles; larger counts keep 256).  Process-wide; the workspace
 * size above covers every choice, and the timesteps are the same bits
 * whatever the unit.  No reference counterpart. */
int hhs_timesteps_sym_set_unit(int u);
int hhs_timesteps_sym_f64(const hhs_particles_f64 *sys, int n, double eta, double eps2, int mode,
                          void *ws, size_t ws_bytes, void *stream);
int hhs_timesteps_f32(const hhs_particles_f32 *sys, int n, float eta, float eps2,
                      void *stream);

/* Timestep evaluation modes for the HOLD drivers. */
#define HHS_TIMESTEP_EXACT 0 /* IEEE sqrt/divide, bit-exact with the formula   */
#define HHS_TIMESTEP_FAST 1  /* rsqrt + Newton, one divide per particle; agrees
                                with EXACT to a few ulp                         */
#define HHS_TIMESTEP_FP32 2  /* FP64 differences rounded to FP32, criterion in
                                FP32; relative error ~1e-7                      */
/* Same criterion as hhs_timesteps_f64 in the FAST form (no counterpart in
 * the reference). */
int hhs_timesteps_fast_f64(const hhs_particles_f64 *sys, int n, double eta, double eps2,
                           void *stream);
/* Same criterion in any of the three modes. */
int hhs_timesteps_mode_f64(const hhs_particles_f64 *sys, int n, double eta, double eps2,
                           int mode, void *stream);

/* ------------------------------------------------------ scan / split -- */
/* Device workspace for the single-pass scan (tile counter, epoch, tile
 * status words).  Allocate hhs_scan_workspace_bytes(n_max) bytes, call
 * hhs_scan_workspace_init once, then reuse it for any n <= n_max.  Calls
 * that share one workspace must be ordered (same stream). */
size_t hhs_scan_workspace_bytes(int n_max);
int hhs_scan_workspace_init(void *ws, size_t bytes, void *stream);
/* Library-owned, grow-only workspace of the calling thread, initialised and
 * ready for n <= n_max (for callers that, like the reference's scan
 * functions, carry no workspace).  Valid until the thread exits; calls from
 * one thread must be stream-ordered. */
int hhs_scan_workspace_thread(int n_max, void **ws, size_t *bytes);

/* scan_predicate_{small,large,large_ev} (src/split.cuh:437-565):
 * exclusive prefix sum of n ints: one pass (decoupled look-back) up to
 * 2^18 elements, reduce-then-scan above and for in-place scans (in == out). */
int hhs_scan_exclusive_i32(const int *in, int *out, int n, void *ws, size_t ws_bytes,
                           void *stream);

/* split (src/split.cuh:573-610): predicate timestep[i] <= pivot marks a
 * particle fast; fast ones go to fast[scan[i]], the others to
 * slow[i - scan[i]] (stable), copying pos, vel, m, timestep, id and
 * setting parent_id = i.  A lone fast particle stays slow
 * (src/split.cuh:321-324).  predicate / scanned (n ints each) are written
 * when non-NULL.  part_dev (device memory) receives the counts; the call
 * is asynchronous (no host synchronisation). */
int hhs_split_f64(const hhs_particles_f64 *total, hhs_particles_f64 *slow,
                  hhs_particles_f64 *fast, int *predicate, int *scanned, double pivot,
                  int n, hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream);
int hhs_split_f32(const hhs_particles_f32 *total, hhs_particles_f32 *slow,
                  hhs_particles_f32 *fast, int *predicate, int *scanned, float pivot,
                  int n, hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream);

/* hhs_split_f64 without a workspace argument and without any allocation.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The scan's workspace is a library-owned static device
 * array of 8256 bytes per GPU (a 64-byte header and one count per chunk,
 * at most 2048 chunks), which is what a split of any n needs; a larger need
 * would return HHS_ERR_WORKSPACE.  Calls nothing that allocates.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (different streams, or several host threads) race and must be ordered
 * by the caller; calls on one stream are ordered by it.  Same results as
 * hhs_split_f64, bit for bit. */
int hhs_split_f64_noalloc(const hhs_particles_f64 *total, hhs_particles_f64 *slow,
                          hhs_particles_f64 *fast, int *predicate, int *scanned, double pivot,
                          int n, hhs_partition *part_dev, void *stream);

/* Scalar-array split of the reference's scan tests
 * (src/scan_optimized_3.cuh:485-515, split.cuh naming): values <= pivot go
 * to fast, the rest to slow, both stable; no lone-particle rule. */
int hhs_split_values_f32(const float *values, float *slow, float *fast, float pivot,
                         int n, hhs_partition *part_dev, void *ws, size_t ws_bytes,
                         void *stream);
int hhs_split_values_f64(const double *values, double *slow, double *fast,
                         double pivot, int n, hhs_partition *part_dev, void *ws,
                         size_t ws_bytes, void *stream);

/* ---------------------------------------------------- drift / update -- */
/* drift (src/drift.cuh:30-43): pos += vel * (dt * fac), FP64 bit-exact. */
int hhs_drift_f64(double *pos, const double *vel, int n, double dt, double fac,
                  void *stream);
int hhs_drift_f32(float *pos, const float *vel, int n, float dt, float fac,
                  void *stream);

/* update_system (src/update_system.cuh:29-43):
 * prev[cur.parent_id[i]] <- cur[i] for pos, vel, m, timestep. */
int hhs_update_f64(hhs_particles_f64 *prev, const hhs_particles_f64 *cur, int n,
                   void *stream);
int hhs_update_f32(hhs_particles_f32 *prev, const hhs_particles_f32 *cur, int n,
                   void *stream);

/* ---------------------------------------------------------- energy -- */
/* FP64 kinetic and potential energy, E = sum 1/2 m v^2 -
 * sum_{i<j} m_i m_j / sqrt(r^2 + eps2) (no reference counterpart; the
 * diagnostic required for |dE/E|).  Deterministic for a given n.
 * Blocks the calling thread until the result is on the host.  Returns
 * HHS_ERR_NONFINITE (with both sums still written) when either sum is NaN
 * or infinite, i.e. when the state holds a NaN or Inf. */
int hhs_energy_f64(const double *pos, const double *vel, const double *m, int n,
                   double eps2, double *kinetic, double *potential, void *stream);

/* ------------------------------------------------------------- HOLD -- */
/* nbodysystem_globals<T> (src/particle_system.cuh:65-104) plus the
 * per-level buffers of nbodysystem_buffers<T> (:107-132), allocated once
 * and reused (the reference allocates them on every hold_step call and
 * never frees them, src/hold.cuh:56). */
typedef struct hhs_globals hhs_globals;

int hhs_globals_create(hhs_globals **out, int n_total, double eta, double eps,
                       int max_level);
int hhs_globals_destroy(hhs_globals *g);
/* Scratch exposed as in the reference (device pointers, n_total ints). */
int *hhs_globals_predicate(hhs_globals *g);
int *hhs_globals_scanned_predicate(hhs_globals *g);
double hhs_globals_simtime(const hhs_globals *g);
void hhs_globals_set_simtime(hhs_globals *g, double t);
double hhs_globals_eta(const hhs_globals *g);
double hhs_globals_eps(const hhs_globals *g);

typedef struct hhs_hold_stats {
    long long calls;        /* hold_step invocations                       */
    long long interactions; /* kick pair evaluations                        */
    long long ts_pairs;     /* timestep pair evaluations (ordered)          */
    int max_level;          /* deepest level reached                        */
    long long calls_per_level[64];
} hhs_hold_stats;
int hhs_globals_stats(const hhs_globals *g, hhs_hold_stats *out);
/* HHS_TIMESTEP_EXACT (default), HHS_TIMESTEP_FAST or HHS_TIMESTEP_FP32
 * for hold_step's timestep evaluations. */
int hhs_globals_set_timestep_mode(hhs_globals *g, int mode);
/* Kick arithmetic for this hold_step (an HHS_KICK_* mode), or -1 (default)
 * for the process-wide hhs_kick_set_precision setting.  HHS_KICK_F64_EXACT
 * with HHS_TIMESTEP_EXACT is the FP64-only HOLD mode: every kernel of the
 * recursion is bit-exact with the FP64 CPU HOLD, so the whole integration
 * is.  No reference counterpart (the reference has one arithmetic per T). */
int hhs_globals_set_kick_precision(hhs_globals *g, int mode);
/* Calls of hold_step with at most n_max particles run their whole
 * recursion subtree in one device launch (one workgroup; MIXED32 kicks and
 * the globals' timestep mode), with one host sync per subtree instead of
 * one per call.  0 disables it.  Used only when the kick precision in force
 * is HHS_KICK_MIXED32.  Default 256.  No reference counterpart. */
int hhs_globals_set_subtree_max(hhs_globals *g, int n_max);
int hhs_globals_subtree_max(const hhs_globals *g);
/* Speculative launches (default on): the first fast recursion of a call --
 * its split or its device subtree -- is queued before the host has read the
 * call's counts, with the set size read on the device; a launch that turns
 * out not to be the one the recursion makes does nothing.  Results are the
 * same bits either way.  No reference counterpart. */
int hhs_globals_set_speculation(hhs_globals *g, int on);
/* Kick partials applied by their readers (default on): a call whose second
 * fast recursion runs as a device subtree (or that has no fast set) leaves
 * its split-K kick's partial sums to that subtree (fast sinks) and to its
 * scatter-back (slow sinks) instead of a reduce launch.  Results are the
 * same bits either way.  No reference counterpart. */
int hhs_globals_set_fused_reduce(hhs_globals *g, int on);
/* Chained scatter-backs (default on): a call that is its parent's second
 * fast recursion leaves its scatter-back to the parent, whose own follows
 * at once, and the chain of them runs as one launch (up to 8 levels).
 * Results are the same bits either way.  No reference counterpart. */
int hhs_globals_set_chain_scatter(hhs_globals *g, int on);
int hhs_globals_timestep_mode(const hhs_globals *g);
int hhs_globals_kick_precision(const hhs_globals *g);
/* Root steps completed by hhs_evolve_f64 (restored from a snapshot on
 * resume). */
long long hhs_globals_root_steps(const hhs_globals *g);
void hhs_globals_set_root_steps(hhs_globals *g, long long k);
int hhs_globals_reset_stats(hhs_globals *g);

/* hold_step<T,BLOCKSIZE> (src/hold.cuh:51-181) on total[0..n), at
 * recursion level clevel.  Levels deeper than max_level return
 * HHS_ERR_DEPTH (the reference has no guard).  Synchronises the stream
 * once per call to read the slow/fast counts, as the reference does
 * (src/split.cuh:430). */
int hhs_hold_step_f64(hhs_globals *g, hhs_particles_f64 *total, int n, int clevel,
                      double stime, double etime, double dt, int calc_timestep,
                      void *stream);

/* hold_step<float,BLOCKSIZE> (src/hold.cuh:51-53 with T = float): the same
 * recursion on an FP32 particle set.  Timesteps in FP32 with the formula's
 * IEEE operations (the globals' timestep mode applies to FP64 sets only),
 * MIXED32 kicks on FP32 separations, FP32 drifts; host recursion only (no
 * device subtrees).  The globals' per-level buffers hold either precision,
 * one hold_step call at a time. */
int hhs_hold_step_f32(hhs_globals *g, hhs_particles_f32 *total, int n, int clevel, float stime,
                      float etime, float dt, int calc_timestep, void *stream);

/* hold<T> (declared, never defined: src/hold.cuh:20-35): root steps of
 * dt_root (timesteps recomputed at each root step) until t_end. */
int hhs_hold_f64(hhs_globals *g, hhs_particles_f64 *sys, double t_end, double dt_root,
                 void *stream);
int hhs_hold_f32(hhs_globals *g, hhs_particles_f32 *sys, float t_end, float dt_root,
                 void *stream);

/* ------------------------------------------- driver and snapshots -- */
/* Snapshot file (no reference counterpart: the reference exports state
 * only through gpu_to_host, src/particle_system.cuh:230-247, which also
 * leaves out parent_id).  Layout, little-endian:
 *   hhs_snapshot_header   128 bytes
 *   pos                   3n doubles (AoS x, y, z)
 *   vel                   3n doubles
 *   m                     n doubles
 *   id                    n int32
 * i.e. 128 + 60 n bytes.  Snapshots are taken only at root-step
 * boundaries: there every particle is at simtime, and nothing else is
 * carried from one root step to the next (timesteps are recomputed at the
 * root, the per-level buffers are scratch), so an integration resumed from
 * a snapshot continues bit for bit. */
#define HHS_SNAPSHOT_VERSION 1
typedef struct hhs_snapshot_header {
    char magic[8];          /* "HHSSNAP" and a NUL                          */
    uint32_t version;       /* HHS_SNAPSHOT_VERSION                         */
    uint32_t header_bytes;  /* sizeof(hhs_snapshot_header) = 128            */
    int64_t n;              /* particles                                    */
    int64_t root_steps;     /* root steps completed                         */
    double simtime;         /* globals simtime at the snapshot              */
    double dt_root;         /* root step of the run that wrote it           */
    double eta, eps;        /* integrator parameters                        */
    int32_t timestep_mode;  /* HHS_TIMESTEP_*                               */
    int32_t kick_precision; /* HHS_KICK_* or -1                             */
    uint64_t payload_fnv1a; /* FNV-1a 64 of the payload bytes               */
    uint8_t reserved[48];   /* zero                                         */
} hhs_snapshot_header;

/* Writes sys[0..n) (pos, vel, m; id if non-NULL, else 0..n-1) and g's
 * state to `path` after the work queued on `stream` has finished.  Writes
 * "<path>.tmp" and renames it over `path`, so an interrupted write leaves
 * the previous snapshot intact.  Blocking. */
int hhs_snapshot_write_f64(const char *path, const hhs_globals *g, const hhs_particles_f64 *sys,
                           int n, double dt_root, void *stream);
/* Reads and checks the header (magic, version, sizes against the file
 * length); HHS_ERR_FORMAT when it does not describe a whole snapshot. */
int hhs_snapshot_read_header(const char *path, hhs_snapshot_header *out);
/* Loads a snapshot into device arrays sys (sys->n >= the snapshot's n;
 * pos, vel, m required, id written when non-NULL), verifying the payload
 * checksum.  If g is non-NULL its simtime and root-step count are set from
 * the header, and its eta/eps and modes must equal the snapshot's
 * (HHS_ERR_INVALID otherwise), so that the resumed run is the same
 * integration.  `hdr` (optional) receives the header.  Blocking. */
int hhs_snapshot_read_f64(const char *path, hhs_particles_f64 *sys, hhs_globals *g,
                          hhs_snapshot_header *hdr, void *stream);

/* Top-level driver: hold<T> (declared, never defined: src/hold.cuh:20-35)
 * with root-step bookkeeping.  Runs root steps k = root_steps(g), ... of
 * dt_root over [k dt_root, (k+1) dt_root] while k dt_root < t_end,
 * recomputing the timesteps at each root step.  When snapshot_path is
 * non-NULL, writes a snapshot after every snapshot_every-th root step
 * (k+1 divisible by it; 0 = none on the way) and after the last one.
 * Resume = hhs_snapshot_read_f64 into fresh arrays and globals, then
 * hhs_evolve_f64 again with the same t_end and dt_root. */
int hhs_evolve_f64(hhs_globals *g, hhs_particles_f64 *sys, double t_end, double dt_root,
                   const char *snapshot_path, int snapshot_every, void *stream);

/* --------------------------------------------------------- ensembles -- */
/* Many independent systems integrated in one persistent launch (no
 * reference counterpart: the reference integrates one host-driven system).
 * Each system runs the same HOLD recursion as hhs_hold_f64 (root steps of
 * its dt_root until its t_end, timesteps recomputed at every root step),
 * entirely on the device: a workgroup (or, for systems of at most
 * wave_team_max_n particles, one wave) owns a system at a time.  A system's
 * result does not depend on what else is in the batch. */
typedef struct hhs_system_result {
    double e0, e1;       /* total energy before / after (FP64, softened)     */
    double t;            /* time reached                                     */
    long long calls;     /* hold_step invocations                            */
    long long interactions;
    long long ts_pairs;
    int max_level;
    int status;          /* HHS_OK, HHS_ERR_DEPTH (level or call cap), ...   */
} hhs_system_result;

typedef struct hhs_ensemble_desc {
    int n_systems;
    const int *offsets;    /* device, n_systems + 1 particle offsets         */
    double *pos, *vel;     /* device, AoS vec3, all systems back to back;
                              updated in place, original order kept          */
    const double *m;       /* device                                          */
    const double *t_end;   /* device, per system                              */
    const double *dt_root; /* device, per system                              */
    double eta, eps;
    int max_level;         /* <= 47; 0 selects 47                             */
    long long max_calls;   /* per-system cap on hold_step calls, 0 = none     */
    int wave_team_max_n;   /* 0 selects 512                                   */
    hhs_system_result *results; /* device, per system                        */
    int timestep_mode;     /* HHS_TIMESTEP_EXACT, _FAST or _FP32               */
    int *done;             /* optional, per system: set to 1 (system-scope
                              release) once that system's pos, vel and result
                              are written, while the launch still runs; with
                              pos/vel/results/done in host-mapped memory a
                              host thread can finish systems one by one      */
    int approximate_decisions; /* FP32/FAST timesteps: 0 (default) recomputes
                              exactly every timestep within the error band of
                              a pivot, so the slow/fast decisions are the
                              exact mode's, as in the per-system HOLD; 1
                              keeps the approximate values                   */
} hhs_ensemble_desc;

size_t hhs_ensemble_workspace_bytes(int n_systems, long long n_particles);
/* Asynchronous on `stream`; offsets_host is the host copy of d->offsets
 * (used to order the work queues). */
int hhs_ensemble_run_f64(const hhs_ensemble_desc *d, const int *offsets_host, void *ws,
                         size_t ws_bytes, void *stream);

/* ----------------------------------------------------------- serving -- */
/* Continuous-batching request server (no reference counterpart).  A
 * background scheduler admits requests in arrival order while their device
 * footprint fits hbm_budget_bytes; systems of at most small_max_n particles
 * run in ensemble batches (one persistent launch per batch, up to
 * batch_streams batches in flight), larger ones run hhs_hold_f64 on one of
 * big_workers streams.  Requests that arrive while others run are admitted
 * as memory and slots free up, and their work overlaps on the device.  Each
 * request's result depends only on its own data. */
typedef struct hhs_server hhs_server;

typedef struct hhs_server_config {
    double eta, eps;
    size_t hbm_budget_bytes;
    int small_max_n;    /* 0 selects 512                                    */
    int max_n;          /* largest accepted system, 0 selects 65536         */
    int max_batch;      /* systems per ensemble batch, 0 selects 4096       */
    int batch_streams;  /* ensemble batches in flight, 0 selects 2          */
    int big_workers;    /* concurrent large systems, 0 selects 4            */
    int max_level;      /* 0 selects 40                                     */
    int timestep_mode;  /* HHS_TIMESTEP_EXACT, _FAST or _FP32                */
} hhs_server_config;

typedef struct hhs_request_info {
    double t_submit, t_admit, t_start, t_done; /* seconds on the server clock */
    hhs_system_result result;
    int path;  /* 0: ensemble batch, 1: per-system HOLD                     */
    int state; /* 0 pending, 1 admitted, 2 running, 3 done, 4 failed        */
} hhs_request_info;

typedef struct hhs_server_stats_t {
    double now;
    size_t used_bytes, peak_bytes;
    int pending, running_batches, running_big;
} hhs_server_stats_t;

int hhs_server_create(hhs_server **out, const hhs_server_config *cfg);
/* Finishes every submitted request, then stops. */
int hhs_server_destroy(hhs_server *s);
/* Copies the host arrays (pos, vel: 3n; m: n); HHS_ERR_BUDGET if the
 * request alone exceeds the budget. */
int hhs_server_submit(hhs_server *s, int n, const double *pos, const double *vel,
                      const double *m, double t_end, double dt_root, long long *ticket);
int hhs_server_poll(hhs_server *s, long long ticket, hhs_request_info *info);
/* Blocks until the request is done; returns its status. */
int hhs_server_wait(hhs_server *s, long long ticket, hhs_request_info *info);
/* Final pos/vel (original particle order) of a finished request. */
int hhs_server_fetch(hhs_server *s, long long ticket, double *pos_out, double *vel_out);
/* Forget a finished request. */
int hhs_server_release(hhs_server *s, long long ticket);
int hhs_server_stats(hhs_server *s, hhs_server_stats_t *out);

/* ------------------------------------------- initial conditions (host) -- */
/* Host-only (no GPU use).  Plummer sphere in Henon units (G = M = 1,
 * E = -1/4; Aarseth, Henon & Wielen 1974), SplitMix64 stream `seed`,
 * radii beyond 10 scale lengths rejected, centre of mass at rest at the
 * origin.  mass_model 0: equal masses 1/n; 1: Salpeter (slope -2.35,
 * m in [0.1, 100], normalised to total mass 1).  virialize != 0 rescales
 * velocities to virial ratio 1/2 with the exact O(n^2) potential (eps2).
 * pos, vel: 3n doubles; m: n doubles. */
int hhs_ic_plummer(int n, uint64_t seed, int mass_model, int virialize, double eps2,
                   double *pos, double *vel, double *m);

#ifdef __cplusplus
}
#endif
#endif /* HHS_H */
