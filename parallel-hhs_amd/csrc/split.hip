// Single-pass exclusive scan and fused slow/fast split for gfx950.
//
// Reference: _get_predicate, the Blelloch scans, add and _set_predicate
// (src/split.cuh:107-610).  The reference runs 3-7 launches, 2 cudaMalloc
// and 2 cudaFree per level plus a blocking cudaDeviceSynchronize
// (src/split.cuh:430, :513-514, :563-564).  Here one kernel does
// predicate + scan + stable scatter + counts, using decoupled look-back
// (Merrill & Garland 2016, "Single-pass Parallel Prefix Scan with
// Decoupled Look-back"):
//
//  * a workgroup takes the next tile index from an atomic counter, so tiles
//    are numbered in the order they start and a tile only ever waits on
//    tiles that are already running (forward progress);
//  * the tile is loaded striped (element base + k*256 + t for item k), so
//    consecutive lanes touch consecutive elements (coalesced HBM access),
//    and scanned in index order with wave ballots / shuffles plus LDS;
//  * the tile publishes its aggregate, walks back over predecessors'
//    64-bit status words (epoch | flag | value, agent-scope acquire loads),
//    then publishes its inclusive prefix;
//  * status words carry a 30-bit epoch so the workspace needs no clearing
//    between calls; the last tile resets the counter and bumps the epoch.
//
// The lone-fast-particle rule (src/split.cuh:321-324) needs the final count,
// which only the last tile knows; a follow-up kernel rewrites the slow set
// as a plain copy when the raw fast count is 1 (it exits at once otherwise).
#include "hhs_internal.h"

namespace hhs {
namespace {

constexpr int kBlock = 256;  // threads per tile
constexpr int kWaves = kBlock / 64;

constexpr uint32_t kFlagAggregate = 1u;
constexpr uint32_t kFlagPrefix = 2u;
constexpr uint32_t kEpochMask = 0x3FFFFFFFu;

struct ScanWs {
    uint32_t *header;  // [0] tile counter, [1] epoch, [2] raw fast count
    uint64_t *status;  // one word per tile
};

__device__ __forceinline__ uint64_t make_status(uint32_t epoch, uint32_t flag, uint32_t v) {
    return (static_cast<uint64_t>(epoch & kEpochMask) << 34) |
           (static_cast<uint64_t>(flag) << 32) | static_cast<uint64_t>(v);
}

__device__ __forceinline__ void publish(uint64_t *p, uint64_t w) {
    __hip_atomic_store(p, w, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Exclusive prefix of the tiles before `tile` (single lane).
__device__ int look_back(const uint64_t *status, int tile, uint32_t epoch) {
    int prefix = 0;
    for (int k = tile - 1; k >= 0; --k) {
        uint64_t w;
        for (;;) {
            w = __hip_atomic_load(&status[k], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t e = static_cast<uint32_t>(w >> 34);
            const uint32_t f = static_cast<uint32_t>(w >> 32) & 3u;
            if (e == (epoch & kEpochMask) && f != 0u) break;
            __builtin_amdgcn_s_sleep(1);
        }
        prefix += static_cast<int>(static_cast<uint32_t>(w));
        if (((w >> 32) & 3u) == kFlagPrefix) break;
    }
    return prefix;
}

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ int wave_inclusive(int x, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

// Generic tile scan; Op supplies value(e) (an int), emit(e, v, excl) and
// finish(total) (called once, by the last tile, with the grand total).
template <int ITEMS, bool BINARY, class Op>
__global__ __launch_bounds__(kBlock) void scan_tiles_kernel(Op op, int n, ScanWs ws,
                                                            int ntiles) {
    __shared__ int s_tile;
    __shared__ uint32_t s_epoch;
    __shared__ int s_wave_tot[ITEMS][kWaves];
    __shared__ int s_prefix;

    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    if (t == 0) {
        const uint32_t ep = __hip_atomic_load(&ws.header[1], __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t id = __hip_atomic_fetch_add(&ws.header[0], 1u, __ATOMIC_ACQ_REL,
                                                   __HIP_MEMORY_SCOPE_AGENT);
        s_tile = static_cast<int>(id);
        s_epoch = ep;
    }
    __syncthreads();
    const int tile = s_tile;
    const uint32_t epoch = s_epoch;
    const int base = tile * (kBlock * ITEMS);

    int v[ITEMS], incl[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const int e = base + k * kBlock + t;
        v[k] = e < n ? op.value(e) : 0;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        int x;
        if constexpr (BINARY) {
            const uint64_t b = __ballot(v[k] != 0);
            const uint64_t le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
            x = __popcll(b & le);
        } else {
            x = wave_inclusive(v[k], lane);
        }
        incl[k] = x;
        if (lane == 63) s_wave_tot[k][wave] = x;
    }
    __syncthreads();
    int running = 0;
    int excl[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        int before = 0, slot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const int c = s_wave_tot[k][w];
            before += (w < wave) ? c : 0;
            slot += c;
        }
        excl[k] = running + before + incl[k] - v[k];
        running += slot;
    }
    const int tile_total = running;

    if (t == 0) {
        int prefix = 0;
        if (tile == 0) {
            publish(&ws.status[0], make_status(epoch, kFlagPrefix, static_cast<uint32_t>(tile_total)));
        } else {
            publish(&ws.status[tile], make_status(epoch, kFlagAggregate, static_cast<uint32_t>(tile_total)));
            prefix = look_back(ws.status, tile, epoch);
            publish(&ws.status[tile],
                    make_status(epoch, kFlagPrefix, static_cast<uint32_t>(prefix + tile_total)));
        }
        s_prefix = prefix;
        if (tile == ntiles - 1) {
            op.finish(prefix + tile_total, ws.header);
            // Every tile took its index before this one did, and read the
            // epoch before that: safe to recycle the counter and epoch.
            __hip_atomic_store(&ws.header[0], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&ws.header[1], (epoch + 1u) & kEpochMask, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    __syncthreads();
    const int prefix = s_prefix;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
        const int e = base + k * kBlock + t;
        if (e < n) op.emit(e, v[k], prefix + excl[k]);
    }
}

// ------------------------------------------------------------------ ops

struct ScanI32Op {
    const int *in;
    int *out;
    __device__ int value(int e) const { return in[e]; }
    __device__ void emit(int e, int, int excl) const { out[e] = excl; }
    __device__ void finish(int, uint32_t *) const {}
};

template <typename T> struct SplitOp {
    Particles<T> total, slow, fast;
    int *predicate, *scanned;
    T pivot;
    int n;
    hhs_partition *part;
    // src/split.cuh:115 -- NaN compares false and counts as slow.
    __device__ int value(int e) const { return total.timestep[e] <= pivot ? 1 : 0; }
    __device__ void emit(int e, int p, int f) const {
        if (predicate) predicate[e] = p;
        if (scanned) scanned[e] = f;
        const Particles<T> &d = p ? fast : slow;
        const int o = p ? f : e - f;  // src/split.cuh:336,350
        d.pos[3 * o + 0] = total.pos[3 * e + 0];
        d.pos[3 * o + 1] = total.pos[3 * e + 1];
        d.pos[3 * o + 2] = total.pos[3 * e + 2];
        d.vel[3 * o + 0] = total.vel[3 * e + 0];
        d.vel[3 * o + 1] = total.vel[3 * e + 1];
        d.vel[3 * o + 2] = total.vel[3 * e + 2];
        d.m[o] = total.m[e];
        d.timestep[o] = total.timestep[e];
        if (d.id) d.id[o] = total.id ? total.id[e] : e;
        d.parent_id[o] = e;
    }
    __device__ void finish(int nf, uint32_t *header) const {
        header[2] = static_cast<uint32_t>(nf);
        const int reported = (nf == 1) ? 0 : nf;  // src/split.cuh:321-324
        part->N_fast = reported;
        part->N_slow = n - reported;
    }
};

template <typename T> struct SplitValuesOp {
    const T *values;
    T *slow, *fast;
    T pivot;
    int n;
    hhs_partition *part;
    __device__ int value(int e) const { return values[e] <= pivot ? 1 : 0; }
    __device__ void emit(int e, int p, int f) const {
        if (p)
            fast[f] = values[e];
        else
            slow[e - f] = values[e];
    }
    __device__ void finish(int nf, uint32_t *header) const {
        header[2] = static_cast<uint32_t>(nf);
        part->N_fast = nf;
        part->N_slow = n - nf;
    }
};

// Lone fast particle: the whole set stays slow, in order (the intended
// behaviour of src/split.cuh:321-324; the reference's :361-368 races).
template <typename T>
__global__ __launch_bounds__(256) void lone_fast_fixup_kernel(Particles<T> total,
                                                              Particles<T> slow, int n,
                                                              const uint32_t *header) {
    if (header[2] != 1u) return;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    slow.pos[3 * e + 0] = total.pos[3 * e + 0];
    slow.pos[3 * e + 1] = total.pos[3 * e + 1];
    slow.pos[3 * e + 2] = total.pos[3 * e + 2];
    slow.vel[3 * e + 0] = total.vel[3 * e + 0];
    slow.vel[3 * e + 1] = total.vel[3 * e + 1];
    slow.vel[3 * e + 2] = total.vel[3 * e + 2];
    slow.m[e] = total.m[e];
    slow.timestep[e] = total.timestep[e];
    if (slow.id) slow.id[e] = total.id ? total.id[e] : e;
    slow.parent_id[e] = e;
}

__global__ void zero_partition_kernel(hhs_partition *part) {
    part->N_fast = 0;
    part->N_slow = 0;
}

__global__ void fill_u32_kernel(uint32_t *p, size_t words, uint32_t v) {
    const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (i < words) p[i] = v;
}

int items_for(int n) {
    // Small inputs: short tiles so several workgroups share the work; large
    // inputs: 4 items per lane to amortise the look-back.
    if (n <= 64 * 1024) return 1;
    if (n <= 512 * 1024) return 2;
    return 4;
}

template <bool BINARY, class Op>
int launch_scan_op(const Op &op, int n, void *ws_ptr, size_t ws_bytes, hipStream_t st) {
    const int items = items_for(n);
    const int tile = kBlock * items;
    const int ntiles = (n + tile - 1) / tile;
    HHS_REQUIRE(ws_ptr, "scan: null workspace");
    if (ws_bytes < kScanHeaderBytes + sizeof(uint64_t) * static_cast<size_t>(ntiles)) {
        set_error("scan: workspace of %zu bytes too small for n=%d (need %zu)", ws_bytes, n,
                  scan_ws_bytes(n));
        return HHS_ERR_WORKSPACE;
    }
    ScanWs ws{static_cast<uint32_t *>(ws_ptr),
              reinterpret_cast<uint64_t *>(static_cast<char *>(ws_ptr) + kScanHeaderBytes)};
    switch (items) {
    case 1:
        hipLaunchKernelGGL((scan_tiles_kernel<1, BINARY, Op>), dim3(ntiles), dim3(kBlock), 0, st,
                           op, n, ws, ntiles);
        break;
    case 2:
        hipLaunchKernelGGL((scan_tiles_kernel<2, BINARY, Op>), dim3(ntiles), dim3(kBlock), 0, st,
                           op, n, ws, ntiles);
        break;
    default:
        hipLaunchKernelGGL((scan_tiles_kernel<4, BINARY, Op>), dim3(ntiles), dim3(kBlock), 0, st,
                           op, n, ws, ntiles);
        break;
    }
    HHS_LAUNCH_CHECK("scan_tiles_kernel");
    return HHS_OK;
}

}  // namespace

size_t scan_ws_bytes(int n_max) {
    const size_t tiles = n_max > 0 ? (static_cast<size_t>(n_max) + kBlock - 1) / kBlock : 1;
    return kScanHeaderBytes + sizeof(uint64_t) * tiles;
}

int scan_ws_init(void *ws, size_t bytes, hipStream_t st) {
    HHS_REQUIRE(ws && bytes >= kScanHeaderBytes, "scan_ws_init: workspace too small");
    const size_t words = bytes / 4;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((words + 255) / 256), dim3(256), 0, st,
                       static_cast<uint32_t *>(ws), words, 0u);
    HHS_LAUNCH_CHECK("fill_u32_kernel");
    return HHS_OK;
}

template <typename T>
int launch_split(const Particles<T> &total, const Particles<T> &slow, const Particles<T> &fast,
                 int *predicate, int *scanned, T pivot, int n, hhs_partition *part_dev,
                 void *ws, size_t ws_bytes, hipStream_t st) {
    HHS_REQUIRE(n >= 0, "split: negative n");
    HHS_REQUIRE(part_dev, "split: null partition");
    if (n == 0) {
        hipLaunchKernelGGL(zero_partition_kernel, dim3(1), dim3(1), 0, st, part_dev);
        HHS_LAUNCH_CHECK("zero_partition_kernel");
        return HHS_OK;
    }
    HHS_REQUIRE(total.pos && total.vel && total.m && total.timestep,
                "split: null field in total");
    HHS_REQUIRE(slow.pos && slow.vel && slow.m && slow.timestep && slow.parent_id,
                "split: null field in slow");
    HHS_REQUIRE(fast.pos && fast.vel && fast.m && fast.timestep && fast.parent_id,
                "split: null field in fast");
    SplitOp<T> op{total, slow, fast, predicate, scanned, pivot, n, part_dev};
    int rc = launch_scan_op<true>(op, n, ws, ws_bytes, st);
    if (rc) return rc;
    hipLaunchKernelGGL((lone_fast_fixup_kernel<T>), dim3((n + 255) / 256), dim3(256), 0, st,
                       total, slow, n, static_cast<const uint32_t *>(ws));
    HHS_LAUNCH_CHECK("lone_fast_fixup_kernel");
    return HHS_OK;
}

template int launch_split<double>(const Particles<double> &, const Particles<double> &,
                                  const Particles<double> &, int *, int *, double, int,
                                  hhs_partition *, void *, size_t, hipStream_t);
template int launch_split<float>(const Particles<float> &, const Particles<float> &,
                                 const Particles<float> &, int *, int *, float, int,
                                 hhs_partition *, void *, size_t, hipStream_t);

}  // namespace hhs

using namespace hhs;

template <typename T, typename P>
static int split_c(const P *total, P *slow, P *fast, int *predicate, int *scanned, T pivot, int n,
                   hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream) {
    HHS_REQUIRE(total && slow && fast, "split: null system");
    HHS_REQUIRE(n >= 0 && n <= total->n, "split: n=%d outside [0, %d]", n, total->n);
    HHS_REQUIRE(slow->n >= n && fast->n >= n, "split: slow/fast capacity below n=%d", n);
    return launch_split<T>(from_c(*total), from_c(*slow), from_c(*fast), predicate, scanned,
                           pivot, n, part_dev, ws, ws_bytes, as_stream(stream));
}

template <typename T>
static int split_values(const T *values, T *slow, T *fast, T pivot, int n,
                        hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream) {
    HHS_REQUIRE(n >= 0, "split_values: negative n");
    HHS_REQUIRE(part_dev, "split_values: null partition");
    hipStream_t st = as_stream(stream);
    if (n == 0) {
        hipLaunchKernelGGL(zero_partition_kernel, dim3(1), dim3(1), 0, st, part_dev);
        HHS_LAUNCH_CHECK("zero_partition_kernel");
        return HHS_OK;
    }
    HHS_REQUIRE(values && slow && fast, "split_values: null array");
    SplitValuesOp<T> op{values, slow, fast, pivot, n, part_dev};
    return launch_scan_op<true>(op, n, ws, ws_bytes, st);
}

extern "C" {

size_t hhs_scan_workspace_bytes(int n_max) { return scan_ws_bytes(n_max); }

int hhs_scan_workspace_init(void *ws, size_t bytes, void *stream) {
    return scan_ws_init(ws, bytes, as_stream(stream));
}

int hhs_scan_workspace_thread(int n_max, void **ws, size_t *bytes) {
    struct ThreadWs {
        void *ptr = nullptr;
        size_t bytes = 0;
        ~ThreadWs() {
            if (ptr) (void)hipFree(ptr);
        }
    };
    static thread_local ThreadWs tw;
    HHS_REQUIRE(ws && bytes, "scan_workspace_thread: null output");
    const size_t need = scan_ws_bytes(n_max);
    if (need > tw.bytes) {
        // Wait for work queued on the old buffer before replacing it.
        HHS_HIP(hipDeviceSynchronize());
        if (tw.ptr) HHS_HIP(hipFree(tw.ptr));
        tw.ptr = nullptr;
        tw.bytes = 0;
        HHS_HIP(hipMalloc(&tw.ptr, need));
        HHS_HIP(hipMemset(tw.ptr, 0, need));
        tw.bytes = need;
    }
    *ws = tw.ptr;
    *bytes = tw.bytes;
    return HHS_OK;
}

int hhs_scan_exclusive_i32(const int *in, int *out, int n, void *ws, size_t ws_bytes,
                           void *stream) {
    HHS_REQUIRE(n >= 0, "scan: negative n");
    if (n == 0) return HHS_OK;
    HHS_REQUIRE(in && out, "scan: null array");
    ScanI32Op op{in, out};
    return launch_scan_op<false>(op, n, ws, ws_bytes, as_stream(stream));
}

int hhs_split_f64(const hhs_particles_f64 *total, hhs_particles_f64 *slow,
                  hhs_particles_f64 *fast, int *predicate, int *scanned, double pivot, int n,
                  hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream) {
    return split_c<double>(total, slow, fast, predicate, scanned, pivot, n, part_dev, ws,
                           ws_bytes, stream);
}

int hhs_split_f32(const hhs_particles_f32 *total, hhs_particles_f32 *slow,
                  hhs_particles_f32 *fast, int *predicate, int *scanned, float pivot, int n,
                  hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream) {
    return split_c<float>(total, slow, fast, predicate, scanned, pivot, n, part_dev, ws,
                          ws_bytes, stream);
}

int hhs_split_values_f32(const float *values, float *slow, float *fast, float pivot, int n,
                         hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream) {
    return split_values<float>(values, slow, fast, pivot, n, part_dev, ws, ws_bytes, stream);
}

int hhs_split_values_f64(const double *values, double *slow, double *fast, double pivot, int n,
                         hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream) {
    return split_values<double>(values, slow, fast, pivot, n, part_dev, ws, ws_bytes, stream);
}

}  // extern "C"
