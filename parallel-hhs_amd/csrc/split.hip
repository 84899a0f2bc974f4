// Exclusive scan and fused slow/fast split for gfx950.
//
// Reference: _get_predicate, the Blelloch scans, add and _set_predicate
// (src/split.cuh:107-610).  The reference runs 3-7 launches, 2 cudaMalloc
// and 2 cudaFree per level plus a blocking cudaDeviceSynchronize
// (src/split.cuh:430, :513-514, :563-564).  Here the predicate, the scan,
// the stable scatter and the counts are fused, in one of two forms:
//
//  * up to kLookbackMaxN elements, a single-pass decoupled look-back scan
//    (scan_lookback_kernel): one launch, one read of the input; every wait
//    for a predecessor is bounded, with a fallback that sums the
//    predecessor's input;
//  * above, a reduce-then-scan over large contiguous chunks (count_kernel:
//    workgroup g sums the predicates of chunk g; scan_chunks_kernel:
//    workgroup g adds up counts[0..g) for its offset and walks its chunk in
//    tiles of kBlock * kItems elements in index order -- striped loads,
//    wave ballots / shuffles plus LDS for the tile scan -- emitting every
//    element with its exclusive prefix; the last chunk reports the grand
//    total).  No inter-workgroup waiting at all; one extra read of the
//    predicate source.
//
// The lone-fast-particle rule (src/split.cuh:321-324) needs the final count,
// which only the last tile or chunk knows; a follow-up kernel rewrites the
// slow set as a plain copy when the raw fast count is 1 (it exits at once
// otherwise).
#include <mutex>
#include <type_traits>

#include "hhs_internal.h"
#include "pair_math.h"

namespace hhs {
namespace {

constexpr int kBlock = 256;  // threads per workgroup
constexpr int kWaves = kBlock / 64;
constexpr int kItems = 4;    // elements per lane per tile
constexpr int kTile = kBlock * kItems;
constexpr int kMaxChunks = 2048;

// Workspace: u32 header[16] ([2] = raw fast count of the last split), then
// one int count per chunk.
struct ScanWs {
    uint32_t *header;
    int *counts;
};

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ int wave_inclusive(int x, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    return x;
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
    return x;
}

// Sum of x over the workgroup (every thread gets it).
__device__ __forceinline__ int block_sum(int x, int *s_red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    x = wave_sum(x);
    if (lane == 0) s_red[wave] = x;
    __syncthreads();
    int t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += s_red[w];
    __syncthreads();
    return t;
}

// Decision guard of the HOLD split (pair_math.h) on one tile: thread t holds
// the timesteps ts[k] of elements base + k * kBlock + t (valid below end) of
// the call's set.  Every timestep within the band of a pivot dt * 2^-k is
// recomputed with the exact formula over all n particles of the set (the
// sources of its timestep evaluation), the workgroup's threads splitting the
// sources, and replaced in its owner's register -- before any predicate is
// formed from it -- and, with write_back, in set.timestep.  The common case
// (nothing flagged) costs one compare per element and one barrier.  Called
// by every thread of the workgroup.
struct SplitGuard {
    TsGuard g;
    double eta, c_ff, eps2;  // the timestep criterion's parameters
};

__device__ void guard_tile(const SplitGuard &gd, const Particles<double> &set, int n, int base,
                           int end, double (&ts)[kItems], bool write_back) {
    __shared__ int s_list[kTile];
    __shared__ double s_res[kTile];
    __shared__ int s_cnt;
    __shared__ double s_min[kWaves];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    bool fl[kItems];
    bool any = false;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        fl[k] = base + k * kBlock + t < end && guard_near_pivot(gd.g, ts[k]);
        any = any || fl[k];
    }
    if (!__syncthreads_or(any)) return;
    if (t == 0) s_cnt = 0;
    __syncthreads();
    int slot[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        slot[k] = fl[k] ? atomicAdd(&s_cnt, 1) : -1;
        if (fl[k]) s_list[slot[k]] = base + k * kBlock + t;
    }
    __syncthreads();
    const int cnt = s_cnt;
    for (int q = 0; q < cnt; ++q) {
        const int s = s_list[q];
        const double xi = set.pos[3 * s], yi = set.pos[3 * s + 1], zi = set.pos[3 * s + 2];
        const double vxi = set.vel[3 * s], vyi = set.vel[3 * s + 1], vzi = set.vel[3 * s + 2];
        const double mi = set.m[s];
        double best = TsConst<double>::huge;
        for (int j = t; j < n; j += kBlock)
            pair_timestep<double>(xi, yi, zi, vxi, vyi, vzi, mi, set.pos[3 * j],
                                  set.pos[3 * j + 1], set.pos[3 * j + 2], set.vel[3 * j],
                                  set.vel[3 * j + 1], set.vel[3 * j + 2], set.m[j], gd.eta,
                                  gd.c_ff, gd.eps2, j != s, best);
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const double o = __shfl_xor(best, d, 64);
            best = o < best ? o : best;
        }
        if (lane == 0) s_min[wave] = best;
        __syncthreads();
        if (t == 0) {
            double m = s_min[0];
#pragma unroll
            for (int w = 1; w < kWaves; ++w) m = s_min[w] < m ? s_min[w] : m;
            s_res[q] = m;
            if (write_back) set.timestep[s] = m;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k)
        if (fl[k]) ts[k] = s_res[slot[k]];
}

template <class Op>
__global__ __launch_bounds__(kBlock) void count_kernel(Op op, int n, int per_chunk,
                                                       int *counts) {
    __shared__ int s_red[kWaves];
    const int b = blockIdx.x * per_chunk;
    const int e_end = min(b + per_chunk, n);
    int c = 0;
    bool guarded = false;
    if constexpr (Op::kGuard) {
        if (op.guard_on()) {  // tiles of kItems per thread, the guard on each
            guarded = true;
            for (int base = b; base < e_end; base += kTile) {
                double ts[kItems];
#pragma unroll
                for (int k = 0; k < kItems; ++k) {
                    const int e = base + k * kBlock + static_cast<int>(threadIdx.x);
                    ts[k] = e < e_end ? op.total.timestep[e] : 0.0;
                }
                guard_tile(op.gd, op.total, n, base, e_end, ts, true);
#pragma unroll
                for (int k = 0; k < kItems; ++k) {
                    const int e = base + k * kBlock + static_cast<int>(threadIdx.x);
                    c += (e < e_end && ts[k] <= op.pivot) ? 1 : 0;
                }
            }
        }
    }
    if (!guarded) {
#pragma unroll 8
        for (int e = b + static_cast<int>(threadIdx.x); e < e_end; e += kBlock)
            c += op.value_at(e);
    }
    c = block_sum(c, s_red);
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// Op supplies value_at(e) (the int summed), an Item type with load(e) (every
// field the element needs, so all of a tile's loads are in flight before any
// store), value(item), emit(e, item, v, excl), and finish(total, n, header)
// (called once, by the last chunk, with the grand total and the count).
template <bool BINARY, class Op>
__global__ __launch_bounds__(kBlock) void scan_chunks_kernel(Op op, int n, int per_chunk,
                                                             ScanWs ws) {
    __shared__ int s_red[kWaves];
    __shared__ int s_wave_tot[kItems][kWaves];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int g = blockIdx.x;

    int offset = 0;
    for (int j = t; j < g; j += kBlock) offset += ws.counts[j];
    offset = block_sum(offset, s_red);

    const int c_end = min((g + 1) * per_chunk, n);
    for (int base = g * per_chunk; base < c_end; base += kTile) {
        typename Op::Item it[kItems];
        int v[kItems], incl[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const int e = base + k * kBlock + t;
            if (e < c_end) it[k] = op.load(e);
        }
        // a single chunk has no count pass: the guard runs on the loaded
        // records (the emitted timesteps are the refined ones)
        if constexpr (Op::kGuard) {
            if (gridDim.x == 1 && op.guard_on()) {
                double ts[kItems];
#pragma unroll
                for (int k = 0; k < kItems; ++k) ts[k] = base + k * kBlock + t < c_end ? it[k].ts : 0.0;
                guard_tile(op.gd, op.total, n, base, c_end, ts, false);
#pragma unroll
                for (int k = 0; k < kItems; ++k) it[k].ts = ts[k];
            }
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const int e = base + k * kBlock + t;
            v[k] = e < c_end ? op.value(it[k]) : 0;
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            int x;
            if constexpr (BINARY) {
                const uint64_t bal = __ballot(v[k] != 0);
                const uint64_t le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
                x = __popcll(bal & le);
            } else {
                x = wave_inclusive(v[k], lane);
            }
            incl[k] = x;
            if (lane == 63) s_wave_tot[k][wave] = x;
        }
        __syncthreads();
        int running = 0;
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            int before = 0, slot = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const int c = s_wave_tot[k][w];
                before += (w < wave) ? c : 0;
                slot += c;
            }
            const int e = base + k * kBlock + t;
            if (e < c_end) op.emit(e, it[k], v[k], offset + running + before + incl[k] - v[k]);
            running += slot;
        }
        offset += running;
        __syncthreads();  // s_wave_tot is rewritten by the next tile
    }
    if (g == static_cast<int>(gridDim.x) - 1 && t == 0) op.finish(offset, n, ws.header);
}

// ------------------------------------------- single-pass look-back scan
// Merrill & Garland's decoupled look-back (2016): tile t (kTile elements,
// blockIdx order) loads its items once, forms its aggregate, publishes it,
// reads the status words of its predecessors back to the nearest inclusive
// prefix, publishes its own inclusive prefix and emits its items -- one
// launch and one read of the input.  Status word of a tile (64 bits): the
// call's epoch (31 bits), a prefix flag, the 32-bit value (the aggregate,
// or with the flag the inclusive prefix), stored and loaded with agent-scope
// atomics.  A word of another epoch is "not yet published".
//
// Every wait is bounded: wave 0 polls a window of 64 predecessors (one per
// lane) at most kSpinLimit times; if a needed predecessor has still not
// published by then, the whole workgroup computes that tile's aggregate from
// the input itself (with the decision guard when the op has one, so the value
// is the one its owner publishes) and the look-back continues below it.  No
// workgroup ever waits on another without a bound.  (Workgroups are
// dispatched in blockIdx order, so a predecessor has started when its
// successor runs and the polls normally end within microseconds.)
//
// The epoch lives in the workspace header: every tile reads it at its start
// and uses epoch + 1; the last tile to finish (a device-scope counter) stores
// the new epoch and resets the counter, so the next call -- stream-ordered
// after this one -- sees fresh status.  Before the 31-bit epoch would wrap,
// the last tile clears every status word, so a stale word never matches.
//
// On MI355X it wins while the input is a few hundred tiles (one launch
// instead of two, no second read): n = 131072 scan 11.7 -> 7.4 us, record
// split 18.9 -> 16.8 us; n = 262144 split 18.6 -> 14.9 us.  Above that it
// loses: at 2^24 (16384 tiles, ~2000 resident at once) most tiles find
// only aggregates in their window and walk back window by window, each poll
// a cross-XCD atomic round trip, so the prefix frontier serialises: scan
// 41 -> 637 us, split 0.54 -> 1.02 ms.  Larger inputs therefore take the
// reduce-then-scan (kLookbackMaxN).
constexpr int kLookbackMaxN = 1 << 18;
constexpr int kHdrEpoch = 0, kHdrDone = 1;  // header words; [2] raw fast count
constexpr uint32_t kEpochMax = (1u << 31) - 1;
constexpr int kSpinLimit = 4096;

__device__ __forceinline__ uint64_t lb_word(uint32_t epoch, bool prefix, int value) {
    return (static_cast<uint64_t>(epoch) << 33) | (static_cast<uint64_t>(prefix) << 32) |
           static_cast<uint32_t>(value);
}
__device__ __forceinline__ uint32_t lb_epoch(uint64_t w) { return static_cast<uint32_t>(w >> 33); }
__device__ __forceinline__ bool lb_prefix(uint64_t w) { return (w >> 32) & 1ull; }
__device__ __forceinline__ int lb_value(uint64_t w) {
    return static_cast<int>(static_cast<uint32_t>(w));
}
__device__ __forceinline__ void lb_store(uint64_t *p, uint64_t w) {
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t lb_load(uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum of the values of tile k (elements [k kTile, min((k+1) kTile, n))) over
// the workgroup, with the decision guard applied as the tile's owner applies
// it; every thread gets the sum.  The fallback of the look-back.
template <class Op>
__device__ int tile_value_sum(const Op &op, int n, int k, int *s_red) {
    const int base = k * kTile, end = min(base + kTile, n);
    const int t = threadIdx.x;
    int c = 0;
    if constexpr (Op::kGuard) {
        if (op.guard_on()) {
            double ts[kItems];
#pragma unroll
            for (int q = 0; q < kItems; ++q) {
                const int e = base + q * kBlock + t;
                ts[q] = e < end ? op.total.timestep[e] : 0.0;
            }
            guard_tile(op.gd, op.total, n, base, end, ts, false);
#pragma unroll
            for (int q = 0; q < kItems; ++q) {
                const int e = base + q * kBlock + t;
                c += (e < end && ts[q] <= op.pivot) ? 1 : 0;
            }
            return block_sum(c, s_red);
        }
    }
#pragma unroll
    for (int q = 0; q < kItems; ++q) {
        const int e = base + q * kBlock + t;
        c += e < end ? op.value_at(e) : 0;
    }
    return block_sum(c, s_red);
}

// status: one word per tile after the 64-byte header; status_cap words fit
// the workspace (>= tiles).  n_dev non-null: the element count is read from
// device memory (a count an earlier kernel on the stream wrote), the grid
// covers an upper bound of it, and the launch does nothing at all unless
// the count exceeds n_min; workgroups past the count's tiles exit at once
// and take no part in the protocol.
template <bool BINARY, class Op>
__global__ __launch_bounds__(kBlock) void scan_lookback_kernel(Op op, int n, uint32_t *header,
                                                               uint64_t *status, int tiles,
                                                               int status_cap,
                                                               const int *n_dev, int n_min) {
    __shared__ int s_red[kWaves];
    __shared__ int s_wave_tot[kItems][kWaves];
    __shared__ uint32_t s_epoch;
    __shared__ int s_k, s_acc, s_state, s_last;
    const int t = threadIdx.x;
    const int lane = t & 63;
    const int wave = t >> 6;
    const int tile = blockIdx.x;
    if (n_dev) {  // n is the bound the grid covers
        const int nd = __builtin_amdgcn_readfirstlane(*n_dev);
        if (nd <= n_min || nd > n) return;
        n = nd;
        tiles = (n + kTile - 1) / kTile;
        if (tile >= tiles) return;
    }
    if (t == 0)
        s_epoch = __hip_atomic_load(header + kHdrEpoch, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT) + 1u;
    const int base = tile * kTile;
    const int c_end = min(base + kTile, n);
    typename Op::Item it[kItems];
    int v[kItems], incl[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int e = base + k * kBlock + t;
        if (e < c_end) it[k] = op.load(e);
    }
    if constexpr (Op::kGuard) {
        if (op.guard_on()) {  // the guard on the loaded records (uniform branch)
            double ts[kItems];
#pragma unroll
            for (int k = 0; k < kItems; ++k) ts[k] = base + k * kBlock + t < c_end ? it[k].ts : 0.0;
            guard_tile(op.gd, op.total, n, base, c_end, ts, false);
#pragma unroll
            for (int k = 0; k < kItems; ++k) it[k].ts = ts[k];
        }
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int e = base + k * kBlock + t;
        v[k] = e < c_end ? op.value(it[k]) : 0;
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        int x;
        if constexpr (BINARY) {
            const uint64_t bal = __ballot(v[k] != 0);
            const uint64_t le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
            x = __popcll(bal & le);
        } else {
            x = wave_inclusive(v[k], lane);
        }
        incl[k] = x;
        if (lane == 63) s_wave_tot[k][wave] = x;
    }
    __syncthreads();
    const uint32_t ep = s_epoch;
    // in-tile exclusive offsets and the tile's aggregate
    int local[kItems];
    int running = 0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        int before = 0, slot = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const int c = s_wave_tot[k][w];
            before += (w < wave) ? c : 0;
            slot += c;
        }
        local[k] = running + before + incl[k] - v[k];
        running += slot;
    }
    const int agg = running;
    int excl = 0;
    if (tile == 0) {
        if (t == 0) lb_store(status, lb_word(ep, true, agg));
    } else {
        if (t == 0) {
            lb_store(status + tile, lb_word(ep, false, agg));
            s_k = tile - 1;
            s_acc = 0;
        }
        __syncthreads();
        // Rounds: wave 0 resolves one window (or gives up on it, bounded);
        // every thread reads the round's outcome into registers and passes a
        // second barrier before wave 0 may write the next round's.
        for (;;) {
            // wave 0: one window of predecessors k, k-1, ..., k-63
            if (wave == 0) {
                const int k = s_k;
                const int idx = k - lane;
                bool ok = idx < 0;  // below tile 0: never needed (tile 0 is a prefix)
                uint64_t w = 0;
                int spins = 0;
                for (;;) {
                    if (!ok) {
                        w = lb_load(status + idx);
                        ok = lb_epoch(w) == ep;
                    }
                    const uint64_t okm = __ballot(ok);
                    const uint64_t pm = __ballot(ok && idx >= 0 && lb_prefix(w));
                    const int fp = pm ? __ffsll(static_cast<unsigned long long>(pm)) - 1 : 63;
                    const uint64_t need = (fp >= 63) ? ~0ull : ((2ull << fp) - 1ull);
                    if ((okm & need) == need) {
                        // lanes 0..fp: aggregates and, at fp (when pm != 0), an
                        // inclusive prefix; without a prefix the whole window
                        // is aggregates (tile 0, a prefix, lies further back)
                        int x = (lane <= fp && idx >= 0) ? lb_value(w) : 0;
#pragma unroll
                        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
                        if (lane == 0) {
                            s_acc += x;
                            s_k = k - 64;
                            s_state = pm != 0 ? 1 : 0;  // 1: done, 0: next window
                        }
                        break;
                    }
                    if (++spins >= kSpinLimit) {
                        // the nearest needed predecessor that has not published:
                        // add the aggregates before it; the workgroup sums its
                        // input, and the look-back resumes below it
                        const uint64_t miss = need & ~okm;
                        const int L = __ffsll(static_cast<unsigned long long>(miss)) - 1;
                        int x = (lane < L) ? lb_value(w) : 0;
#pragma unroll
                        for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, 64);
                        if (lane == 0) {
                            s_acc += x;
                            s_k = k - L;  // the tile to sum from its input
                            s_state = 2;
                        }
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            __syncthreads();
            const int state = s_state;
            const int kk = s_k;
            __syncthreads();
            if (state == 1) break;
            if (state == 2) {
                const int sum = tile_value_sum(op, n, kk, s_red);
                if (t == 0) {
                    s_acc += sum;
                    s_k = kk - 1;
                }
                if (kk == 0) break;  // the fallback summed tile 0
                __syncthreads();
            }
        }
        __syncthreads();
        excl = s_acc;
        if (t == 0) lb_store(status + tile, lb_word(ep, true, excl + agg));
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int e = base + k * kBlock + t;
        if (e < c_end) op.emit(e, it[k], v[k], excl + local[k]);
    }
    if (tile == tiles - 1 && t == 0) op.finish(excl + agg, n, header);
    // completion: the last tile to finish stores the epoch for the next call
    if (t == 0) {
        const uint32_t done = atomicAdd(header + kHdrDone, 1u);
        s_last = done == static_cast<uint32_t>(tiles - 1) ? 1 : 0;
    }
    __syncthreads();
    if (s_last) {
        if (ep >= kEpochMax)  // before the epoch wraps: no stale word may survive
            for (int i = t; i < status_cap; i += kBlock) lb_store(status + i, 0ull);
        __syncthreads();
        if (t == 0) {
            __hip_atomic_store(header + kHdrDone, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(header + kHdrEpoch, ep >= kEpochMax ? 0u : ep, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// ------------------------------------------------------------------ ops

struct ScanI32Op {
    static constexpr bool kGuard = false;
    const int *__restrict__ in;
    int *__restrict__ out;
    using Item = int;
    __device__ int value_at(int e) const { return in[e]; }
    __device__ Item load(int e) const { return in[e]; }
    __device__ int value(Item x) const { return x; }
    __device__ void emit(int e, Item, int, int excl) const { out[e] = excl; }
    __device__ void finish(int, int, uint32_t *) const {}
};

// One particle record as read by the split (68 B in FP64).
template <typename T> struct SplitRec {
    T px, py, pz, vx, vy, vz, m, ts;
    int id;
};

// GUARD: the HOLD split's decision guard (FP64 sets after an approximate
// timestep evaluation), a separate instantiation so that the plain split's
// kernels carry none of its code.
template <typename T, bool GUARD = false> struct SplitOp {
    static constexpr bool kGuard = GUARD && std::is_same<T, double>::value;
    Particles<T> total, slow, fast;
    int *predicate, *scanned;
    T pivot;
    int n;
    hhs_partition *part;
    // HOLD only: drift the slow records as they are written (pos + vel * h,
    // the drift kernel's expression) and publish the raw fast count in a
    // host-mapped ticket, seq << 32 | count, stored last with system-scope
    // release so that a host that sees seq also sees the partition.
    int drift;
    T h;
    unsigned long long *ticket;
    uint32_t seq;
    int *fast_count;  // device word: the reported fast count, the next level's n
    // HOLD only (GUARD): the decision guard, when guard != 0 and, with
    // guard_dev, *guard_dev != 0 (the timestep call ran it already unless
    // its stage overflowed; the word is read once per workgroup)
    int guard;
    SplitGuard gd;
    const int *guard_dev;
    __device__ bool guard_on() const {
        return guard && (!guard_dev || __builtin_amdgcn_readfirstlane(*guard_dev) != 0);
    }
    using Item = SplitRec<T>;
    // src/split.cuh:115 -- NaN compares false and counts as slow.
    __device__ int value_at(int e) const { return total.timestep[e] <= pivot ? 1 : 0; }
    __device__ Item load(int e) const {
        Item r;
        r.ts = total.timestep[e];
        r.px = total.pos[3 * e + 0];
        r.py = total.pos[3 * e + 1];
        r.pz = total.pos[3 * e + 2];
        r.vx = total.vel[3 * e + 0];
        r.vy = total.vel[3 * e + 1];
        r.vz = total.vel[3 * e + 2];
        r.m = total.m[e];
        r.id = total.id ? total.id[e] : e;
        return r;
    }
    __device__ int value(const Item &r) const { return r.ts <= pivot ? 1 : 0; }
    __device__ void emit(int e, const Item &r, int p, int f) const {
#pragma clang fp contract(off)
        if (predicate) predicate[e] = p;
        if (scanned) scanned[e] = f;
        const Particles<T> &d = p ? fast : slow;
        const int o = p ? f : e - f;  // src/split.cuh:336,350
        const bool dr = drift && !p;
        d.pos[3 * o + 0] = dr ? r.px + r.vx * h : r.px;
        d.pos[3 * o + 1] = dr ? r.py + r.vy * h : r.py;
        d.pos[3 * o + 2] = dr ? r.pz + r.vz * h : r.pz;
        d.vel[3 * o + 0] = r.vx;
        d.vel[3 * o + 1] = r.vy;
        d.vel[3 * o + 2] = r.vz;
        d.m[o] = r.m;
        d.timestep[o] = r.ts;
        if (d.id) d.id[o] = r.id;
        d.parent_id[o] = e;
    }
    __device__ void finish(int nf, int n, uint32_t *header) const {
        header[2] = static_cast<uint32_t>(nf);
        const int reported = (nf == 1) ? 0 : nf;  // src/split.cuh:321-324
        part->N_fast = reported;
        part->N_slow = n - reported;
        if (fast_count) *fast_count = reported;
        if (ticket)
            __hip_atomic_store(ticket,
                               (static_cast<unsigned long long>(seq) << 32) |
                                   static_cast<uint32_t>(nf),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
};

template <typename T> struct SplitValuesOp {
    static constexpr bool kGuard = false;
    const T *__restrict__ values;
    T *slow, *fast;
    T pivot;
    int n;
    hhs_partition *part;
    using Item = T;
    __device__ int value_at(int e) const { return values[e] <= pivot ? 1 : 0; }
    __device__ Item load(int e) const { return values[e]; }
    __device__ int value(Item x) const { return x <= pivot ? 1 : 0; }
    __device__ void emit(int e, Item x, int p, int f) const {
        if (p)
            fast[f] = x;
        else
            slow[e - f] = x;
    }
    __device__ void finish(int nf, int n, uint32_t *header) const {
        header[2] = static_cast<uint32_t>(nf);
        part->N_fast = nf;
        part->N_slow = n - nf;
    }
};

// Lone fast particle: the whole set stays slow, in order (the intended
// behaviour of src/split.cuh:321-324; the reference's :361-368 races).
// header null: unconditional (HOLD, which reads the raw count itself);
// drift: positions written as pos + vel * h, as the HOLD split does.
template <typename T>
__global__ __launch_bounds__(256) void lone_fast_fixup_kernel(Particles<T> total,
                                                              Particles<T> slow, int n,
                                                              const uint32_t *header, int drift,
                                                              T h) {
#pragma clang fp contract(off)
    if (header && header[2] != 1u) return;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        const T x = total.pos[3 * e + 0], y = total.pos[3 * e + 1], z = total.pos[3 * e + 2];
        const T vx = total.vel[3 * e + 0], vy = total.vel[3 * e + 1], vz = total.vel[3 * e + 2];
        slow.pos[3 * e + 0] = drift ? x + vx * h : x;
        slow.pos[3 * e + 1] = drift ? y + vy * h : y;
        slow.pos[3 * e + 2] = drift ? z + vz * h : z;
        slow.vel[3 * e + 0] = total.vel[3 * e + 0];
        slow.vel[3 * e + 1] = total.vel[3 * e + 1];
        slow.vel[3 * e + 2] = total.vel[3 * e + 2];
        slow.m[e] = total.m[e];
        slow.timestep[e] = total.timestep[e];
        if (slow.id) slow.id[e] = total.id ? total.id[e] : e;
        slow.parent_id[e] = e;
    }
}

__global__ void zero_partition_kernel(hhs_partition *part) {
    part->N_fast = 0;
    part->N_slow = 0;
}

__global__ void fill_u32_kernel(uint32_t *p, size_t words, uint32_t v) {
    const size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x;
    if (i < words) p[i] = v;
}

// Chunking: inputs of at most two tiles take one workgroup (one launch);
// larger ones are cut into at most kMaxChunks chunks of whole tiles.
struct Chunks {
    int count, per_chunk;
};
Chunks chunks_for(int n) {
    const int tiles = (n + kTile - 1) / kTile;
    if (tiles <= 2) return {1, tiles * kTile};
    const int g = tiles < kMaxChunks ? tiles : kMaxChunks;
    const int tiles_per = (tiles + g - 1) / g;
    return {(tiles + tiles_per - 1) / tiles_per, tiles_per * kTile};
}

// Workspace use.  The look-back scan (n <= kLookbackMaxN): the 64-byte
// header, then one status word per tile.  The reduce-then-scan (larger n,
// in-place scans, and the allocation-free split's fixed scratch,
// `lookback_ws` false): the header,
// then one count per chunk -- placed after the status words of a workspace
// that the look-back also uses, so that no count is ever read as a status
// word.  A workspace starts zeroed (hhs_scan_workspace_init).
// n_dev non-null: n is an upper bound of the count read on the device (the
// look-back form only), and nothing runs unless that count exceeds n_min.
template <bool BINARY, class Op>
int launch_scan_op(const Op &op, int n, void *ws_ptr, size_t ws_bytes, hipStream_t st,
                   bool lookback_ws = true, bool in_place = false, const int *n_dev = nullptr,
                   int n_min = 0) {
    HHS_REQUIRE(ws_ptr, "scan: null workspace");
    HHS_REQUIRE(!n_dev || (lookback_ws && !in_place && n <= kLookbackMaxN),
                "scan: a device-side count needs the look-back form (n <= %d)", kLookbackMaxN);
    uint32_t *header = static_cast<uint32_t *>(ws_ptr);
    char *after = static_cast<char *>(ws_ptr) + kScanHeaderBytes;
    const int tiles = (n + kTile - 1) / kTile;
    const size_t status_bytes = sizeof(uint64_t) * static_cast<size_t>(tiles);
    if (lookback_ws && !in_place && n <= kLookbackMaxN) {
        if (ws_bytes < kScanHeaderBytes + status_bytes) {
            set_error("scan: workspace of %zu bytes too small for n=%d (need %zu)", ws_bytes, n,
                      kScanHeaderBytes + status_bytes);
            return HHS_ERR_WORKSPACE;
        }
        const int cap = static_cast<int>((ws_bytes - kScanHeaderBytes) / sizeof(uint64_t));
        hipLaunchKernelGGL((scan_lookback_kernel<BINARY, Op>), dim3(tiles), dim3(kBlock), 0, st, op,
                           n, header, reinterpret_cast<uint64_t *>(after), tiles, cap, n_dev,
                           n_min);
        HHS_LAUNCH_CHECK("scan_lookback_kernel");
        return HHS_OK;
    }
    const Chunks c = chunks_for(n);
    const size_t skip = lookback_ws ? status_bytes : 0;
    const size_t need = kScanHeaderBytes + skip + sizeof(int) * static_cast<size_t>(c.count);
    if (ws_bytes < need) {
        set_error("scan: workspace of %zu bytes too small for n=%d (need %zu)", ws_bytes, n,
                  need);
        return HHS_ERR_WORKSPACE;
    }
    ScanWs ws{header, reinterpret_cast<int *>(after + skip)};
    if (c.count > 1) {
        hipLaunchKernelGGL((count_kernel<Op>), dim3(c.count), dim3(kBlock), 0, st, op, n,
                           c.per_chunk, ws.counts);
        HHS_LAUNCH_CHECK("count_kernel");
    }
    hipLaunchKernelGGL((scan_chunks_kernel<BINARY, Op>), dim3(c.count), dim3(kBlock), 0, st, op,
                       n, c.per_chunk, ws);
    HHS_LAUNCH_CHECK("scan_chunks_kernel");
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
                 void *ws, size_t ws_bytes, hipStream_t st, bool lookback_ws) {
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
    int rc = launch_scan_op<true>(op, n, ws, ws_bytes, st, lookback_ws);
    if (rc) return rc;
    const int fix_groups = (n + 255) / 256 < 512 ? (n + 255) / 256 : 512;
    hipLaunchKernelGGL((lone_fast_fixup_kernel<T>), dim3(fix_groups), dim3(256), 0, st, total,
                       slow, n, static_cast<const uint32_t *>(ws), 0, T(0));
    HHS_LAUNCH_CHECK("lone_fast_fixup_kernel");
    return HHS_OK;
}

// The HOLD form of the split (one call of src/hold.cuh:79-83 plus the
// first slow half-drift of :123-124, which the fast recursion does not
// read): slow records are written drifted by h_slow, the counts go to
// `part` and the raw fast count to `raw` (host-mapped memory, read after a
// stream sync), and no lone-particle follow-up is launched -- the caller
// runs launch_lone_fast_fixup when *raw == 1.
//
// guard_band > 0 (FP64 sets; calls that have just computed their timesteps
// in an approximate mode): the decision guard with that band, the pivot as
// the call's dt and the criterion's eta and eps2, runs in the count pass.
template <typename T>
int launch_split_hold(const Particles<T> &total, const Particles<T> &slow,
                      const Particles<T> &fast, T pivot, int n, hhs_partition *part,
                      unsigned long long *ticket, uint32_t seq, int *fast_count, T h_slow,
                      void *ws, size_t ws_bytes, hipStream_t st, double guard_band, double eta,
                      double eps2, const int *n_dev, int n_min, const int *guard_dev) {
    HHS_REQUIRE(n > 0 && part && ticket, "split_hold: bad arguments");
    HHS_REQUIRE(!n_dev || guard_band <= 0.0, "split_hold: a device-side count takes no guard");
    if constexpr (std::is_same<T, double>::value) {
        if (guard_band > 0.0) {
            SplitOp<T, true> op{total, slow, fast, nullptr, nullptr, pivot, n, part, 1, h_slow,
                                ticket, seq, fast_count};
            op.guard = 1;
            op.gd.g = make_ts_guard(static_cast<double>(pivot), guard_band);
            op.gd.eta = eta;
            op.gd.c_ff = eta / TsConst<double>::sqrt2;
            op.gd.eps2 = eps2;
            op.guard_dev = guard_dev;
            return launch_scan_op<true>(op, n, ws, ws_bytes, st);
        }
    }
    SplitOp<T> op{total, slow, fast, nullptr, nullptr, pivot, n, part, 1, h_slow, ticket, seq,
                  fast_count};
    return launch_scan_op<true>(op, n, ws, ws_bytes, st, true, false, n_dev, n_min);
}

template <typename T>
int launch_lone_fast_fixup(const Particles<T> &total, const Particles<T> &slow, int n, T h_slow,
                           hipStream_t st) {
    const int fix_groups = (n + 255) / 256 < 512 ? (n + 255) / 256 : 512;
    hipLaunchKernelGGL((lone_fast_fixup_kernel<T>), dim3(fix_groups), dim3(256), 0, st, total,
                       slow, n, static_cast<const uint32_t *>(nullptr), 1, h_slow);
    HHS_LAUNCH_CHECK("lone_fast_fixup_kernel");
    return HHS_OK;
}

template int launch_split_hold<double>(const Particles<double> &, const Particles<double> &,
                                       const Particles<double> &, double, int, hhs_partition *,
                                       unsigned long long *, uint32_t, int *, double, void *,
                                       size_t, hipStream_t, double, double, double, const int *,
                                       int, const int *);
template int launch_split_hold<float>(const Particles<float> &, const Particles<float> &,
                                      const Particles<float> &, float, int, hhs_partition *,
                                      unsigned long long *, uint32_t, int *, float, void *,
                                      size_t, hipStream_t, double, double, double, const int *,
                                      int, const int *);
template int launch_lone_fast_fixup<double>(const Particles<double> &, const Particles<double> &,
                                            int, double, hipStream_t);
template int launch_lone_fast_fixup<float>(const Particles<float> &, const Particles<float> &,
                                           int, float, hipStream_t);

template int launch_split<double>(const Particles<double> &, const Particles<double> &,
                                  const Particles<double> &, int *, int *, double, int,
                                  hhs_partition *, void *, size_t, hipStream_t, bool);
template int launch_split<float>(const Particles<float> &, const Particles<float> &,
                                 const Particles<float> &, int *, int *, float, int,
                                 hhs_partition *, void *, size_t, hipStream_t, bool);

// ------------------------------------------------ allocation-free split
// File-scope device scratch of hhs_split_f64_noalloc: the reduce-then-scan's
// workspace (the 64-byte header and one count per chunk), whose need is
// kScanHeaderBytes + 4 * chunks with at most kMaxChunks chunks for any n --
// 8256 bytes.  Each GPU holds its own copy from the library's load until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: the per-device address comes from hipGetSymbolAddress
// (noalloc_scratch) and is the only one passed to kernels.
constexpr size_t kNoAllocScratchBytes = kScanHeaderBytes + sizeof(int) * kMaxChunks;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device (with
// that device current) on its first call.
int noalloc_scratch(void **out) {
    int dev = 0;
    HHS_HIP(hipGetDevice(&dev));
    HHS_REQUIRE(dev >= 0 && dev < kMaxDevices, "split_noalloc: device %d beyond %d", dev,
                kMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        HHS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return HHS_OK;
}

}  // namespace hhs

using namespace hhs;

template <typename T, typename P>
static int split_c(const P *total, P *slow, P *fast, int *predicate, int *scanned, T pivot, int n,
                   hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream,
                   bool lookback_ws = true) {
    HHS_REQUIRE(total && slow && fast, "split: null system");
    HHS_REQUIRE(n >= 0 && n <= total->n, "split: n=%d outside [0, %d]", n, total->n);
    HHS_REQUIRE(slow->n >= n && fast->n >= n, "split: slow/fast capacity below n=%d", n);
    return launch_split<T>(from_c(*total), from_c(*slow), from_c(*fast), predicate, scanned,
                           pivot, n, part_dev, ws, ws_bytes, as_stream(stream), lookback_ws);
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
    // the look-back may sum a predecessor's input after that tile has
    // written its output: an in-place scan takes the reduce-then-scan
    const bool in_place = in == out;
    return launch_scan_op<false>(op, n, ws, ws_bytes, as_stream(stream), true, in_place);
}

int hhs_split_f64(const hhs_particles_f64 *total, hhs_particles_f64 *slow,
                  hhs_particles_f64 *fast, int *predicate, int *scanned, double pivot, int n,
                  hhs_partition *part_dev, void *ws, size_t ws_bytes, void *stream) {
    return split_c<double>(total, slow, fast, predicate, scanned, pivot, n, part_dev, ws,
                           ws_bytes, stream);
}

int hhs_split_f64_noalloc(const hhs_particles_f64 *total, hhs_particles_f64 *slow,
                          hhs_particles_f64 *fast, int *predicate, int *scanned, double pivot,
                          int n, hhs_partition *part_dev, void *stream) {
    HHS_REQUIRE(n >= 0, "split_noalloc: negative n");
    const size_t need = kScanHeaderBytes + sizeof(int) * static_cast<size_t>(chunks_for(n).count);
    if (need > kNoAllocScratchBytes) {
        set_error("split_noalloc: n=%d needs %zu bytes of scratch, the static scratch holds %zu",
                  n, need, kNoAllocScratchBytes);
        return HHS_ERR_WORKSPACE;
    }
    void *ws = nullptr;
    if (int rc = noalloc_scratch(&ws)) return rc;
    return split_c<double>(total, slow, fast, predicate, scanned, pivot, n, part_dev, ws,
                           kNoAllocScratchBytes, stream, false);
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
