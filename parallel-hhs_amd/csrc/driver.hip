// Top-level driver and FP64 snapshots (host side).
//
// The reference declares hold() but never defines it (src/hold.cuh:20-35)
// and has no checkpointing; its only state export is gpu_to_host
// (src/particle_system.cuh:230-247).  hhs_evolve_f64 runs root steps of
// hold_step with calc_timestep = true and writes snapshots at root-step
// boundaries, the only points where every particle is at the same time and
// no state besides pos, vel, m (and the ids) carries over to the next step.
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <vector>

#include "hhs_internal.h"

namespace hhs {
namespace {

static_assert(sizeof(hhs_snapshot_header) == 128, "snapshot header must be 128 bytes");

constexpr char kMagic[8] = {'H', 'H', 'S', 'S', 'N', 'A', 'P', '\0'};

// FNV-1a, 64-bit (Fowler, Noll, Vo), over a byte range, continuing from h.
uint64_t fnv1a(const void *data, size_t bytes, uint64_t h) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < bytes; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ULL;
    }
    return h;
}
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ULL;

size_t payload_bytes(int64_t n) { return static_cast<size_t>(n) * (3 * 8 + 3 * 8 + 8 + 4); }

// Pinned host staging for one snapshot's payload.
struct HostBuffer {
    void *p = nullptr;
    size_t bytes = 0;
    ~HostBuffer() {
        if (p) (void)hipHostFree(p);
    }
    int reserve(size_t b) {
        if (b <= bytes) return HHS_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        HHS_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
        bytes = b;
        return HHS_OK;
    }
};

int write_snapshot(const char *path, const hhs_globals *g, const hhs_particles_f64 &sys, int n,
                   double dt_root, hipStream_t st, HostBuffer &buf) {
    HHS_REQUIRE(path && *path, "snapshot_write: empty path");
    HHS_REQUIRE(n >= 0 && n <= sys.n, "snapshot_write: n=%d outside [0, %d]", n, sys.n);
    HHS_REQUIRE(n == 0 || (sys.pos && sys.vel && sys.m), "snapshot_write: null pos/vel/m");
    const size_t nn = static_cast<size_t>(n);
    int rc = buf.reserve(payload_bytes(n) + 1);
    if (rc) return rc;
    char *pos = static_cast<char *>(buf.p);
    char *vel = pos + 24 * nn;
    char *m = vel + 24 * nn;
    int *id = reinterpret_cast<int *>(m + 8 * nn);
    if (n > 0) {
        HHS_HIP(hipMemcpyAsync(pos, sys.pos, 24 * nn, hipMemcpyDeviceToHost, st));
        HHS_HIP(hipMemcpyAsync(vel, sys.vel, 24 * nn, hipMemcpyDeviceToHost, st));
        HHS_HIP(hipMemcpyAsync(m, sys.m, 8 * nn, hipMemcpyDeviceToHost, st));
        if (sys.id) HHS_HIP(hipMemcpyAsync(id, sys.id, 4 * nn, hipMemcpyDeviceToHost, st));
    }
    HHS_HIP(hipStreamSynchronize(st));
    if (!sys.id)
        for (int i = 0; i < n; ++i) id[i] = i;

    hhs_snapshot_header h;
    std::memset(&h, 0, sizeof(h));
    std::memcpy(h.magic, kMagic, sizeof(kMagic));
    h.version = HHS_SNAPSHOT_VERSION;
    h.header_bytes = sizeof(h);
    h.n = n;
    h.root_steps = g ? hhs_globals_root_steps(g) : 0;
    h.simtime = g ? hhs_globals_simtime(g) : 0.0;
    h.dt_root = dt_root;
    h.eta = g ? hhs_globals_eta(g) : 0.0;
    h.eps = g ? hhs_globals_eps(g) : 0.0;
    h.timestep_mode = g ? hhs_globals_timestep_mode(g) : HHS_TIMESTEP_EXACT;
    h.kick_precision = g ? hhs_globals_kick_precision(g) : -1;
    h.payload_fnv1a = fnv1a(buf.p, payload_bytes(n), kFnvBasis);

    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) {
        set_error("snapshot_write: cannot open %s: %s", tmp.c_str(), std::strerror(errno));
        return HHS_ERR_IO;
    }
    bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1;
    if (ok && n > 0) ok = std::fwrite(buf.p, payload_bytes(n), 1, f) == 1;
    ok = (std::fflush(f) == 0) && ok;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path) != 0) {
        set_error("snapshot_write: writing %s failed: %s", path, std::strerror(errno));
        std::remove(tmp.c_str());
        return HHS_ERR_IO;
    }
    return HHS_OK;
}

int read_header(FILE *f, const char *path, hhs_snapshot_header &h) {
    if (std::fseek(f, 0, SEEK_END) != 0) {
        set_error("snapshot: cannot seek %s", path);
        return HHS_ERR_IO;
    }
    const long size = std::ftell(f);
    std::rewind(f);
    if (size < static_cast<long>(sizeof(h)) || std::fread(&h, sizeof(h), 1, f) != 1) {
        set_error("snapshot: %s is shorter than a snapshot header", path);
        return HHS_ERR_FORMAT;
    }
    if (std::memcmp(h.magic, kMagic, sizeof(kMagic)) != 0) {
        set_error("snapshot: %s is not a snapshot (bad magic)", path);
        return HHS_ERR_FORMAT;
    }
    if (h.version != HHS_SNAPSHOT_VERSION || h.header_bytes != sizeof(h)) {
        set_error("snapshot: %s has version %u / header %u bytes (expected %d / %zu)", path,
                  h.version, h.header_bytes, HHS_SNAPSHOT_VERSION, sizeof(h));
        return HHS_ERR_FORMAT;
    }
    if (h.n < 0 || h.n > INT32_MAX ||
        static_cast<unsigned long>(size) != sizeof(h) + payload_bytes(h.n)) {
        set_error("snapshot: %s holds %ld bytes, not the %zu of n=%lld", path, size,
                  sizeof(h) + payload_bytes(h.n < 0 ? 0 : h.n), static_cast<long long>(h.n));
        return HHS_ERR_FORMAT;
    }
    return HHS_OK;
}

}  // namespace
}  // namespace hhs

using namespace hhs;

extern "C" {

int hhs_snapshot_write_f64(const char *path, const hhs_globals *g, const hhs_particles_f64 *sys,
                           int n, double dt_root, void *stream) {
    HHS_REQUIRE(sys, "snapshot_write: null system");
    HostBuffer buf;
    return write_snapshot(path, g, *sys, n, dt_root, as_stream(stream), buf);
}

int hhs_snapshot_read_header(const char *path, hhs_snapshot_header *out) {
    HHS_REQUIRE(path && out, "snapshot_read_header: null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_error("snapshot: cannot open %s: %s", path, std::strerror(errno));
        return HHS_ERR_IO;
    }
    const int rc = read_header(f, path, *out);
    std::fclose(f);
    return rc;
}

int hhs_snapshot_read_f64(const char *path, hhs_particles_f64 *sys, hhs_globals *g,
                          hhs_snapshot_header *hdr, void *stream) {
    HHS_REQUIRE(path && sys, "snapshot_read: null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_error("snapshot: cannot open %s: %s", path, std::strerror(errno));
        return HHS_ERR_IO;
    }
    hhs_snapshot_header h;
    int rc = read_header(f, path, h);
    HostBuffer buf;
    if (!rc) rc = buf.reserve(payload_bytes(h.n) + 1);
    if (!rc && h.n > 0 && std::fread(buf.p, payload_bytes(h.n), 1, f) != 1) {
        set_error("snapshot: reading the payload of %s failed", path);
        rc = HHS_ERR_IO;
    }
    std::fclose(f);
    if (rc) return rc;
    if (fnv1a(buf.p, payload_bytes(h.n), kFnvBasis) != h.payload_fnv1a) {
        set_error("snapshot: %s payload checksum mismatch (damaged file)", path);
        return HHS_ERR_FORMAT;
    }
    const int n = static_cast<int>(h.n);
    HHS_REQUIRE(n <= sys->n, "snapshot_read: snapshot has n=%d, arrays hold %d", n, sys->n);
    HHS_REQUIRE(n == 0 || (sys->pos && sys->vel && sys->m), "snapshot_read: null pos/vel/m");
    if (g) {
        HHS_REQUIRE(hhs_globals_eta(g) == h.eta && hhs_globals_eps(g) == h.eps,
                    "snapshot_read: globals eta/eps (%g, %g) differ from the snapshot's (%g, %g)",
                    hhs_globals_eta(g), hhs_globals_eps(g), h.eta, h.eps);
        HHS_REQUIRE(hhs_globals_timestep_mode(g) == h.timestep_mode &&
                        hhs_globals_kick_precision(g) == h.kick_precision,
                    "snapshot_read: globals timestep/kick modes (%d, %d) differ from the "
                    "snapshot's (%d, %d)",
                    hhs_globals_timestep_mode(g), hhs_globals_kick_precision(g), h.timestep_mode,
                    h.kick_precision);
    }
    const size_t nn = static_cast<size_t>(n);
    const char *pos = static_cast<const char *>(buf.p);
    const char *vel = pos + 24 * nn;
    const char *m = vel + 24 * nn;
    const char *id = m + 8 * nn;
    const hipStream_t st = as_stream(stream);
    if (n > 0) {
        HHS_HIP(hipMemcpyAsync(sys->pos, pos, 24 * nn, hipMemcpyHostToDevice, st));
        HHS_HIP(hipMemcpyAsync(sys->vel, vel, 24 * nn, hipMemcpyHostToDevice, st));
        HHS_HIP(hipMemcpyAsync(sys->m, m, 8 * nn, hipMemcpyHostToDevice, st));
        if (sys->id) HHS_HIP(hipMemcpyAsync(sys->id, id, 4 * nn, hipMemcpyHostToDevice, st));
    }
    HHS_HIP(hipStreamSynchronize(st));  // the staging buffer is freed on return
    if (g) {
        hhs_globals_set_simtime(g, h.simtime);
        hhs_globals_set_root_steps(g, h.root_steps);
    }
    if (hdr) *hdr = h;
    return HHS_OK;
}

int hhs_evolve_f64(hhs_globals *g, hhs_particles_f64 *sys, double t_end, double dt_root,
                   const char *snapshot_path, int snapshot_every, void *stream) {
    HHS_REQUIRE(g && sys, "evolve: null argument");
    HHS_REQUIRE(dt_root > 0.0, "evolve: dt_root must be positive");
    HHS_REQUIRE(snapshot_every >= 0, "evolve: snapshot_every must be >= 0");
    HostBuffer buf;
    long long k = hhs_globals_root_steps(g);
    bool written = true;
    while (static_cast<double>(k) * dt_root < t_end) {
        const double t0 = static_cast<double>(k) * dt_root;
        int rc = hhs_hold_step_f64(g, sys, sys->n, 0, t0, t0 + dt_root, dt_root, 1, stream);
        if (rc) return rc;
        hhs_globals_set_root_steps(g, ++k);
        written = false;
        if (snapshot_path && snapshot_every > 0 && k % snapshot_every == 0) {
            if ((rc = write_snapshot(snapshot_path, g, *sys, sys->n, dt_root, as_stream(stream),
                                     buf)))
                return rc;
            written = true;
        }
    }
    if (snapshot_path && !written)
        return write_snapshot(snapshot_path, g, *sys, sys->n, dt_root, as_stream(stream), buf);
    return HHS_OK;
}

}  // extern "C"
