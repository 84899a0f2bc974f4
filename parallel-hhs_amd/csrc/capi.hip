// Error reporting, version and device-memory helpers of the C ABI.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "hhs_internal.h"

namespace hhs {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void clear_error() { g_err[0] = '\0'; }

int hip_status(hipError_t e, const char *what, const char *file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return HHS_ERR_NO_DEVICE;
    if (e == hipErrorOutOfMemory) return HHS_ERR_ALLOC;
    return HHS_ERR_HIP;
}

// No event is recorded per use: an event record after each kernel that
// uses the scratch put a ~5 us bubble before the next kernel on MI355X.
// A change of stream or a regrow instead waits for the whole device, which
// covers the last use whatever stream it was on (and that stream need not
// still exist).
ThreadScratch::~ThreadScratch() {
    if (ptr_) (void)hipFree(ptr_);
}

int ThreadScratch::acquire(size_t bytes, hipStream_t st, void **out) {
    if (bytes > bytes_) {
        if (used_) HHS_HIP(hipDeviceSynchronize());  // the last user is done
        if (ptr_) HHS_HIP(hipFree(ptr_));
        ptr_ = nullptr;
        bytes_ = 0;
        used_ = false;
        HHS_HIP(hipMalloc(&ptr_, bytes));
        bytes_ = bytes;
    } else if (used_ && stream_ != st) {
        HHS_HIP(hipDeviceSynchronize());
    }
    *out = ptr_;
    return HHS_OK;
}

int ThreadScratch::release(hipStream_t st) {
    stream_ = st;
    used_ = true;
    return HHS_OK;
}

}  // namespace hhs

using namespace hhs;

extern "C" {

const char *hhs_last_error(void) { return g_err; }

const char *hhs_version(void) { return "hhs-mi355x 0.1.0 (gfx950)"; }

int hhs_runtime_check(void) {
    FILE *f = std::fopen("/proc/self/maps", "r");
    if (!f) return HHS_OK;  // nothing to inspect
    char line[1024], first[512] = "";
    int rc = HHS_OK;
    while (std::fgets(line, sizeof line, f)) {
        const char *path = std::strchr(line, '/');
        if (!path || !std::strstr(path, "libamdhip64.so")) continue;
        char p[512];
        std::snprintf(p, sizeof p, "%s", path);
        p[std::strcspn(p, "\n")] = '\0';
        if (!first[0]) {
            std::snprintf(first, sizeof first, "%s", p);
        } else if (std::strcmp(first, p) != 0) {
            set_error("two HIP runtimes in the process: %s and %s (load PyTorch's first)", first,
                      p);
            rc = HHS_ERR_HIP;
            break;
        }
    }
    std::fclose(f);
    return rc;
}

int hhs_device_count(int *count) {
    HHS_REQUIRE(count, "hhs_device_count: null output");
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess) {
        *count = 0;
        return hip_status(e, "hipGetDeviceCount", __FILE__, __LINE__);
    }
    *count = c;
    return HHS_OK;
}

int hhs_malloc(void **ptr, size_t bytes) {
    HHS_REQUIRE(ptr, "hhs_malloc: null output");
    *ptr = nullptr;
    if (bytes == 0) return HHS_OK;
    HHS_HIP(hipMalloc(ptr, bytes));
    return HHS_OK;
}

int hhs_free(void *ptr) {
    if (ptr) HHS_HIP(hipFree(ptr));
    return HHS_OK;
}

int hhs_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return HHS_OK;
    HHS_REQUIRE(dst && src, "hhs_memcpy_h2d: null pointer");
    HHS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return HHS_OK;
}

int hhs_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return HHS_OK;
    HHS_REQUIRE(dst && src, "hhs_memcpy_d2h: null pointer");
    HHS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return HHS_OK;
}

int hhs_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    if (bytes == 0) return HHS_OK;
    HHS_REQUIRE(dst && src, "hhs_memcpy_d2d: null pointer");
    HHS_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return HHS_OK;
}

int hhs_memset(void *dst, int value, size_t bytes) {
    if (bytes == 0) return HHS_OK;
    HHS_REQUIRE(dst, "hhs_memset: null pointer");
    HHS_HIP(hipMemset(dst, value, bytes));
    return HHS_OK;
}

int hhs_device_synchronize(void) {
    HHS_HIP(hipDeviceSynchronize());
    return HHS_OK;
}

int hhs_stream_synchronize(void *stream) {
    HHS_HIP(hipStreamSynchronize(as_stream(stream)));
    return HHS_OK;
}

int hhs_host_alloc_mapped(void **ptr, size_t bytes) {
    HHS_REQUIRE(ptr, "hhs_host_alloc_mapped: null output");
    *ptr = nullptr;
    if (bytes == 0) return HHS_OK;
    HHS_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return HHS_OK;
}

int hhs_host_free_mapped(void *ptr) {
    if (ptr) HHS_HIP(hipHostFree(ptr));
    return HHS_OK;
}

}  // extern "C"
