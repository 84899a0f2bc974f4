// Drives the continuous-batching server (parallel-hhs_amd/csrc/server.hip)
// through its C ABI -- submits from two host threads, polling, waiting,
// fetching, releasing, an oversized request, and destruction with requests
// still in flight -- for a build of the library whose host code runs under
// AddressSanitizer and UndefinedBehaviorSanitizer (`make -C parallel-hhs_amd
// asan`; device code is not instrumented).  Test infrastructure: run by
// tests/test_gpu_sanitize.py on the GPU.  Exit status 0 when every call
// returned as documented and the sanitizers reported nothing.
#include <sanitizer/lsan_interface.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "hhs.h"

namespace {

int failures = 0;
#define CHECK(c, msg)                                                         \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "FAIL %s:%d: %s (%s)\n", __FILE__, __LINE__, \
                         msg, hhs_last_error());                              \
            ++failures;                                                       \
        }                                                                     \
    } while (0)

struct System {
    int n;
    std::vector<double> pos, vel, m;
};

System plummer(int n, unsigned long long seed) {
    System s{n, std::vector<double>(3 * n), std::vector<double>(3 * n), std::vector<double>(n)};
    const int rc = hhs_ic_plummer(n, seed, 0, 1, 1e-8, s.pos.data(), s.vel.data(), s.m.data());
    CHECK(rc == HHS_OK, "ic_plummer");
    return s;
}

void submit_batch(hhs_server *srv, const std::vector<System> &sys, int first, int stride,
                  std::vector<long long> &tickets) {
    for (size_t k = first; k < sys.size(); k += stride) {
        long long t = -1;
        const int rc = hhs_server_submit(srv, sys[k].n, sys[k].pos.data(), sys[k].vel.data(),
                                         sys[k].m.data(), 0.03125, 1.0 / 64, &t);
        CHECK(rc == HHS_OK && t >= 0, "submit");
        tickets[k] = t;
    }
}

}  // namespace

int main() {
    if (hhs_runtime_check() != HHS_OK) {
        std::fprintf(stderr, "runtime check: %s\n", hhs_last_error());
        return 1;
    }
    const int sizes[] = {8, 40, 130, 300, 700, 1500, 2600, 64, 900, 17, 513, 3000};
    std::vector<System> sys;
    for (size_t k = 0; k < sizeof(sizes) / sizeof(sizes[0]); ++k)
        sys.push_back(plummer(sizes[k], 900 + k));

    hhs_server_config cfg{};
    cfg.eta = 0.01;
    cfg.eps = 1e-4;
    cfg.hbm_budget_bytes = size_t{1} << 30;
    cfg.small_max_n = 512;
    cfg.max_n = 4096;
    cfg.batch_streams = 2;
    cfg.big_workers = 2;
    cfg.timestep_mode = HHS_TIMESTEP_FP32;

    // 1: two submitting threads, then wait / fetch / release every request
    {
        hhs_server *srv = nullptr;
        CHECK(hhs_server_create(&srv, &cfg) == HHS_OK && srv, "create");
        std::vector<long long> tickets(sys.size(), -1);
        std::thread a(submit_batch, srv, std::cref(sys), 0, 2, std::ref(tickets));
        std::thread b(submit_batch, srv, std::cref(sys), 1, 2, std::ref(tickets));
        a.join();
        b.join();
        for (size_t k = 0; k < sys.size(); ++k) {
            hhs_request_info info{};
            CHECK(hhs_server_poll(srv, tickets[k], &info) == HHS_OK, "poll");
            CHECK(hhs_server_wait(srv, tickets[k], &info) == HHS_OK, "wait");
            CHECK(info.state == 3, "request done");
            std::vector<double> p(3 * sys[k].n), v(3 * sys[k].n);
            CHECK(hhs_server_fetch(srv, tickets[k], p.data(), v.data()) == HHS_OK, "fetch");
            bool finite = true;
            for (double x : p) finite = finite && std::isfinite(x);
            CHECK(finite, "finite positions");
            CHECK(hhs_server_release(srv, tickets[k]) == HHS_OK, "release");
        }
        hhs_server_stats_t st{};
        CHECK(hhs_server_stats(srv, &st) == HHS_OK && st.used_bytes == 0, "stats after release");
        // a request that can never fit the budget is refused
        hhs_server_config tiny = cfg;
        tiny.hbm_budget_bytes = 1000;
        hhs_server *s2 = nullptr;
        CHECK(hhs_server_create(&s2, &tiny) == HHS_OK, "create tiny");
        long long t = -1;
        CHECK(hhs_server_submit(s2, sys[4].n, sys[4].pos.data(), sys[4].vel.data(),
                                sys[4].m.data(), 0.1, 0.1, &t) == HHS_ERR_BUDGET,
              "budget refusal");
        CHECK(hhs_server_destroy(s2) == HHS_OK, "destroy tiny");
        CHECK(hhs_server_destroy(srv) == HHS_OK, "destroy");
    }
    // 2: destruction with requests still in flight and never waited for
    {
        hhs_server *srv = nullptr;
        CHECK(hhs_server_create(&srv, &cfg) == HHS_OK && srv, "create 2");
        std::vector<long long> tickets(sys.size(), -1);
        submit_batch(srv, sys, 0, 1, tickets);
        CHECK(hhs_server_destroy(srv) == HHS_OK, "destroy with pending requests");
    }
    // Leak check now, then leave without running library destructors: the
    // ROCm runtime's own teardown at exit frees memory through ASan's device
    // allocator after that allocator has been told the runtime is unloaded,
    // and ASan aborts there (sanitizer_allocator_device.h), in no code of
    // this library.  A leak report ends the process inside the call.
    __lsan_do_leak_check();
    if (failures) std::fprintf(stderr, "%d check(s) failed\n", failures);
    else std::printf("server sanitizer driver: ok\n");
    std::fflush(stdout);
    std::fflush(stderr);
    std::_Exit(failures ? 1 : 0);
}
