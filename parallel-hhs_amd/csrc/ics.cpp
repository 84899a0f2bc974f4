// Initial conditions (host only; no GPU use).
//
// The reference has no IC generator.  Plummer spheres follow Aarseth, Henon
// & Wielen (1974, A&A 37, 183): radius from the inverted cumulative mass,
// speed by von Neumann rejection on g(q) = q^2 (1 - q^2)^{7/2}, isotropic
// directions; then scaled to Henon units (G = M = 1, E = -1/4: length
// 3*pi/16, velocity sqrt(16/(3*pi)) of the Plummer-unit values).  Radii
// beyond 10 Plummer scale lengths are redrawn.  Salpeter masses use the
// inverse CDF of dN/dm ~ m^-2.35 on [0.1, 100].  Random numbers come from
// SplitMix64 (Steele, Lea & Flood 2014), one stream per seed.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "hhs.h"

namespace hhs {
void set_error(const char *fmt, ...);
}

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0, 1)
};

void isotropic(SplitMix64 &rng, double len, double *out) {
    const double ct = 1.0 - 2.0 * rng.uniform();
    const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
    const double ph = 2.0 * M_PI * rng.uniform();
    out[0] = len * st * std::cos(ph);
    out[1] = len * st * std::sin(ph);
    out[2] = len * ct;
}

// Exact O(n^2) potential with per-row partials summed in index order, so
// the value does not depend on the number of threads.
double potential(const double *pos, const double *m, int n, double eps2) {
    std::vector<double> row(static_cast<size_t>(n), 0.0);
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t]() {
            for (int i = static_cast<int>(t); i < n; i += static_cast<int>(nt)) {
                double s = 0.0;
                for (int j = i + 1; j < n; ++j) {
                    const double dx = pos[3 * j] - pos[3 * i];
                    const double dy = pos[3 * j + 1] - pos[3 * i + 1];
                    const double dz = pos[3 * j + 2] - pos[3 * i + 2];
                    const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                    if (r2 > 0.0) s += m[j] / std::sqrt(r2);
                }
                row[static_cast<size_t>(i)] = m[i] * s;
            }
        });
    }
    for (auto &x : th) x.join();
    double pe = 0.0;
    for (int i = 0; i < n; ++i) pe -= row[static_cast<size_t>(i)];
    return pe;
}

}  // namespace

extern "C" int hhs_ic_plummer(int n, uint64_t seed, int mass_model, int virialize, double eps2,
                              double *pos, double *vel, double *m) {
    if (n <= 0 || !pos || !vel || !m || (mass_model != 0 && mass_model != 1)) {
        hhs::set_error("ic_plummer: need n > 0, non-null arrays and mass_model 0 or 1");
        return HHS_ERR_INVALID;
    }
    SplitMix64 rng{seed};
    const double rscale = 3.0 * M_PI / 16.0;
    const double vscale = std::sqrt(16.0 / (3.0 * M_PI));
    for (int i = 0; i < n; ++i) {
        double r;
        do {
            double x1;
            do {
                x1 = rng.uniform();
            } while (x1 <= 0.0);
            r = 1.0 / std::sqrt(std::pow(x1, -2.0 / 3.0) - 1.0);
        } while (!(r <= 10.0));
        isotropic(rng, r * rscale, &pos[3 * i]);
        double q, g;
        do {
            q = rng.uniform();
            g = rng.uniform() * 0.1;
        } while (g > q * q * std::pow(1.0 - q * q, 3.5));
        const double ve = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
        isotropic(rng, q * ve * vscale, &vel[3 * i]);
    }
    if (mass_model == 0) {
        for (int i = 0; i < n; ++i) m[i] = 1.0 / n;
    } else {
        const double a = 1.0 - 2.35, lo = std::pow(0.1, a), hi = std::pow(100.0, a);
        double tot = 0.0;
        for (int i = 0; i < n; ++i) {
            m[i] = std::pow(lo + (hi - lo) * rng.uniform(), 1.0 / a);
            tot += m[i];
        }
        for (int i = 0; i < n; ++i) m[i] /= tot;
    }
    // centre of mass at rest at the origin
    double cm[6] = {0, 0, 0, 0, 0, 0}, mt = 0.0;
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            cm[k] += m[i] * pos[3 * i + k];
            cm[3 + k] += m[i] * vel[3 * i + k];
        }
        mt += m[i];
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k) {
            pos[3 * i + k] -= cm[k] / mt;
            vel[3 * i + k] -= cm[3 + k] / mt;
        }
    if (virialize && n > 1) {
        double ke = 0.0;
        for (int i = 0; i < n; ++i)
            ke += 0.5 * m[i] *
                  (vel[3 * i] * vel[3 * i] + vel[3 * i + 1] * vel[3 * i + 1] +
                   vel[3 * i + 2] * vel[3 * i + 2]);
        const double pe = potential(pos, m, n, eps2);
        if (ke > 0.0 && pe < 0.0) {
            const double f = std::sqrt(0.5 * -pe / ke);
            for (int i = 0; i < 3 * n; ++i) vel[i] *= f;
        }
    }
    return HHS_OK;
}
