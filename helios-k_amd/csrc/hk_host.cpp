// hk_host.cpp -- host-side parts of the C ABI: error text, version, and the
// seeded synthetic line generator used by the drivers, tests and benchmark.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "hk_internal.h"

namespace hk {

static thread_local char g_err[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

// SplitMix64 (Steele, Lea & Flood 2014) as a counter-based generator.
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform double in [0, 1) for (seed, species, line, field).
static inline double u01(uint64_t seed, int species, int64_t i, int field) {
    uint64_t k = seed ^ ((uint64_t)(uint32_t)species << 48) ^ ((uint64_t)i << 4) ^ (uint64_t)field;
    return (double)(splitmix64(splitmix64(k)) >> 11) * 0x1.0p-53;
}

}  // namespace hk

extern "C" const char *hk_last_error(void) { return hk::g_err; }

extern "C" const char *hk_version(void) { return "heliosk-mi355x 0.1 (gfx950)"; }

extern "C" int hk_synth_lines(uint64_t seed, int species, int64_t n, int kind, double *nu,
                              double *S, double *EL, double *gamma, double *nexp) {
    HK_CHECK_ARG(n >= 0 && nu && S && EL && gamma && nexp, "hk_synth_lines: bad arguments");
    HK_CHECK_ARG(kind == 0 || kind == 1, "hk_synth_lines: kind %d", kind);
    auto work = [=](int64_t a, int64_t b) {
        using hk::u01;
        for (int64_t i = a; i < b; ++i) {
            nu[i] = 30000.0 * u01(seed, species, i, 0);
            double lg = -30.0 + 11.0 * u01(seed, species, i, 1);
            EL[i] = 20000.0 * u01(seed, species, i, 2);
            double g = 0.01 + 0.09 * u01(seed, species, i, 3);
            nexp[i] = 0.3 + 0.5 * u01(seed, species, i, 4);
            if (kind == 1) {
                g = pow(10.0, -4.0 + 4.0 * u01(seed, species, i, 3));
                if (u01(seed, species, i, 5) < 0.05) lg = -20.0 + 3.0 * u01(seed, species, i, 1);
            }
            S[i] = pow(10.0, lg);
            gamma[i] = g;
        }
    };
    const int nt = (int)std::max<int64_t>(1, std::min<int64_t>(16, n / 100000));
    std::vector<std::thread> th;
    for (int k = 0; k < nt; ++k) th.emplace_back(work, n * k / nt, n * (k + 1) / nt);
    for (auto &x : th) x.join();
    return HK_OK;
}
