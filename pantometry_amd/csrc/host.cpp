// Host-side pieces of the library: error reporting, device checks, the violation record, the
// seeded generator of crates/pantometry-core/src/rng.rs and the seeded workloads built on it.
//
// Seeded inputs are made here, on the host, with the host's libm (as Rust's f64::exp, ln and cos
// are on linux-gnu), and uploaded: the device never recomputes a transcendental.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"

namespace pm {

static thread_local std::string g_error;

void set_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_error = buf;
}

int require_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error("no HIP device is available (%s); this library has no CPU fallback",
                  e == hipSuccess ? "device count is zero" : hipGetErrorString(e));
        return PM_ERR_DEVICE;
    }
    int dev = 0;
    PM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PM_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s; this library is built for gfx950 only", dev,
                  prop.gcnArchName);
        return PM_ERR_DEVICE;
    }
    return PM_OK;
}

void fill_violation(pm_violation *v, const char *quantity, const char *site, double before,
                    double after, double scale, double tolerance, int64_t step) {
    if (!v) return;
    std::memset(v, 0, sizeof *v);
    std::snprintf(v->quantity, sizeof v->quantity, "%s", quantity);
    std::snprintf(v->site, sizeof v->site, "%s", site);
    v->before = before, v->after = after, v->scale = scale, v->tolerance = tolerance;
    v->step = step;
}

// Runs body(lo, hi) over [0, n) split into contiguous chunks on the host's cores.  Each index is
// computed independently, so the split never changes a value.
template <class F> static void parallel_chunks(size_t n, F body) {
    unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < (size_t)1 << 16) threads = 1;
    size_t chunk = (n + threads - 1) / threads;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < threads; t++) {
        size_t lo = t * chunk, hi = std::min(n, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([=] { body(lo, hi); });
    }
    for (auto &th : pool) th.join();
}

// SplitMix64 finaliser, rng.rs:37-42.
static inline uint64_t mix64(uint64_t z) {
    z = z + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

}  // namespace pm

using namespace pm;

extern "C" {

const char *pm_last_error(void) { return g_error.c_str(); }

int pm_device_info(char *name, size_t name_len, int *compute_units, size_t *hbm_bytes) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        set_error("no HIP device is available");
        return PM_ERR_DEVICE;
    }
    int dev = 0;
    PM_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    PM_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_len) std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    return PM_OK;
}

// Rng::new, rng.rs:49-55.
uint64_t pm_rng_new(uint64_t seed) { return seed == 0 ? 0x9E3779B97F4A7C15ULL : seed; }

// Rng::for_index, rng.rs:66-68.
uint64_t pm_rng_for_index(uint64_t seed, uint64_t index) {
    return pm_rng_new(mix64(seed ^ mix64(index)));
}

// xorshift64*, rng.rs:81-88.
static inline uint64_t next_u64(uint64_t *s) {
    uint64_t x = *s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1DULL;
}

// rng.rs:91-93.
double pm_rng_unit(uint64_t *s) { return (double)(next_u64(s) >> 11) / (double)(1ULL << 53); }

// rng.rs:96-98.
double pm_rng_range(uint64_t *s, double lo, double hi) { return lo + (hi - lo) * pm_rng_unit(s); }

// Box-Muller, rng.rs:146-151.
double pm_rng_gaussian(uint64_t *s) {
    double u1 = std::fmax(pm_rng_unit(s), 2.2250738585072014e-308);
    double u2 = pm_rng_unit(s);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// rng.rs:75-78.
uint64_t pm_rng_split(uint64_t *s) { return pm_rng_new(mix64(next_u64(s))); }

int pm_workload_hot_spot(size_t nx, size_t ny, size_t nz, double base, double amp, double sigma,
                         double noise, uint64_t noise_seed, double *cells) {
    if (!cells || nx == 0 || ny == 0 || nz == 0) {
        set_error("pm_workload_hot_spot: empty grid or null output");
        return PM_ERR_ARGUMENT;
    }
    const double hx = 0.5 * (double)nx, hy = 0.5 * (double)ny, hz = 0.5 * (double)nz;
    const double two_s2 = 2.0 * sigma * sigma;
    parallel_chunks(nx * ny * nz, [=](size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; c++) {
            size_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
            double ddx = ((double)i + 0.5) - hx, ddy = ((double)j + 0.5) - hy,
                   ddz = ((double)k + 0.5) - hz;
            double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
            uint64_t s = pm_rng_for_index(noise_seed, c);
            double t = base + amp * std::exp(-r2 / two_s2);
            cells[c] = t + pm_rng_range(&s, -noise, noise);
        }
    });
    return PM_OK;
}

int pm_workload_heterogeneous(size_t nx, size_t ny, size_t nz, uint64_t k_seed,
                              uint64_t spot_seed, int spots, double *conductivity,
                              double *cells) {
    if (!cells || !conductivity || nx == 0 || ny == 0 || nz == 0 || spots < 0 || spots > 64) {
        set_error("pm_workload_heterogeneous: empty grid, null output or spots outside 0..64");
        return PM_ERR_ARGUMENT;
    }
    struct Spot { double amp, two_s2, cx, cy, cz; };
    std::vector<Spot> sp((size_t)spots);
    for (int m = 0; m < spots; m++) {
        uint64_t s = pm_rng_for_index(spot_seed, (uint64_t)m);
        double amp = pm_rng_range(&s, 50.0, 200.0);
        double sig = pm_rng_range(&s, 4.0, 16.0);
        double cx = pm_rng_range(&s, 0.1 * (double)nx, 0.9 * (double)nx);
        double cy = pm_rng_range(&s, 0.1 * (double)ny, 0.9 * (double)ny);
        double cz = pm_rng_range(&s, 0.1 * (double)nz, 0.9 * (double)nz);
        sp[(size_t)m] = {amp, 2.0 * sig * sig, cx, cy, cz};
    }
    const Spot *spp = sp.data();
    parallel_chunks(nx * ny * nz, [=](size_t lo, size_t hi) {
        for (size_t c = lo; c < hi; c++) {
            size_t i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
            uint64_t s = pm_rng_for_index(k_seed, c);
            conductivity[c] = std::exp(0.0 + 0.5 * pm_rng_gaussian(&s));
            double t = 300.0;
            for (int m = 0; m < spots; m++) {
                double ddx = ((double)i + 0.5) - spp[m].cx, ddy = ((double)j + 0.5) - spp[m].cy,
                       ddz = ((double)k + 0.5) - spp[m].cz;
                double r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                t = t + spp[m].amp * std::exp(-r2 / spp[m].two_s2);
            }
            cells[c] = t;
        }
    });
    return PM_OK;
}

}  // extern "C"
