// host.cpp -- host-only parts of the library: sticky error code, the
// CopyToConstantMemory coefficients, and initial-condition generators.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/mgad.h"

namespace mgad {

namespace {
std::atomic<int> g_last_error{0};
double g_kx = 0.0, g_ky = 0.0, g_kz = 0.0;
}  // namespace

struct DiffCoef { double kx, ky, kz; };

std::recursive_mutex &host_mutex() {
    static std::recursive_mutex m;
    return m;
}

void set_error(int code) {
    int expected = 0;
    g_last_error.compare_exchange_strong(expected, code);
}

int check_hip(hipError_t e) {
    if (e == hipSuccess) return MGAD_OK;
    std::fprintf(stderr, "mgad: HIP error %d (%s)\n", static_cast<int>(e), hipGetErrorString(e));
    set_error(MGAD_ERR_HIP);
    return MGAD_ERR_HIP;
}

int check_launch() { return check_hip(hipGetLastError()); }

DiffCoef diff_coef() { return DiffCoef{g_kx, g_ky, g_kz}; }

// SplitMix64 (Steele, Lea, Flood 2014) -> uniform double in [0,1).
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

template <class F>
void parallel_planes(unsigned nz, F &&f) {
    const unsigned nt = std::max(1u, std::min(nz, std::min(32u, std::thread::hardware_concurrency())));
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < nt; t++)
        ts.emplace_back([&, t] {
            for (unsigned k = t; k < nz; k += nt) f(k);
        });
    for (auto &th : ts) th.join();
}

}  // namespace mgad

using namespace mgad;

extern "C" int mgad_get_last_error(void) { return g_last_error.exchange(0); }

extern "C" const char *mgad_error_string(int code) {
    switch (code) {
    case MGAD_OK: return "ok";
    case MGAD_ERR_INVALID_ARG: return "invalid argument";
    case MGAD_ERR_UNSUPPORTED: return "unsupported grid shape";
    case MGAD_ERR_HIP: return "HIP runtime error";
    case MGAD_ERR_NONFINITE: return "non-finite max|u|";
    case MGAD_ERR_SCRATCH_FULL: return "work list larger than the static device scratch";
    default: return "unknown error";
    }
}

extern "C" const char *mgad_version(void) { return "mgad 0.1.0 (gfx950)"; }

extern "C" void CopyToConstantMemory(double kx, double ky, double kz) {
    g_kx = kx;
    g_ky = ky;
    g_kz = kz;
}

extern "C" int mgad_ic_fourier(double *u, unsigned int nx, unsigned int ny, unsigned int nz,
                               double lo, double hi, uint64_t seed, int nmodes, int kmax,
                               double amp, double offset) {
    if (!u || nx == 0 || ny == 0 || nz == 0 || nmodes < 0 || kmax < 1 || !(hi > lo)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    struct Mode { double kx, ky, kz, a, phi; };
    std::vector<Mode> modes(nmodes);
    SplitMix64 rng{seed};
    const double pi = 3.14159265358979323846;
    for (auto &m : modes) {
        m.kx = 1.0 + std::floor(kmax * rng.uniform());
        m.ky = 1.0 + std::floor(kmax * rng.uniform());
        m.kz = 1.0 + std::floor(kmax * rng.uniform());
        m.a = amp * rng.uniform();
        m.phi = 2.0 * pi * rng.uniform();
    }
    const double dx = (hi - lo) / nx, dy = (hi - lo) / ny, dz = (hi - lo) / nz;
    parallel_planes(nz, [&](unsigned k) {
        const double z = lo + k * dz;
        for (unsigned j = 0; j < ny; j++) {
            const double y = lo + j * dy;
            double *row = u + (static_cast<size_t>(k) * ny + j) * nx;
            for (unsigned i = 0; i < nx; i++) {
                const double x = lo + i * dx;
                double s = offset;
                for (const auto &m : modes) s += m.a * std::sin(pi * (m.kx * x + m.ky * y + m.kz * z) + m.phi);
                row[i] = s;
            }
        }
    });
    return MGAD_OK;
}

extern "C" int mgad_ic_gauss_noise(double *u, unsigned int nx, unsigned int ny, unsigned int nz,
                                   double sigma, double noise, uint64_t seed) {
    if (!u || nx < 2 || ny < 2 || nz < 2 || !(sigma > 0)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const double dx = 1.0 / (nx - 1), dy = 1.0 / (ny - 1), dz = 1.0 / (nz - 1);
    // noise drawn serially in storage order so the field does not depend on threading
    SplitMix64 rng{seed};
    const size_t n = static_cast<size_t>(nx) * ny * nz;
    for (size_t q = 0; q < n; q++) u[q] = noise * (2.0 * rng.uniform() - 1.0);
    parallel_planes(nz, [&](unsigned k) {
        const double z = k * dz - 0.5;
        for (unsigned j = 0; j < ny; j++) {
            const double y = j * dy - 0.5;
            double *row = u + (static_cast<size_t>(k) * ny + j) * nx;
            for (unsigned i = 0; i < nx; i++) {
                const double x = i * dx - 0.5;
                const bool edge = i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1;
                row[i] = edge ? 0.0 : std::exp(-(x * x + y * y + z * z) / (2.0 * sigma * sigma)) + row[i];
            }
        }
    });
    return MGAD_OK;
}
