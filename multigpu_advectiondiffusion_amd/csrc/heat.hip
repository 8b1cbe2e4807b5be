// heat.hip -- 3-D heat equation, 7-point Laplacian + forward Euler fused in
// one pass: u' = u + dt*L7(u) on the Dirichlet interior, boundary copied.
// L7 is Compute_Laplace3d_O2 (SingleGPU/Diffusion3d_baselineCode/
// kernels.cu:59-62) and the update is stage 1 of Compute_RK (:130-132), so a
// step reads u once and writes u' once (16 B per node) instead of the
// reference's Lu round trip.
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>

#include "mgad_common.hpp"

namespace mgad {

__global__ __launch_bounds__(256) void k_heat7_fe(const double *__restrict__ u,
                                                  double *__restrict__ v, int nx, int ny, int nz,
                                                  double kx, double ky, double kz, double dt) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    const unsigned xy = static_cast<unsigned>(nx) * ny;
    const unsigned o = i + static_cast<unsigned>(nx) * j + xy * k;
    const double uo = u[o];
    if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
        v[o] = uo;
        return;
    }
    const double Lu = 12 * kx * (u[o + 1] - 2 * uo + u[o - 1]) +
                      12 * ky * (u[o + nx] - 2 * uo + u[o - nx]) +
                      12 * kz * (u[o + xy] - 2 * uo + u[o - xy]);
    v[o] = uo + dt * Lu;
}

namespace {

int enqueue_heat(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                 double kz, double dt, double *u, double *w, int nsteps) {
    const dim3 grid((nx + 63) / 64, (ny + 3) / 4, nz);
    for (int it = 0; it < nsteps; it++) {
        const double *src = (it & 1) ? w : u;
        double *dst = (it & 1) ? u : w;
        hipLaunchKernelGGL(k_heat7_fe, grid, dim3(256), 0, s, src, dst, (int)nx, (int)ny, (int)nz,
                           kx, ky, kz, dt);
        const int st = check_launch();
        if (st) return st;
    }
    if (nsteps & 1)
        return check_hip(hipMemcpyAsync(u, w, sizeof(double) * nx * ny * nz,
                                        hipMemcpyDeviceToDevice, s));
    return MGAD_OK;
}

struct HeatKey {
    hipStream_t s;
    unsigned nx, ny, nz;
    double kx, ky, kz, dt;
    double *u, *w;
    int nsteps;
    bool operator<(const HeatKey &o) const {
        return std::tie(s, nx, ny, nz, kx, ky, kz, dt, u, w, nsteps) <
               std::tie(o.s, o.nx, o.ny, o.nz, o.kx, o.ky, o.kz, o.dt, o.u, o.w, o.nsteps);
    }
};

}  // namespace
}  // namespace mgad

using namespace mgad;

extern "C" int mgad_heat7_fe_steps(hipStream_t stream, unsigned int nx, unsigned int ny,
                                   unsigned int nz, double kx, double ky, double kz, double dt,
                                   double *u, double *w, int nsteps, int use_graph) {
    if (nx < 3 || ny < 3 || nz < 3 || 1ull * nx * ny * nz >= (1ull << 31) || !u || !w || u == w ||
        nsteps < 0) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    if (!use_graph || stream == nullptr || nsteps == 0)
        return enqueue_heat(stream, nx, ny, nz, kx, ky, kz, dt, u, w, nsteps);
    static std::map<HeatKey, hipGraphExec_t> cache;
    const HeatKey key{stream, nx, ny, nz, kx, ky, kz, dt, u, w, nsteps};
    auto it = cache.find(key);
    hipGraphExec_t exec = nullptr;
    if (it == cache.end()) {
        if (cache.size() > 16) {
            for (auto &kv : cache) (void)hipGraphExecDestroy(kv.second);
            cache.clear();
        }
        hipGraph_t g = nullptr;
        int st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        if (st) return st;
        const int est = enqueue_heat(stream, nx, ny, nz, kx, ky, kz, dt, u, w, nsteps);
        st = check_hip(hipStreamEndCapture(stream, &g));
        if (est) {
            if (g) (void)hipGraphDestroy(g);
            return est;
        }
        if (st) return st;
        st = check_hip(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
        (void)hipGraphDestroy(g);
        if (st) return st;
        cache[key] = exec;
    } else {
        exec = it->second;
    }
    return check_hip(hipGraphLaunch(exec, stream));
}
