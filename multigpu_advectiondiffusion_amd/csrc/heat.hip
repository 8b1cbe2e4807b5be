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

// One forward-Euler step, u' = u + dt*L7(u) on the Dirichlet interior and
// u' = u on the boundary nodes.  Each thread marches a chunk of one (x, y)
// column holding u of planes k-1, k, k+1 in registers, so every node is read
// from memory once per step by its own thread; x and y neighbours are the
// same plane's loads of the adjacent lanes and rows, served by the caches.
__global__ __launch_bounds__(256) void k_heat7_fe(const double *__restrict__ u, double *__restrict__ v, int nx,
                                                  int ny, int nz, int kchunk, double kx, double ky, double kz,
                                                  double dt) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z0 = blockIdx.z * kchunk, z1 = min(z0 + kchunk, nz);
    if (i >= nx || j >= ny) return;
    const unsigned xy = static_cast<unsigned>(nx) * ny, col = i + static_cast<unsigned>(nx) * j;
    const bool side = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
    // 12*kx etc. as the source's 12 * kx * (...) evaluates them
    const double ckx = 12 * kx, cky = 12 * ky, ckz = 12 * kz;
    double below = z0 > 0 ? u[col + xy * static_cast<unsigned>(z0 - 1)] : 0.0;
    double cur = u[col + xy * static_cast<unsigned>(z0)];
#pragma unroll 4
    for (int k = z0; k < z1; k++) {
        const unsigned o = col + xy * static_cast<unsigned>(k);
        const double above = k + 1 < nz ? u[o + xy] : 0.0;
        if (side || k == 0 || k == nz - 1) {
            v[o] = cur;
        } else {
            // (u_e - 2 u_o) + u_w etc., the 2*u_o product exact
            const double Lu = ckx * (p2fma(-2.0, cur, u[o + 1]) + u[o - 1]) +
                              cky * (p2fma(-2.0, cur, u[o + nx]) + u[o - nx]) +
                              ckz * (p2fma(-2.0, cur, above) + below);
            v[o] = cur + dt * Lu;
        }
        below = cur;
        cur = above;
    }
}

// One SSP-RK3 stage of the reference heat step fused: the 13-point Laplacian
// of Compute_Laplace3d (Diffusion3d_baselineCode/kernels.cu:94-97, with the
// CopyToConstantMemory coefficients) and Compute_RK (kernels.cu:127-140) on
// the interior 2..n-3 of every axis, Lu kept in a register.  Each thread
// marches a chunk of one (x, y) column with a 5-plane queue of u for the z
// neighbours; x and y neighbours come through the caches.
__global__ __launch_bounds__(256) void k_heat13_rk(const double *__restrict__ uin, const double *un, double *uout,
                                                   int nx, int ny, int nz, int kchunk, double kx, double ky,
                                                   double kz, double dt, int stage) {
    const int i = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z0 = 2 + blockIdx.z * kchunk, z1 = min(z0 + kchunk, nz - 2);
    if (i >= nx - 2 || j >= ny - 2) return;
    const unsigned xy = static_cast<unsigned>(nx) * ny, col = i + static_cast<unsigned>(nx) * j;
    const unsigned px = nx, px2 = 2u * nx;
    double q[5];  // u at planes k-2..k+2
#pragma unroll
    for (int t = 0; t < 4; t++) q[t] = uin[col + xy * static_cast<unsigned>(z0 - 2 + t)];
    for (int k = z0; k < z1; k++) {
        const unsigned o = col + xy * static_cast<unsigned>(k);
        q[4] = uin[o + 2u * xy];
        const double c = q[2];
        const double X = d4(uin[o + 2], uin[o + 1], c, uin[o - 1], uin[o - 2]);
        const double Y = d4(uin[o + px2], uin[o + px], c, uin[o - px], uin[o - px2]);
        const double Z = d4(q[4], q[3], c, q[1], q[0]);
        const double Lu = (kx * X + ky * Y) + kz * Z;
        const double uo = stage == 1 ? c : un[o];
        uout[o] = rk_stage(stage, dt, c, uo, Lu);
#pragma unroll
        for (int t = 0; t < 4; t++) q[t] = q[t + 1];
    }
}

namespace {

int enqueue_heat13(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, DiffCoef c, double dt, double *u,
                   double *w1, double *w2) {
    const int kchunk = 16;
    const dim3 grid((nx - 4 + 63) / 64, (ny - 4 + 3) / 4, (nz - 4 + kchunk - 1) / kchunk);
    hipLaunchKernelGGL(k_heat13_rk, grid, dim3(256), 0, s, u, u, w1, (int)nx, (int)ny, (int)nz, kchunk, c.kx, c.ky,
                       c.kz, dt, 1);
    hipLaunchKernelGGL(k_heat13_rk, grid, dim3(256), 0, s, w1, u, w2, (int)nx, (int)ny, (int)nz, kchunk, c.kx, c.ky,
                       c.kz, dt, 2);
    hipLaunchKernelGGL(k_heat13_rk, grid, dim3(256), 0, s, w2, u, u, (int)nx, (int)ny, (int)nz, kchunk, c.kx, c.ky,
                       c.kz, dt, 3);
    return check_launch();
}

int enqueue_heat(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                 double kz, double dt, double *u, double *w, int nsteps) {
    // z chunks: enough blocks to fill the device several times over
    const long bxy = static_cast<long>((nx + 63) / 64) * ((ny + 3) / 4);
    int kc = static_cast<int>(nz);
    while (kc > 1 && bxy * ((nz + kc - 1) / kc) < 4096) kc = (kc + 1) / 2;
    const dim3 grid((nx + 63) / 64, (ny + 3) / 4, (nz + kc - 1) / kc);
    for (int it = 0; it < nsteps; it++) {
        const double *src = (it & 1) ? w : u;
        double *dst = (it & 1) ? u : w;
        hipLaunchKernelGGL(k_heat7_fe, grid, dim3(256), 0, s, src, dst, (int)nx, (int)ny, (int)nz, kc,
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

// The reference's heat step fused (SingleGPU/Diffusion3d_baselineCode/
// main.cpp:108-121): per step uo <- u, then for rk = 1..3 Call_Lu3d and
// Call_RK3d(rk), one launch per stage; the two ghost layers are not written.
extern "C" int mgad_heat13_rk3_steps(hipStream_t stream, unsigned int nx, unsigned int ny, unsigned int nz, double dt,
                                     double *u, double *w1, double *w2, int nsteps, int use_graph) {
    if (nx < 5 || ny < 5 || nz < 5 || 1ull * nx * ny * nz >= (1ull << 31) || !u || !w1 || !w2 || u == w1 ||
        u == w2 || w1 == w2 || nsteps < 0) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const size_t bytes = sizeof(double) * nx * ny * nz;
    int st = check_hip(hipMemcpyAsync(w1, u, bytes, hipMemcpyDeviceToDevice, stream));
    if (!st) st = check_hip(hipMemcpyAsync(w2, u, bytes, hipMemcpyDeviceToDevice, stream));
    if (st || nsteps == 0) return st;
    const DiffCoef c = diff_coef();
    if (!use_graph || stream == nullptr) {
        for (int s = 0; s < nsteps && !st; s++) st = enqueue_heat13(stream, nx, ny, nz, c, dt, u, w1, w2);
        return st;
    }
    using Key = std::tuple<hipStream_t, unsigned, unsigned, unsigned, double, double, double, double, double *,
                           double *, double *>;
    static std::map<Key, hipGraphExec_t> cache;
    const Key key{stream, nx, ny, nz, c.kx, c.ky, c.kz, dt, u, w1, w2};
    hipGraphExec_t exec = nullptr;
    auto it = cache.find(key);
    if (it == cache.end()) {
        if (cache.size() > 16) {
            for (auto &kv : cache) (void)hipGraphExecDestroy(kv.second);
            cache.clear();
        }
        hipGraph_t g = nullptr;
        if ((st = check_hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal)))) return st;
        const int est = enqueue_heat13(stream, nx, ny, nz, c, dt, u, w1, w2);
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
    for (int s = 0; s < nsteps && !st; s++) st = check_hip(hipGraphLaunch(exec, stream));
    return st;
}
