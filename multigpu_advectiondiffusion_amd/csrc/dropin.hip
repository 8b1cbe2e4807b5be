// dropin.hip -- the reference's device-wrapper entry points with their
// original signatures and update sets (frozen ghost layers), re-implemented
// for gfx950.
//
// Sweeps (Call_Adv_x/y/z): one thread marches a chunk of one grid line,
// computing each face flux once and reusing it for the two nodes that share
// it, as Compute_dF/dG/dH do (SingleGPU/Burgers3d_WENO5/kernels.cu:197-394).
// Lines are split into chunks so that even the short x lines of a 256^3 grid
// give every CU several waves; a chunk starts by computing the face on its
// low side (the reference's fu_old), except at the line start where fu_old is
// 0 as in the reference (kernels.cu:222).  For the y and z sweeps adjacent
// lanes own adjacent x columns, so every load and store is coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mgad_common.hpp"

namespace mgad {

namespace {

struct AdvScheme {
    int weno = kWenoZ;  // reference: WENO-Z (kernels.cu:241)
    int p = 1;          // reference: tau5/(B+EPS) to the first power (:163)
    double eps = 1e-6;  // Burgers.h:26
    int lf = MGAD_LF_LOCAL;
    double alpha = 0.0;
};
AdvScheme g_adv;

constexpr int kChunk = 32;  // faces per thread and line chunk

}  // namespace

// AXIS 0: x lines (rows), 1: y lines, 2: z lines.  MODE 0 writes Lu, 1 subtracts.
// px is the row pitch in elements (px >= nx; px = nx for unpitched arrays).
template <int AXIS, int WENO, int P, bool LOCAL>
__global__ __launch_bounds__(256) void k_adv_sweep(const double *__restrict__ u,
                                                   double *__restrict__ Lu, int nx, int ny, int nz,
                                                   int px, double d, double eps, double alpha,
                                                   int nchunks) {
    // transverse extents (interior 3..n-4) and line length
    const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const int na = (AXIS == 0 ? ny : nx) - 6;        // first transverse coordinate count
    const int nb = (AXIS == 2 ? ny : nz) - 6;        // second transverse coordinate count
    const long gid = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
    const long nlines = static_cast<long>(na) * nb;
    if (gid >= nlines * nchunks) return;
    // lanes run fastest over the first transverse coordinate (x for y/z sweeps)
    const int a = 3 + static_cast<int>(gid % na);
    const long rest = gid / na;
    const int b = 3 + static_cast<int>(rest % nb);
    const int c = static_cast<int>(rest / nb);

    const unsigned plane = static_cast<unsigned>(px) * ny;
    unsigned base, st;
    if (AXIS == 0) {
        base = static_cast<unsigned>(px) * a + plane * b;
        st = 1;
    } else if (AXIS == 1) {
        base = static_cast<unsigned>(a) + plane * b;
        st = px;
    } else {
        base = static_cast<unsigned>(a) + static_cast<unsigned>(px) * b;
        st = plane;
    }
    // faces i+1/2 for i in [i0, i1): the reference marches i = 2..n-4
    const int i0 = 2 + c * kChunk;
    const int i1 = min(i0 + kChunk, n - 3);
    if (i0 >= i1) return;

    auto split = [&](double v, double &fp, double &fm) {
        if (LOCAL)
            lf_split_local(v, fp, fm);
        else
            lf_split(v, alpha, fp, fm);
    };
    // queue: fp[q], fm[q] hold node i-3+q for the face being computed
    // (node i0-3 exists only for chunks after the first; the first chunk's
    // low face is the reference's fu_old = 0 and never reads it)
    double fp[7], fm[7];
    fp[0] = fm[0] = 0.0;
    if (i0 > 2) split(u[base + st * static_cast<unsigned>(i0 - 3)], fp[0], fm[0]);
#pragma unroll
    for (int q = 1; q < 6; q++) split(u[base + st * static_cast<unsigned>(i0 - 3 + q)], fp[q], fm[q]);
    double fu_old = 0.0;
    if (i0 > 2)
        fu_old = weno_face<WENO, P>(fp[0], fp[1], fp[2], fp[3], fp[4], fm[1], fm[2], fm[3], fm[4],
                                    fm[5], eps);
    for (int i = i0; i < i1; i++) {
        split(u[base + st * static_cast<unsigned>(i + 3)], fp[6], fm[6]);
        const double fu = weno_face<WENO, P>(fp[1], fp[2], fp[3], fp[4], fp[5], fm[2], fm[3],
                                             fm[4], fm[5], fm[6], eps);
        const unsigned o = base + st * static_cast<unsigned>(i);
        if (AXIS == 0)
            Lu[o] = -(fu - fu_old) / d;
        else
            Lu[o] -= (fu - fu_old) / d;
        fu_old = fu;
#pragma unroll
        for (int q = 0; q < 6; q++) {
            fp[q] = fp[q + 1];
            fm[q] = fm[q + 1];
        }
    }
}

// Compute_Laplace_Async arithmetic, kernels.cu:474-476: Lu += 13-point operator.
__global__ __launch_bounds__(256) void k_diff13_b5(const double *__restrict__ u,
                                                   double *__restrict__ Lu, int nx, int ny, int nz,
                                                   int pitch, double kx, double ky, double kz) {
    const int i = 3 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 3 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 3 + blockIdx.z;
    if (i >= nx - 3 || j >= ny - 3 || k >= nz - 2) return;
    const unsigned px = pitch, px2 = 2u * px, XY = px * ny, XY2 = 2u * XY;
    const unsigned o = i + px * j + XY * k;
    const double center = u[o];
    const double below2 = u[o - XY2], below = u[o - XY], above = u[o + XY], above2 = u[o + XY2];
    Lu[o] += kx * (-u[o - 2] + 16 * u[o - 1] - 30 * center + 16 * u[o + 1] - u[o + 2]) +
             ky * (-u[o - px2] + 16 * u[o - px] - 30 * center + 16 * u[o + px] - u[o + px2]) +
             kz * (-below2 + 16 * below - 30 * center + 16 * above - above2);
}

// SSP-RK3 stage on the interior [lo, n-lo) of every axis: lo = 3 is
// Compute_RK of the Burgers code (kernels.cu:518), lo = 2 the heat one
// (Diffusion3d_baselineCode/kernels.cu:127).
__global__ __launch_bounds__(256) void k_rk_interior(double *u, const double *uo,
                                                     const double *__restrict__ Lu, int nx, int ny,
                                                     int nz, int px, int lo, int step, double dt) {
    const int i = lo + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = lo + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = lo + blockIdx.z;
    if (i >= nx - lo || j >= ny - lo || k >= nz - lo) return;
    const unsigned o = i + static_cast<unsigned>(px) * j + static_cast<unsigned>(px) * ny * k;
    u[o] = rk_stage(step, dt, u[o], uo[o], Lu[o]);
}

// Compute_Laplace3d, Diffusion3d_baselineCode/kernels.cu:94-97 (13-point, Lu =).
__global__ __launch_bounds__(256) void k_lu3d_13(const double *__restrict__ u,
                                                 double *__restrict__ Lu, int nx, int ny, int nz,
                                                 int px, double kx, double ky, double kz) {
    const int i = 2 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 2 + blockIdx.z;
    if (i >= nx - 2 || j >= ny - 2 || k >= nz - 2) return;
    const unsigned xy = static_cast<unsigned>(px) * ny;
    const unsigned o = i + static_cast<unsigned>(px) * j + xy * k;
    const unsigned nn = o + px + px, n = o + px, s = o - px, ss = o - px - px;
    const unsigned ee = o + 2, e = o + 1, w = o - 1, ww = o - 2;
    const unsigned tt = o + xy + xy, t = o + xy, b = o - xy, bb = o - xy - xy;
    Lu[o] = kx * (-u[ee] + 16 * u[e] - 30 * u[o] + 16 * u[w] - u[ww]) +
            ky * (-u[nn] + 16 * u[n] - 30 * u[o] + 16 * u[s] - u[ss]) +
            kz * (-u[tt] + 16 * u[t] - 30 * u[o] + 16 * u[b] - u[bb]);
}

// Compute_Laplace3d_O2, Diffusion3d_baselineCode/kernels.cu:59-62 (7-point, Lu =).
__global__ __launch_bounds__(256) void k_lu3d_o2(const double *__restrict__ u,
                                                 double *__restrict__ Lu, int nx, int ny, int nz,
                                                 int px, double kx, double ky, double kz) {
    const int i = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = 1 + blockIdx.z;
    if (i >= nx - 1 || j >= ny - 1 || k >= nz - 1) return;
    const unsigned xy = static_cast<unsigned>(px) * ny;
    const unsigned o = i + static_cast<unsigned>(px) * j + xy * k;
    const unsigned n = o + px, s = o - px, e = o + 1, w = o - 1, t = o + xy, b = o - xy;
    Lu[o] = 12 * kx * (u[e] - 2 * u[o] + u[w]) + 12 * ky * (u[n] - 2 * u[o] + u[s]) +
            12 * kz * (u[t] - 2 * u[o] + u[b]);
}

namespace {

template <int AXIS, int WENO, int P>
void launch_sweep_t(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, double d,
                    double *u, double *Lu) {
    const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const long na = (AXIS == 0 ? ny : nx) - 6, nb = (AXIS == 2 ? ny : nz) - 6;
    const int nchunks = (n - 5 + kChunk - 1) / kChunk;  // faces 2..n-4
    const long total = na * nb * nchunks;
    const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
    if (g_adv.lf == MGAD_LF_LOCAL)
        hipLaunchKernelGGL((k_adv_sweep<AXIS, WENO, P, true>), dim3(blocks), dim3(256), 0, s, u, Lu,
                           (int)nx, (int)ny, (int)nz, (int)px, d, g_adv.eps, g_adv.alpha, nchunks);
    else
        hipLaunchKernelGGL((k_adv_sweep<AXIS, WENO, P, false>), dim3(blocks), dim3(256), 0, s, u,
                           Lu, (int)nx, (int)ny, (int)nz, (int)px, d, g_adv.eps, g_adv.alpha,
                           nchunks);
}

// Arrays of px*ny*nz elements (row pitch px >= nx) addressable with 32-bit offsets.
bool grid_ok(unsigned nx, unsigned ny, unsigned nz, unsigned px, unsigned minn) {
    return nx >= minn && ny >= minn && nz >= minn && px >= nx && 1ull * px * ny * nz < (1ull << 31);
}

template <int AXIS>
void launch_sweep(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, double d,
                  double *u, double *Lu) {
    if (!grid_ok(nx, ny, nz, px, 7) || !u || !Lu) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    if (g_adv.weno == MGAD_WENO_JS)
        launch_sweep_t<AXIS, kWenoJS, 2>(s, nx, ny, nz, px, d, u, Lu);
    else if (g_adv.p == 2)
        launch_sweep_t<AXIS, kWenoZ, 2>(s, nx, ny, nz, px, d, u, Lu);
    else
        launch_sweep_t<AXIS, kWenoZ, 1>(s, nx, ny, nz, px, d, u, Lu);
    check_launch();
}

void diff13(hipStream_t s, unsigned Nx, unsigned Ny, unsigned Nz, unsigned px, double *q, double *Lq) {
    if (!grid_ok(Nx, Ny, Nz, px, 7) || !q || !Lq) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    const DiffCoef c = diff_coef();
    const dim3 grid((Nx - 6 + 63) / 64, (Ny - 6 + 3) / 4, Nz - 5);
    hipLaunchKernelGGL(k_diff13_b5, grid, dim3(256), 0, s, q, Lq, (int)Nx, (int)Ny, (int)Nz, (int)px,
                       c.kx, c.ky, c.kz);
    check_launch();
}

void rk_burgers(hipStream_t s, unsigned Nx, unsigned Ny, unsigned Nz, unsigned px, unsigned step,
                double dt, double *u, double *uo, double *Lu) {
    if (!grid_ok(Nx, Ny, Nz, px, 7) || step < 1 || step > 3 || !u || !uo || !Lu) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    const dim3 grid((Nx - 6 + 63) / 64, (Ny - 6 + 3) / 4, Nz - 6);
    hipLaunchKernelGGL(k_rk_interior, grid, dim3(256), 0, s, u, uo, Lu, (int)Nx, (int)Ny, (int)Nz,
                       (int)px, 3, (int)step, dt);
    check_launch();
}

void lu3d(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, double *u, double *Lu) {
    if (!grid_ok(nx, ny, nz, px, 5) || !u || !Lu) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    const DiffCoef c = diff_coef();
    const dim3 grid((nx - 4 + 63) / 64, (ny - 4 + 3) / 4, nz - 4);
    hipLaunchKernelGGL(k_lu3d_13, grid, dim3(256), 0, s, u, Lu, (int)nx, (int)ny, (int)nz, (int)px,
                       c.kx, c.ky, c.kz);
    check_launch();
}

void rk_heat(hipStream_t s, unsigned step, unsigned nx, unsigned ny, unsigned nz, unsigned px,
             double dt, double *u, double *uo, double *Lu) {
    if (!grid_ok(nx, ny, nz, px, 5) || step < 1 || step > 3 || !u || !uo || !Lu) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    const dim3 grid((nx - 4 + 63) / 64, (ny - 4 + 3) / 4, nz - 4);
    hipLaunchKernelGGL(k_rk_interior, grid, dim3(256), 0, s, u, uo, Lu, (int)nx, (int)ny, (int)nz,
                       (int)px, 2, (int)step, dt);
    check_launch();
}

}  // namespace

}  // namespace mgad

using namespace mgad;

extern "C" int mgad_set_adv_scheme(int weno, int p, double eps, int lf, double alpha) {
    if ((weno != MGAD_WENO_Z && weno != MGAD_WENO_JS) || (p != 1 && p != 2) ||
        (lf != MGAD_LF_LOCAL && lf != MGAD_LF_GLOBAL) || !(eps >= 0.0)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    g_adv.weno = weno;
    g_adv.p = p;
    g_adv.eps = eps;
    g_adv.lf = lf;
    g_adv.alpha = alpha;
    return MGAD_OK;
}

// ---- unpitched (SingleGPU/Burgers3d_WENO5, Diffusion3d_baselineCode) ----

extern "C" void Call_Adv_x(dim3, dim3, hipStream_t aStream, unsigned int Nx, unsigned int Ny,
                           unsigned int Nz, double dx, double *u, double *Lu) {
    launch_sweep<0>(aStream, Nx, Ny, Nz, Nx, dx, u, Lu);
}

extern "C" void Call_Adv_y(dim3, dim3, hipStream_t aStream, unsigned int Nx, unsigned int Ny,
                           unsigned int Nz, double dy, double *v, double *Lv) {
    launch_sweep<1>(aStream, Nx, Ny, Nz, Nx, dy, v, Lv);
}

extern "C" void Call_Adv_z(dim3, dim3, hipStream_t aStream, unsigned int Nx, unsigned int Ny,
                           unsigned int Nz, double dz, double *w, double *Lw) {
    launch_sweep<2>(aStream, Nx, Ny, Nz, Nx, dz, w, Lw);
}

extern "C" void Call_Diff_(dim3, dim3, hipStream_t aStream, unsigned int Nx, unsigned int Ny,
                           unsigned int Nz, double *q, double *Lq) {
    diff13(aStream, Nx, Ny, Nz, Nx, q, Lq);
}

extern "C" void Call_sspRK(dim3, dim3, hipStream_t aStream, unsigned int Nx, unsigned int Ny,
                           unsigned int Nz, unsigned int step, double dt, double *u, double *uo,
                           double *Lu) {
    rk_burgers(aStream, Nx, Ny, Nz, Nx, step, dt, u, uo, Lu);
}

extern "C" void Call_Lu3d(dim3, dim3, hipStream_t aStream, unsigned int nx, unsigned int ny,
                          unsigned int nz, double *u, double *Lu) {
    lu3d(aStream, nx, ny, nz, nx, u, Lu);
}

extern "C" void Call_RK3d(dim3, dim3, hipStream_t aStream, unsigned int step, unsigned int nx,
                          unsigned int ny, unsigned int nz, double dt, double *u, double *uo,
                          double *Lu) {
    rk_heat(aStream, step, nx, ny, nz, nx, dt, u, uo, Lu);
}

// ---- pitched (SingleGPU/Burgers3d_WENO5_PitchedMem, _SharedMem,
//      Diffusion3d_PitchedMem, Diffusion3d_Blocking): pitch in elements ----

extern "C" void Call_Adv_x_pitched(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                   unsigned int Ny, unsigned int Nz, double dx, double *u, double *Lu) {
    launch_sweep<0>(aStream, Nx, Ny, Nz, pitch, dx, u, Lu);
}

extern "C" void Call_Adv_y_pitched(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                   unsigned int Ny, unsigned int Nz, double dy, double *v, double *Lv) {
    launch_sweep<1>(aStream, Nx, Ny, Nz, pitch, dy, v, Lv);
}

extern "C" void Call_Adv_z_pitched(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                   unsigned int Ny, unsigned int Nz, double dz, double *w, double *Lw) {
    launch_sweep<2>(aStream, Nx, Ny, Nz, pitch, dz, w, Lw);
}

extern "C" void Call_Diff_pitched(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                  unsigned int Ny, unsigned int Nz, double *q, double *Lq) {
    diff13(aStream, Nx, Ny, Nz, pitch, q, Lq);
}

extern "C" void Call_sspRK_pitched(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                   unsigned int Ny, unsigned int Nz, unsigned int step, double dt,
                                   double *u, double *uo, double *Lu) {
    rk_burgers(aStream, Nx, Ny, Nz, pitch, step, dt, u, uo, Lu);
}

extern "C" void Call_Lu3d_pitched(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int nx,
                                  unsigned int ny, unsigned int nz, double *u, double *Lu) {
    lu3d(aStream, nx, ny, nz, pitch, u, Lu);
}

extern "C" void Call_RK3d_pitched(dim3, dim3, hipStream_t aStream, unsigned int step, unsigned int pitch,
                                  unsigned int nx, unsigned int ny, unsigned int nz, double dt,
                                  double *u, double *uo, double *Lu) {
    rk_heat(aStream, step, nx, ny, nz, pitch, dt, u, uo, Lu);
}

extern "C" int mgad_lu3d_o2(hipStream_t stream, unsigned int nx, unsigned int ny, unsigned int nz,
                            const double *u, double *Lu) {
    if (!grid_ok(nx, ny, nz, nx, 3) || !u || !Lu) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const DiffCoef c = diff_coef();
    const dim3 grid((nx - 2 + 63) / 64, (ny - 2 + 3) / 4, nz - 2);
    hipLaunchKernelGGL(k_lu3d_o2, grid, dim3(256), 0, stream, u, Lu, (int)nx, (int)ny, (int)nz,
                       (int)nx, c.kx, c.ky, c.kz);
    return check_launch();
}
