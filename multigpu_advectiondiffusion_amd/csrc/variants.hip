// variants.hip -- drop-ins for the reference drivers whose wrappers keep the
// base code's names but differ from it in update set or rounding:
//
//  - SingleGPU/Burgers3d_WENO5_SharedMem: Call_Adv_x/y/z update every node
//    of the array, read nodes beyond the array as 0, and scale the WENO-Z
//    weights by *(1/6 * (1/sum)) (kernels.cu:153-205, :212-402);
//  - SingleGPU/Diffusion3d_Blocking: Call_Lu3d writes the 13-point operator
//    on nodes 3..n-4 and planes 3..nz-3 only, each axis summed from the low
//    side (kernels.cu:37-88, :187-191);
//  - MultiGPU/Diffusion3d_Baseline: Call_Diff_ (LaplaceO4_async, or the
//    per-axis second-order LaplaceO2_async its wrapper has commented out)
//    over a plane range of one rank's slab, and Call_sspRK on every plane
//    of the slab (Kernels.cu:152-260, :266-297, :401-414).
//
// They take the reference wrappers' arguments under suffixed names (one
// library cannot export two signatures, or two meanings, under one name).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mgad_common.hpp"

namespace mgad {

namespace {

constexpr int kB5sRow = 63;   // x sweep: nodes per wave (64 faces)
constexpr int kB5sChunk = 32;  // y/z sweeps: nodes per thread and line chunk

__device__ __forceinline__ bool wave_all_ok(bool ok) { return __builtin_amdgcn_ballot_w64(!ok) == 0; }

// Split fluxes of node value v (local LF, kernels.cu:159-163 of the
// shared-memory code: 0.5*(f(u) +- |u| u)).
__device__ __forceinline__ double2 split_local(double v) {
    double fp, fm;
    lf_split_local(v, fp, fm);
    return make_double2(fp, fm);
}

// Face below the node whose stencil is a0..a5 (split fluxes of nodes
// l-3..l+2): f+ of l-3..l+1, f- of l-2..l+2.  fd: the shortened weight
// divisions apply (every node read is at most 2^40 in magnitude, eps in
// range), whose results equal IEEE division.
template <int P>
__device__ __forceinline__ double face_b5s(bool fd, double2 a0, double2 a1, double2 a2, double2 a3, double2 a4,
                                           double2 a5, double eps) {
    if (fd)
        return weno_face2<kWenoZB5S, P, true>(a0.x, a1.x, a2.x, a3.x, a4.x, a1.y, a2.y, a3.y, a4.y, a5.y, eps);
    return weno_face2<kWenoZB5S, P, false>(a0.x, a1.x, a2.x, a3.x, a4.x, a1.y, a2.y, a3.y, a4.y, a5.y, eps);
}

}  // namespace

// Compute_dF of the shared-memory code (kernels.cu:212-272) over every row
// of the array: a block takes nodes i0..i0+62 of 4 rows, one row per wave;
// the rows' split fluxes with a 3-node halo go through LDS (nodes outside
// 0..nx-1 are 0, :248-260), lane q computes the face below node i0+q, and
// Lu = -(F_{i+1/2} - F_{i-1/2})/d on the wave's 63 nodes.
template <int P>
__global__ __launch_bounds__(256) void k_adv_x_b5s(const double *__restrict__ u, double *__restrict__ Lu, int nx,
                                                   int ny, int nz, int px, int ntx, double d, double rd, double eps) {
    __shared__ double2 S[4][69];  // split fluxes of nodes i0-3 .. i0+65
    __shared__ double F[4][64];   // faces below nodes i0 .. i0+63
    const int lane = threadIdx.x & 63, r = threadIdx.x >> 6;
    const int i0 = static_cast<int>(blockIdx.x % ntx) * kB5sRow;
    const long nlines = static_cast<long>(ny) * nz;
    const long line = static_cast<long>(blockIdx.x / ntx) * 4 + r;
    const bool live = line < nlines;
    const long lc = live ? line : nlines - 1;  // rows past the end read a valid row and store nothing
    const unsigned base =
        static_cast<unsigned>(px) * static_cast<unsigned>(lc % ny) + static_cast<unsigned>(px) * ny * (lc / ny);
    bool ok = true;
    auto node = [&](int i) {
        const double v = (i >= 0 && i < nx) ? u[base + static_cast<unsigned>(i)] : 0.0;
        ok = ok && fabs(v) <= 0x1p40;
        return split_local(v);
    };
    S[r][lane] = node(i0 - 3 + lane);
    if (lane < 5) S[r][64 + lane] = node(i0 + 61 + lane);
    const bool fd = fast_division_ok<kWenoZB5S>(0.0, eps) && wave_all_ok(ok);
    __syncthreads();
    F[r][lane] = face_b5s<P>(fd, S[r][lane], S[r][lane + 1], S[r][lane + 2], S[r][lane + 3], S[r][lane + 4],
                             S[r][lane + 5], eps);
    __syncthreads();
    const int i = i0 + lane;
    if (live && lane < kB5sRow && i < nx) Lu[base + static_cast<unsigned>(i)] = -div_by(F[r][lane + 1] - F[r][lane], d, rd);
}

// Compute_dG (AXIS 1, kernels.cu:277-337) and Compute_dH (AXIS 2, :342-402)
// over every line of the array: one thread marches a chunk of kB5sChunk
// nodes of one line with the split fluxes of nodes l-3..l+2 in registers,
// lanes over x (coalesced); nodes outside the line are 0; each face is
// computed once and Lu -= (F_{l+1/2} - F_{l-1/2})/d.
template <int AXIS, int P>
__global__ __launch_bounds__(256) void k_adv_yz_b5s(const double *__restrict__ u, double *__restrict__ Lu, int nx,
                                                    int ny, int nz, int px, int nchunks, double d, double rd,
                                                    double eps) {
    const long gid = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
    const int n = AXIS == 1 ? ny : nz, nb = AXIS == 1 ? nz : ny;
    if (gid >= static_cast<long>(nx) * nb * nchunks) return;
    const int a = static_cast<int>(gid % nx);
    const long rest = gid / nx;
    const int b = static_cast<int>(rest % nb), c = static_cast<int>(rest / nb);
    const unsigned plane = static_cast<unsigned>(px) * ny;
    const unsigned base = static_cast<unsigned>(a) + (AXIS == 1 ? plane * b : static_cast<unsigned>(px) * b);
    const unsigned st = AXIS == 1 ? static_cast<unsigned>(px) : plane;
    const int l0 = c * kB5sChunk, l1 = min(l0 + kB5sChunk, n);
    const bool eps_ok = fast_division_ok<kWenoZB5S>(0.0, eps);
    bool ok = true;
    auto node = [&](int l) {
        const double v = (l >= 0 && l < n) ? u[base + st * static_cast<unsigned>(l)] : 0.0;
        ok = ok && fabs(v) <= 0x1p40;
        return split_local(v);
    };
    double2 q0 = node(l0 - 3), q1 = node(l0 - 2), q2 = node(l0 - 1), q3 = node(l0), q4 = node(l0 + 1),
            q5 = node(l0 + 2);
    double flo = face_b5s<P>(eps_ok && wave_all_ok(ok), q0, q1, q2, q3, q4, q5, eps);
    for (int l = l0; l < l1; l++) {
        q0 = q1;
        q1 = q2;
        q2 = q3;
        q3 = q4;
        q4 = q5;
        q5 = node(l + 3);
        const double fhi = face_b5s<P>(eps_ok && wave_all_ok(ok), q0, q1, q2, q3, q4, q5, eps);
        const unsigned o = base + st * static_cast<unsigned>(l);
        Lu[o] = Lu[o] - div_by(fhi - flo, d, rd);
        flo = fhi;
    }
}

// LaplaceO2_async (ORDER 2, Kernels.cu:152-201) and LaplaceO4_async
// (ORDER 4, :207-260) of the MPI heat code on nodes 3..n-4 in x and y and
// planes kstart..kstart+nk-1 of one rank's slab; one node per thread, lanes
// over x.  O2 sums each axis from its low side, O4 is one left-to-right sum
// of the twelve neighbours and -90 u scaled by diff_x/12.
template <int ORDER>
__global__ __launch_bounds__(256) void k_md3_lap(const double *__restrict__ u, double *__restrict__ Lu, int nx,
                                                 int ny, int nz, int px, int kstart, int ntx, int nty, double dxc,
                                                 double dyc, double dzc) {
    const long tile = blockIdx.x;
    const int i = 3 + static_cast<int>(tile % ntx) * 64 + (threadIdx.x & 63);
    const int j = 3 + static_cast<int>((tile / ntx) % nty) * 4 + (threadIdx.x >> 6);
    const int k = kstart + static_cast<int>(tile / (static_cast<long>(ntx) * nty));
    if (i >= nx - 3 || j >= ny - 3) return;
    const size_t p1 = static_cast<size_t>(px), p2 = 2 * p1, XY = p1 * ny, XY2 = 2 * XY;
    const size_t o = i + p1 * j + XY * k;
    const double center = u[o];
    if (ORDER == 2) {
        Lu[o] = dxc * (p2fma(-2.0, center, u[o - 1]) + u[o + 1]) +
                dyc * (p2fma(-2.0, center, u[o - p1]) + u[o + p1]) +
                dzc * (p2fma(-2.0, center, u[o - XY]) + u[o + XY]);
    } else {
        // -u_ww + 16u_w + 16u_e - u_ee - u_ss + 16u_s - 90u + 16u_n - u_nn - u_bb + 16u_b + 16u_t - u_tt,
        // left to right; the products by 16 are exact and fold into FMAs
        double sum = -u[o - 2];
        sum = p2fma(16.0, u[o - 1], sum);
        sum = p2fma(16.0, u[o + 1], sum);
        sum = sum - u[o + 2];
        sum = sum - u[o - p2];
        sum = p2fma(16.0, u[o - p1], sum);
        sum = sum - 90.0 * center;
        sum = p2fma(16.0, u[o + p1], sum);
        sum = sum - u[o + p2];
        sum = sum - u[o - XY2];
        sum = p2fma(16.0, u[o - XY], sum);
        sum = p2fma(16.0, u[o + XY], sum);
        sum = sum - u[o + XY2];
        Lu[o] = (dxc / 12) * sum;
    }
    (void)nz;
}

// Compute_RK of the MPI heat code (Kernels.cu:266-297): nodes 3..n-4 in x
// and y, every plane of the slab.
__global__ __launch_bounds__(256) void k_md3_rk(double *u, const double *uo, const double *__restrict__ Lu, int nx,
                                                int ny, int px, int ntx, int nty, int step, double dt) {
    const long tile = blockIdx.x;
    const int i = 3 + static_cast<int>(tile % ntx) * 64 + (threadIdx.x & 63);
    const int j = 3 + static_cast<int>((tile / ntx) % nty) * 4 + (threadIdx.x >> 6);
    const int k = static_cast<int>(tile / (static_cast<long>(ntx) * nty));
    if (i >= nx - 3 || j >= ny - 3) return;
    const size_t o = i + static_cast<size_t>(px) * j + static_cast<size_t>(px) * ny * k;
    u[o] = rk_stage(step, dt, u[o], uo[o], Lu[o]);
}

namespace {

// Arrays of px*ny*nz values (px >= nx) addressable with 32-bit offsets.
bool array_ok(unsigned nx, unsigned ny, unsigned nz, unsigned px) {
    return nx >= 1 && ny >= 1 && nz >= 1 && px >= nx && 1ull * px * ny * nz < (1ull << 31);
}

// Blocks of 64 x 4 nodes over nodes 3..n-4 in x and y, times nk planes.
struct Tiles3 {
    int ntx, nty;
    unsigned long long blocks;
};
Tiles3 tiles3(unsigned nx, unsigned ny, unsigned nk) {
    Tiles3 t;
    t.ntx = static_cast<int>((nx - 6 + 63) / 64);
    t.nty = static_cast<int>((ny - 6 + 3) / 4);
    t.blocks = 1ull * t.ntx * t.nty * nk;
    return t;
}

void adv_b5s(int axis, hipStream_t s, unsigned px, unsigned nx, unsigned ny, unsigned nz, double d, double *u,
             double *Lu) {
    if (!array_ok(nx, ny, nz, px) || !u || !Lu || !(d != 0.0)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    const AdvParams a = adv_params();
    const double rd = 1.0 / d;
    if (axis == 0) {
        const int ntx = static_cast<int>((nx + kB5sRow - 1) / kB5sRow);
        const unsigned long long blocks = 1ull * ntx * ((1ull * ny * nz + 3) / 4);
        if (a.p == 2)
            hipLaunchKernelGGL(k_adv_x_b5s<2>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, u, Lu, (int)nx,
                               (int)ny, (int)nz, (int)px, ntx, d, rd, a.eps);
        else
            hipLaunchKernelGGL(k_adv_x_b5s<1>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, u, Lu, (int)nx,
                               (int)ny, (int)nz, (int)px, ntx, d, rd, a.eps);
    } else {
        const unsigned n = axis == 1 ? ny : nz, nb = axis == 1 ? nz : ny;
        const int nchunks = static_cast<int>((n + kB5sChunk - 1) / kB5sChunk);
        const unsigned blocks = static_cast<unsigned>((1ull * nx * nb * nchunks + 255) / 256);
        auto kern = axis == 1 ? (a.p == 2 ? k_adv_yz_b5s<1, 2> : k_adv_yz_b5s<1, 1>)
                              : (a.p == 2 ? k_adv_yz_b5s<2, 2> : k_adv_yz_b5s<2, 1>);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, u, Lu, (int)nx, (int)ny, (int)nz, (int)px, nchunks, d,
                           rd, a.eps);
    }
    check_launch();
}

void md3_diff(int order, hipStream_t s, unsigned px, unsigned nx, unsigned ny, unsigned nz, unsigned kstart,
              unsigned kstop, double dxc, double dyc, double dzc, double *q, double *Lq) {
    const unsigned r = order == 2 ? 1u : 2u;  // planes read on each side
    if (!array_ok(nx, ny, nz, px) || nx < 7 || ny < 7 || kstart < r || kstop > nz - std::min(nz, r) ||
        kstart > kstop || !q || !Lq) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    if (kstart == kstop) return;
    const Tiles3 t = tiles3(nx, ny, kstop - kstart);
    if (order == 2)
        hipLaunchKernelGGL(k_md3_lap<2>, dim3(static_cast<unsigned>(t.blocks)), dim3(256), 0, s, q, Lq, (int)nx,
                           (int)ny, (int)nz, (int)px, (int)kstart, t.ntx, t.nty, dxc, dyc, dzc);
    else
        hipLaunchKernelGGL(k_md3_lap<4>, dim3(static_cast<unsigned>(t.blocks)), dim3(256), 0, s, q, Lq, (int)nx,
                           (int)ny, (int)nz, (int)px, (int)kstart, t.ntx, t.nty, dxc, dyc, dzc);
    check_launch();
}

}  // namespace
}  // namespace mgad

using namespace mgad;

extern "C" void Call_Adv_x_sharedmem(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                     unsigned int Ny, unsigned int Nz, double dx, double *u, double *Lu) {
    adv_b5s(0, aStream, pitch, Nx, Ny, Nz, dx, u, Lu);
}

extern "C" void Call_Adv_y_sharedmem(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                     unsigned int Ny, unsigned int Nz, double dy, double *v, double *Lv) {
    adv_b5s(1, aStream, pitch, Nx, Ny, Nz, dy, v, Lv);
}

extern "C" void Call_Adv_z_sharedmem(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                     unsigned int Ny, unsigned int Nz, double dz, double *w, double *Lw) {
    adv_b5s(2, aStream, pitch, Nx, Ny, Nz, dz, w, Lw);
}

extern "C" void Call_Lu3d_blocking(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int nx,
                                   unsigned int ny, unsigned int nz, double *u, double *Lu) {
    // nz >= 6 and pitch*ny*nz < 2^31 keep a plane below 2^29 values, the
    // z-march kernel's 32-bit plane offsets
    if (!array_ok(nx, ny, nz, pitch) || nx < 7 || ny < 7 || nz < 6 || !u || !Lu) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    lu13_b5_launch(aStream, nx, ny, nz, pitch, u, Lu);
}

extern "C" void Call_Diff_mpi(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                              unsigned int Ny, unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                              double diff_x, double diff_y, double diff_z, double *q, double *Lq) {
    md3_diff(4, aStream, pitch, Nx, Ny, _Nz, kstart, kstop, diff_x, diff_y, diff_z, q, Lq);
}

extern "C" void Call_Diff_O2_mpi(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                 unsigned int Ny, unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                                 double diff_x, double diff_y, double diff_z, double *q, double *Lq) {
    md3_diff(2, aStream, pitch, Nx, Ny, _Nz, kstart, kstop, diff_x, diff_y, diff_z, q, Lq);
}

extern "C" void Call_sspRK_mpi(dim3, dim3, hipStream_t aStream, unsigned int step, unsigned int pitch,
                               unsigned int Nx, unsigned int Ny, unsigned int _Nz, double dt, double *u, double *uo,
                               double *Lu) {
    if (!array_ok(Nx, Ny, _Nz, pitch) || Nx < 7 || Ny < 7 || step < 1 || step > 3 || !u || !uo || !Lu) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    const Tiles3 t = tiles3(Nx, Ny, _Nz);
    hipLaunchKernelGGL(k_md3_rk, dim3(static_cast<unsigned>(t.blocks)), dim3(256), 0, aStream, u, uo, Lu, (int)Nx,
                       (int)Ny, (int)pitch, t.ntx, t.nty, (int)step, dt);
    check_launch();
}
