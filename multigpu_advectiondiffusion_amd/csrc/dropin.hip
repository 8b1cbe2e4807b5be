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
#include <type_traits>

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

AdvParams adv_params() { return AdvParams{g_adv.weno, g_adv.p, g_adv.eps, g_adv.lf, g_adv.alpha}; }

// Which nodes a sweep updates: lines over the first transverse coordinate
// a = 3..n_a-4 and the second b = b0..b0+nb-1; along each line nodes l0..l1-1,
// using faces l0-1/2 .. l1-1/2.  zero_first: the face below node l0 is 0 (the
// reference's fu_old = 0, SingleGPU/Burgers3d_WENO5/kernels.cu:222) instead
// of being computed from nodes l0-3..l0+2.
struct SweepRange {
    int b0, nb, l0, l1, zero_first;
};

// Face flux with the 13/12 terms supplied; fd (the same for the whole wave)
// selects the shortened weight divisions, whose results equal IEEE division
// when every node read and a global LF alpha are at most 2^40 in magnitude
// and eps is in range (fast_division_ok).
template <int WENO, int P>
__device__ __forceinline__ double face_fd(bool fd, double v0, double v1, double v2, double v3, double v4, double w0,
                                          double w1, double w2, double w3, double w4, double Sp0, double Sp1,
                                          double Sp2, double Sm0, double Sm1, double Sm2, double eps) {
    if (fd)
        return weno_face_s<WENO, P, true>(v0, v1, v2, v3, v4, w0, w1, w2, w3, w4, Sp0, Sp1, Sp2, Sm0, Sm1, Sm2, eps);
    return weno_face_s<WENO, P, false>(v0, v1, v2, v3, v4, w0, w1, w2, w3, w4, Sp0, Sp1, Sp2, Sm0, Sm1, Sm2, eps);
}

// True for the whole wave when no active lane has seen a node above 2^40 in
// magnitude (or NaN): a scalar branch, no divergence.
__device__ __forceinline__ bool wave_all(bool ok) { return __builtin_amdgcn_ballot_w64(!ok) == 0; }

// AXIS 0: x lines (rows), 1: y lines, 2: z lines.  AXIS 0 writes Lu, 1 and 2 subtract.
// px is the row pitch in elements (px >= nx; px = nx for unpitched arrays).
// Along its line chunk a thread keeps the split fluxes of nodes i-3..i+3 and
// the 13/12 terms of the sub-stencil centres in register rings, so each
// node's terms are computed once, not once per face that reads them.
// Lu -/= (F_{i+1/2} - F_{i-1/2})/d is div_by with rd = RN(1/d).  The next
// node's u and Lu are loaded one face ahead.
template <int AXIS, int WENO, int P, bool LOCAL>
__global__ __launch_bounds__(256) void k_adv_sweep(const double *__restrict__ u,
                                                   double *__restrict__ Lu, int nx, int ny, int nz,
                                                   int px, double d, double rd, double eps, double alpha,
                                                   SweepRange r, int nchunks) {
    const int na = (AXIS == 0 ? ny : nx) - 6;  // first transverse coordinate count (3..n-4)
    const long gid = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
    const long nlines = static_cast<long>(na) * r.nb;
    if (gid >= nlines * nchunks) return;
    // lanes run fastest over the first transverse coordinate (x for y/z sweeps)
    const int a = 3 + static_cast<int>(gid % na);
    const long rest = gid / na;
    const int b = r.b0 + static_cast<int>(rest % r.nb);
    const int c = static_cast<int>(rest / r.nb);

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
    const int i0 = r.l0 + c * kChunk;
    const int i1 = min(i0 + kChunk, r.l1);
    if (i0 >= i1) return;
    const bool low_face = i0 > r.l0 || !r.zero_first;
    // a global LF alpha enters every split flux, so it is bounded like |u|
    const bool eps_ok = fast_division_ok<WENO>(LOCAL ? 0.0 : alpha, eps);

    bool ok = true;  // every node read so far is at most 2^40 in magnitude
    auto take = [&](double v, double &fp, double &fm) {
        ok = ok && fabs(v) <= 0x1p40;
        if (LOCAL)
            lf_split_local(v, fp, fm);
        else
            lf_split(v, alpha, fp, fm);
    };
    auto load = [&](int i, double &fp, double &fm) { take(u[base + st * static_cast<unsigned>(i)], fp, fm); };
    // rings: fp/fm hold the split fluxes of nodes i-3..i+3 for the face being
    // computed, node n in slot (n - i0 + 3) & 7 (node i0-3 is read only when
    // the chunk's low face is computed); SP/SM the 13/12 terms of f+ / f-
    // centred on node n in slot (n - i0 + 3) & 3.  The march is unrolled by
    // 8, so every slot is a constant and no register moves shift the rings.
    double fp[8], fm[8], SP[4], SM[4];
    fp[0] = fm[0] = 0.0;
    if (low_face) load(i0 - 3, fp[0], fm[0]);
#pragma unroll
    for (int q = 1; q < 6; q++) load(i0 - 3 + q, fp[q], fm[q]);
    SP[2] = sterm(fp[1], fp[2], fp[3]);
    SP[3] = sterm(fp[2], fp[3], fp[4]);
    SM[3] = sterm(fm[2], fm[3], fm[4]);
    SM[0] = sterm(fm[3], fm[4], fm[5]);
    double fu_old = 0.0;
    if (low_face)
        fu_old = face_fd<WENO, P>(eps_ok && wave_all(ok), fp[0], fp[1], fp[2], fp[3], fp[4], fm[1], fm[2], fm[3],
                                  fm[4], fm[5], sterm(fp[0], fp[1], fp[2]), SP[2], SP[3], sterm(fm[1], fm[2], fm[3]),
                                  SM[3], SM[0], eps);
    // u of node i+3 and Lu of node i are loaded one node ahead
    double u_next = u[base + st * static_cast<unsigned>(i0 + 3)];
    double lu_next = AXIS == 0 ? 0.0 : Lu[base + st * static_cast<unsigned>(i0)];
    int ib = i0;
    auto node = [&](auto S) {
        constexpr int s = decltype(S)::value;
        auto R8 = [](int q) { return (s + q) & 7; };
        auto R4 = [](int q) { return (s + q) & 3; };
        const int i = ib + s;
        const double v = u_next, lu = lu_next;
        if (i + 1 < i1) {
            u_next = u[base + st * static_cast<unsigned>(i + 4)];
            if (AXIS != 0) lu_next = Lu[base + st * static_cast<unsigned>(i + 1)];
        }
        take(v, fp[R8(6)], fm[R8(6)]);
        SP[R4(4)] = sterm(fp[R8(3)], fp[R8(4)], fp[R8(5)]);
        SM[R4(5)] = sterm(fm[R8(4)], fm[R8(5)], fm[R8(6)]);
        const double fu = face_fd<WENO, P>(eps_ok && wave_all(ok), fp[R8(1)], fp[R8(2)], fp[R8(3)], fp[R8(4)],
                                           fp[R8(5)], fm[R8(2)], fm[R8(3)], fm[R8(4)], fm[R8(5)], fm[R8(6)],
                                           SP[R4(2)], SP[R4(3)], SP[R4(4)], SM[R4(3)], SM[R4(4)], SM[R4(5)], eps);
        const unsigned o = base + st * static_cast<unsigned>(i);
        const double q = div_by(fu - fu_old, d, rd);
        Lu[o] = AXIS == 0 ? -q : lu - q;
        fu_old = fu;
    };
    for (; ib < i1; ib += 8) {
        node(std::integral_constant<int, 0>{});
        if (ib + 1 >= i1) break;
        node(std::integral_constant<int, 1>{});
        if (ib + 2 >= i1) break;
        node(std::integral_constant<int, 2>{});
        if (ib + 3 >= i1) break;
        node(std::integral_constant<int, 3>{});
        if (ib + 4 >= i1) break;
        node(std::integral_constant<int, 4>{});
        if (ib + 5 >= i1) break;
        node(std::integral_constant<int, 5>{});
        if (ib + 6 >= i1) break;
        node(std::integral_constant<int, 6>{});
        if (ib + 7 >= i1) break;
        node(std::integral_constant<int, 7>{});
    }
}

// The x sweep with lanes along x (coalesced): a block takes 63 consecutive
// nodes i0..i0+62 of 4 rows, one row per wave.  The split fluxes of the
// nodes, with a 3-node halo, go through LDS; lane q computes the face below
// node i0+q (0 at the line start, the reference's fu_old = 0), so every
// lane computes one face and Lu[i] = -(F_{i+1/2} - F_{i-1/2})/d on nodes
// 2..nx-4 of rows 3..ny-4 and planes b0..b0+nb-1 (Compute_dF, SingleGPU/
// Burgers3d_WENO5/kernels.cu:197-260; MultiGPU/Burgers3d_Baseline/
// Kernels.cu:225-293).
constexpr int kXRowNodes = 63;

template <int WENO, int P, bool LOCAL>
__global__ __launch_bounds__(256) void k_adv_x_rows(const double *__restrict__ u, double *__restrict__ Lu, int nx,
                                                    int ny, int px, double d, double rd, double eps, double alpha,
                                                    int b0, int nb) {
    __shared__ double2 S[4][69];  // (f+, f-) of nodes i0-3 .. i0+65
    __shared__ double F[4][64];   // faces i0-1/2 .. i0+62+1/2
    const int lane = threadIdx.x & 63, r = threadIdx.x >> 6;
    const int i0 = 2 + static_cast<int>(blockIdx.x) * kXRowNodes;
    const long line = static_cast<long>(blockIdx.y) * 4 + r;
    const long nlines = static_cast<long>(ny - 6) * nb;
    const bool live = line < nlines;
    const long lc = live ? line : nlines - 1;  // rows past the end load a valid row and store nothing
    const int j = 3 + static_cast<int>(lc % (ny - 6)), b = b0 + static_cast<int>(lc / (ny - 6));
    const unsigned base = static_cast<unsigned>(px) * j + static_cast<unsigned>(px) * ny * b;
    bool ok = true;  // the nodes this lane loads are at most 2^40 in magnitude
    auto split = [&](double v) {
        double fp, fm;
        ok = ok && fabs(v) <= 0x1p40;
        if (LOCAL)
            lf_split_local(v, fp, fm);
        else
            lf_split(v, alpha, fp, fm);
        return make_double2(fp, fm);
    };
    // node i0-3 is -1 on the first tile: its slot feeds only the face below
    // node 2, which is 0, so any valid node will do; nodes past the end clamp
    S[r][lane] = split(u[base + min(max(i0 - 3 + lane, 0), nx - 1)]);
    if (lane < 5) S[r][64 + lane] = split(u[base + min(i0 + 61 + lane, nx - 1)]);
    // one wave per row: the row's fast-division test is wave-uniform
    const bool fd = fast_division_ok<WENO>(LOCAL ? 0.0 : alpha, eps) && __builtin_amdgcn_ballot_w64(!ok) == 0;
    __syncthreads();
    {
        const double2 a0 = S[r][lane], a1 = S[r][lane + 1], a2 = S[r][lane + 2], a3 = S[r][lane + 3],
                      a4 = S[r][lane + 4], a5 = S[r][lane + 5];
        const double f =
            fd ? weno_face2<WENO, P, true>(a0.x, a1.x, a2.x, a3.x, a4.x, a1.y, a2.y, a3.y, a4.y, a5.y, eps)
               : weno_face2<WENO, P, false>(a0.x, a1.x, a2.x, a3.x, a4.x, a1.y, a2.y, a3.y, a4.y, a5.y, eps);
        F[r][lane] = (lane == 0 && i0 == 2) ? 0.0 : f;
    }
    __syncthreads();
    const int i = i0 + lane;
    if (live && lane < kXRowNodes && i < nx - 3) Lu[base + i] = -div_by(F[r][lane + 1] - F[r][lane], d, rd);
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
                    double *u, double *Lu, const AdvScheme &sc, SweepRange r) {
    if (AXIS == 0 && r.l0 == 2 && r.l1 == static_cast<int>(nx) - 3 && r.zero_first) {
        // x lines of the reference sweeps: lanes along x
        const dim3 grid((nx - 5 + kXRowNodes - 1) / kXRowNodes,
                        static_cast<unsigned>((static_cast<long>(ny - 6) * r.nb + 3) / 4));
        if (sc.lf == MGAD_LF_LOCAL)
            hipLaunchKernelGGL((k_adv_x_rows<WENO, P, true>), grid, dim3(256), 0, s, u, Lu, (int)nx, (int)ny, (int)px,
                               d, 1.0 / d, sc.eps, sc.alpha, r.b0, r.nb);
        else
            hipLaunchKernelGGL((k_adv_x_rows<WENO, P, false>), grid, dim3(256), 0, s, u, Lu, (int)nx, (int)ny,
                               (int)px, d, 1.0 / d, sc.eps, sc.alpha, r.b0, r.nb);
        return;
    }
    const long na = (AXIS == 0 ? ny : nx) - 6;
    const int nchunks = (r.l1 - r.l0 + kChunk - 1) / kChunk;
    const long total = na * r.nb * nchunks;
    const unsigned blocks = static_cast<unsigned>((total + 255) / 256);
    if (sc.lf == MGAD_LF_LOCAL)
        hipLaunchKernelGGL((k_adv_sweep<AXIS, WENO, P, true>), dim3(blocks), dim3(256), 0, s, u, Lu,
                           (int)nx, (int)ny, (int)nz, (int)px, d, 1.0 / d, sc.eps, sc.alpha, r, nchunks);
    else
        hipLaunchKernelGGL((k_adv_sweep<AXIS, WENO, P, false>), dim3(blocks), dim3(256), 0, s, u,
                           Lu, (int)nx, (int)ny, (int)nz, (int)px, d, 1.0 / d, sc.eps, sc.alpha, r, nchunks);
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
    // the reference's sweeps: lines over 3..n-4 both ways, nodes 2..n-4 along
    // the line with the face below node 2 taken as 0
    const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const SweepRange r{3, static_cast<int>(AXIS == 2 ? ny : nz) - 6, 2, n - 3, 1};
    if (g_adv.weno == MGAD_WENO_JS)
        launch_sweep_t<AXIS, kWenoJS, 2>(s, nx, ny, nz, px, d, u, Lu, g_adv, r);
    else if (g_adv.p == 2)
        launch_sweep_t<AXIS, kWenoZ, 2>(s, nx, ny, nz, px, d, u, Lu, g_adv, r);
    else
        launch_sweep_t<AXIS, kWenoZ, 1>(s, nx, ny, nz, px, d, u, Lu, g_adv, r);
    check_launch();
}

void diff13(hipStream_t s, unsigned Nx, unsigned Ny, unsigned Nz, unsigned px, double *q, double *Lq) {
    if (!grid_ok(Nx, Ny, Nz, px, 7) || !q || !Lq) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    const DiffCoef c = diff_coef();
    if (1ull * px * Ny < (1ull << 29)) {  // the z-march kernel's 32-bit plane offsets
        diff13_b5_launch(s, Nx, Ny, Nz, px, q, Lq, c.kx, c.ky, c.kz);
        return;
    }
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
    if (1ull * px * ny < (1ull << 29)) {  // planes addressable with 32-bit byte offsets
        heat13_lu_launch(s, nx, ny, nz, px, u, Lu);
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

// ---- MultiGPU/Burgers3d_Baseline: one rank's z slab of _Nz planes (3 ghost
// planes at each end), row pitch `pitch`; its scheme is fixed at compile time
// there: Reconstruct1d (WENO5-JS, Kernels.cu:112-220), EPS 1e-6
// (BurgersMPI.h:27), local Lax-Friedrichs split (Kernels.cu:254-262). ----

const AdvScheme kMb3Scheme{kWenoJS, 2, 1e-6, MGAD_LF_LOCAL, 0.0};

// Compute_dF/dG (Kernels.cu:225-371): lines over 3..n-4 and planes
// kstart..kstop-1, nodes 2..n-4 along the line from fu_old = 0.
// Compute_dH (:376-452): columns 3..n-4 in x and y, nodes kstart..kstop-1,
// the face below kstart computed from planes kstart-3..kstart+2.
template <int AXIS>
void mb3_sweep(hipStream_t s, unsigned pitch, unsigned nx, unsigned ny, unsigned nz, unsigned kstart,
               unsigned kstop, double d, double *u, double *Lu) {
    const bool range_ok = AXIS == 2 ? (kstart >= 3 && kstop + 3 <= nz) : kstop <= nz;
    if (!grid_ok(nx, ny, nz, pitch, 7) || !u || !Lu || kstart >= kstop || !range_ok) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    SweepRange r;
    if (AXIS == 2)
        r = SweepRange{3, static_cast<int>(ny) - 6, static_cast<int>(kstart), static_cast<int>(kstop), 0};
    else
        r = SweepRange{static_cast<int>(kstart), static_cast<int>(kstop - kstart), 2,
                       static_cast<int>(AXIS == 0 ? nx : ny) - 3, 1};
    launch_sweep_t<AXIS, kWenoJS, 2>(s, nx, ny, nz, pitch, d, u, Lu, kMb3Scheme, r);
    check_launch();
}

// Compute_RK (Kernels.cu:457-490): every node of the slab, ghosts included.
__global__ __launch_bounds__(256) void k_rk_all(double *q, const double *qo, const double *__restrict__ Lq, int nx,
                                                int ny, int nz, int px, int step, double dt) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const unsigned o = i + static_cast<unsigned>(px) * j + static_cast<unsigned>(px) * ny * k;
    q[o] = rk_stage(step, dt, q[o], qo[o], Lq[o]);
}

// copy_br_to_gc / copy_gc_to_br (Kernels.cu:40-107): the 3 boundary planes
// of side 0 (top: planes _Nz-6.._Nz-4 out, ghosts _Nz-3.._Nz-1 in) or side 1
// (bottom: planes 3..5 out, ghosts 0..2 in) to or from a 3-plane buffer of
// row pitch gc_pitch.
template <bool TO_BUFFER>
__global__ __launch_bounds__(256) void k_halo_copy(double *q, double *buf, int nx, int ny, int nz, int px,
                                                   int gpx, int side) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return;
    const int k0 = TO_BUFFER ? (side == 0 ? nz - 6 : 3) : (side == 0 ? nz - 3 : 0);
    const unsigned xy = static_cast<unsigned>(px) * ny, gxy = static_cast<unsigned>(gpx) * ny;
#pragma unroll
    for (int g = 0; g < 3; g++) {
        const unsigned o = i + static_cast<unsigned>(px) * j + xy * static_cast<unsigned>(k0 + g);
        const unsigned b = i + static_cast<unsigned>(gpx) * j + gxy * static_cast<unsigned>(g);
        if (TO_BUFFER)
            buf[b] = q[o];
        else
            q[o] = buf[b];
    }
}

template <bool TO_BUFFER>
void halo_copy(hipStream_t s, double *q, double *buf, unsigned nx, unsigned ny, unsigned nz, unsigned pitch,
               unsigned gc_pitch, unsigned side) {
    if (!grid_ok(nx, ny, nz, pitch, 7) || gc_pitch < nx || 3ull * gc_pitch * ny >= (1ull << 31) || !q || !buf ||
        side > 1) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    hipLaunchKernelGGL(k_halo_copy<TO_BUFFER>, dim3((nx + 63) / 64, (ny + 3) / 4), dim3(256), 0, s, q, buf,
                       (int)nx, (int)ny, (int)nz, (int)pitch, (int)gc_pitch, (int)side);
    check_launch();
}

}  // namespace

}  // namespace mgad

using namespace mgad;

extern "C" int mgad_set_adv_scheme(int weno, int p, double eps, int lf, double alpha) {
    // a global LF alpha bounds |f'(u)| = |u| from above: finite and >= 0
    if ((weno != MGAD_WENO_Z && weno != MGAD_WENO_JS) || (p != 1 && p != 2) ||
        (lf != MGAD_LF_LOCAL && lf != MGAD_LF_GLOBAL) || !(eps >= 0.0) ||
        !(alpha >= 0.0 && alpha <= 1.7976931348623157e308)) {
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

// ---- MultiGPU/Burgers3d_Baseline (BurgersMPI.h:112-132): one rank's slab ----

extern "C" void Compute_Adv_x_Async(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                    unsigned int Ny, unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                                    double dx, double *d_s_u, double *d_s_Lu) {
    mb3_sweep<0>(aStream, pitch, Nx, Ny, _Nz, kstart, kstop, dx, d_s_u, d_s_Lu);
}

extern "C" void Compute_Adv_y_Async(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                    unsigned int Ny, unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                                    double dy, double *d_s_u, double *d_s_Lu) {
    mb3_sweep<1>(aStream, pitch, Nx, Ny, _Nz, kstart, kstop, dy, d_s_u, d_s_Lu);
}

extern "C" void Compute_Adv_z_Async(dim3, dim3, hipStream_t aStream, unsigned int pitch, unsigned int Nx,
                                    unsigned int Ny, unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                                    double dz, double *d_s_u, double *d_s_Lu) {
    mb3_sweep<2>(aStream, pitch, Nx, Ny, _Nz, kstart, kstop, dz, d_s_u, d_s_Lu);
}

extern "C" void Compute_sspRK3Async(dim3, dim3, hipStream_t aStream, unsigned int step, unsigned int pitch,
                                    unsigned int Nx, unsigned int Ny, unsigned int _Nz, double dt, double *q,
                                    double *qo, double *Lq) {
    if (!grid_ok(Nx, Ny, _Nz, pitch, 1) || step < 1 || step > 3 || !q || !qo || !Lq || _Nz > 65535) {
        set_error(MGAD_ERR_INVALID_ARG);
        return;
    }
    hipLaunchKernelGGL(k_rk_all, dim3((Nx + 63) / 64, (Ny + 3) / 4, _Nz), dim3(256), 0, aStream, q, qo, Lq, (int)Nx,
                       (int)Ny, (int)_Nz, (int)pitch, (int)step, dt);
    check_launch();
}

extern "C" void CopyBoundaryRegionToGhostCellAsync(dim3, dim3, hipStream_t aStream, double *d_s_q,
                                                   double *d_send_buffer, unsigned int Nx, unsigned int Ny,
                                                   unsigned int _Nz, unsigned int pitch, unsigned int gc_pitch,
                                                   unsigned int side) {
    halo_copy<true>(aStream, d_s_q, d_send_buffer, Nx, Ny, _Nz, pitch, gc_pitch, side);
}

extern "C" void CopyGhostCellToBoundaryRegionAsync(dim3, dim3, hipStream_t aStream, double *d_s_q,
                                                   double *d_recv_buffer, unsigned int Nx, unsigned int Ny,
                                                   unsigned int _Nz, unsigned int pitch, unsigned int gc_pitch,
                                                   unsigned int side) {
    halo_copy<false>(aStream, d_s_q, d_recv_buffer, Nx, Ny, _Nz, pitch, gc_pitch, side);
}
