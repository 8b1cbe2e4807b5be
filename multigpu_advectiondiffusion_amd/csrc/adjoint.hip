// adjoint.hip -- discrete adjoint (reverse mode) of the periodic global-LF
// WENO5 Burgers SSP-RK3 run, with sqrt(N) checkpointing.
//
// J = 1/2 |u(T) - target|^2.  Per step, in reverse (stage inputs u^n, u1, u2):
//   stage 3: u3 = (u^n + 2 (u2 + dt L(u2; a3)))/3
//   stage 2: u2 = 0.75 u^n + 0.25 (u1 + dt L(u1; a2))
//   stage 1: u1 = u^n + dt L(u^n; a1)
// with a_s = max|stage input| (LF alpha; its derivative is routed to the
// first node attaining the maximum) and dt = cfl dx / a1, except a last step
// clipped to T, whose dt = T - sum of the earlier dt's.
//
// The VJP of L along one axis is a line march: the thread owning a chunk of
// nodes walks the faces that touch them, recomputes each face flux and its
// gradient with respect to its ten split-flux inputs, and scatters
// Fbar * gradient into a six-node register accumulator; a node is finished
// (and written, by its owner only) once the face two past it is done, so the
// result does not depend on thread timing.  Block partial sums of
// <lam, L> and <lam, dL/dalpha> feed the dt and alpha adjoints through a
// one-thread kernel per stage, all on the device.
//
// This file is built with FMA contraction and uses a Newton-refined hardware
// reciprocal: the gradient is checked against finite differences and the CPU
// reference adjoint to a relative tolerance, not bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "burgers_periodic.hpp"

namespace mgad {

__device__ __forceinline__ double frcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// Value and gradient of one reconstruction side (reverse mode, unit seed);
// same formulas as the forward weno_side_s.
template <int WENO, int P, bool PLUS>
__device__ __forceinline__ double side_grad(double v0, double v1, double v2, double v3, double v4, double eps,
                                            double *g) {
    const double s0 = (v0 - 2.0 * v1) + v2, t0 = (v0 - 4.0 * v1) + 3.0 * v2;
    const double s1 = (v1 - 2.0 * v2) + v3, t1 = v1 - v3;
    const double s2 = (v2 - 2.0 * v3) + v4, t2 = (3.0 * v2 - 4.0 * v3) + v4;
    const double B0 = kC1312 * s0 * s0 + kC14 * t0 * t0;
    const double B1 = kC1312 * s1 * s1 + kC14 * t1 * t1;
    const double B2 = kC1312 * s2 * s2 + kC14 * t2 * t2;
    double c0, c1, c2, d0, d1, d2;
    if (PLUS) {
        c0 = 2.0 * v0 - 7.0 * v1 + 11.0 * v2; c1 = -v1 + 5.0 * v2 + 2.0 * v3; c2 = 2.0 * v2 + 5.0 * v3 - v4;
        d0 = kD0N; d1 = kD1N; d2 = kD2N;
    } else {
        c0 = -v0 + 5.0 * v1 + 2.0 * v2; c1 = 2.0 * v1 + 5.0 * v2 - v3; c2 = 11.0 * v2 - 7.0 * v3 + 2.0 * v4;
        d0 = kD0P; d1 = kD1P; d2 = kD2P;
    }
    const double r0 = frcp(B0 + eps), r1 = frcp(B1 + eps), r2 = frcp(B2 + eps);
    double a0, a1, a2, q0 = 0, q1 = 0, q2 = 0, tau = 0;
    if (WENO == kWenoZ) {
        tau = fabs(B0 - B2);
        q0 = tau * r0; q1 = tau * r1; q2 = tau * r2;
        a0 = d0 * (1.0 + (P == 2 ? q0 * q0 : q0));
        a1 = d1 * (1.0 + (P == 2 ? q1 * q1 : q1));
        a2 = d2 * (1.0 + (P == 2 ? q2 * q2 : q2));
    } else {
        a0 = d0 * r0 * r0; a1 = d1 * r1 * r1; a2 = d2 * r2 * r2;
    }
    const double S = a0 + a1 + a2;
    const double rS = frcp(S);
    const double h = (a0 * c0 + a1 * c1 + a2 * c2) * (rS * (1.0 / 6.0));
    // reverse
    const double Nb = rS * (1.0 / 6.0), Sb = -h * rS;
    const double ab0 = Nb * c0 + Sb, ab1 = Nb * c1 + Sb, ab2 = Nb * c2 + Sb;
    const double cb0 = Nb * a0, cb1 = Nb * a1, cb2 = Nb * a2;
    double Bb0, Bb1, Bb2;
    if (WENO == kWenoZ) {
        const double qb0 = P == 2 ? ab0 * d0 * 2.0 * q0 : ab0 * d0;
        const double qb1 = P == 2 ? ab1 * d1 * 2.0 * q1 : ab1 * d1;
        const double qb2 = P == 2 ? ab2 * d2 * 2.0 * q2 : ab2 * d2;
        const double taub = qb0 * r0 + qb1 * r1 + qb2 * r2;
        const double sg = (B0 - B2) >= 0.0 ? 1.0 : -1.0;
        Bb0 = -qb0 * q0 * r0 + sg * taub;
        Bb1 = -qb1 * q1 * r1;
        Bb2 = -qb2 * q2 * r2 - sg * taub;
    } else {
        Bb0 = -2.0 * ab0 * a0 * r0;
        Bb1 = -2.0 * ab1 * a1 * r1;
        Bb2 = -2.0 * ab2 * a2 * r2;
    }
    const double s0b = 2.0 * kC1312 * s0 * Bb0, t0b = 2.0 * kC14 * t0 * Bb0;
    const double s1b = 2.0 * kC1312 * s1 * Bb1, t1b = 2.0 * kC14 * t1 * Bb1;
    const double s2b = 2.0 * kC1312 * s2 * Bb2, t2b = 2.0 * kC14 * t2 * Bb2;
    g[0] = s0b + t0b;
    g[1] = -2.0 * s0b - 4.0 * t0b + s1b + t1b;
    g[2] = s0b + 3.0 * t0b - 2.0 * s1b + s2b + 3.0 * t2b;
    g[3] = s1b - t1b - 2.0 * s2b - 4.0 * t2b;
    g[4] = s2b + t2b;
    if (PLUS) {
        g[0] += 2.0 * cb0; g[1] += -7.0 * cb0 - cb1; g[2] += 11.0 * cb0 + 5.0 * cb1 + 2.0 * cb2;
        g[3] += 2.0 * cb1 + 5.0 * cb2; g[4] += -cb2;
    } else {
        g[0] += -cb0; g[1] += 5.0 * cb0 + 2.0 * cb1; g[2] += 2.0 * cb0 + 5.0 * cb1 + 11.0 * cb2;
        g[3] += -cb1 - 7.0 * cb2; g[4] += 2.0 * cb2;
    }
    return h;
}

// Per-step record written during the forward recompute and read in reverse.
struct StepRec {
    double dt, a1, a2, a3, dt_cfl;
};

// Device-side scalars of the reverse sweep.
struct AdjScalars {
    double dtb;                // accumulated dJ/d(dt) of the current step
    double carry;              // dJ/d(dt_last) of a clipped last step
    double J;
    unsigned idx[3];           // first argmax|u| of the stage inputs (stage 1..3)
    unsigned pad;
};

// Records dt (and alpha of stage 1) after stage 1, the alphas of stages 2
// and 3 after stage 2; k_set_dtcfl then stores the unclipped CFL step
// cfl*dx/a1, so a clipped last step is recognised by dt != dt_cfl.
__global__ void k_record(const Ctrl *c, StepRec *rec, int after_stage) {
    if (after_stage == 1) {
        rec->dt = c->dt;
        rec->a1 = bits_double(c->amax[0]);
    } else {
        rec->a2 = bits_double(c->amax[1]);
        rec->a3 = bits_double(c->amax[2]);
    }
}

// Stage VJP along one axis.  FIRST (the x sweep) also writes the stage's
// pass-through terms: stage 3 ubi = 2/3 ubo, acc = 1/3 ubo; stage 2
// ubi = 1/4 ubo, acc += 3/4 ubo; stage 1 ubi (= acc, in place) = acc + ubo.
template <int AXIS, int WENO, int P, bool FIRST>
__global__ __launch_bounds__(256) void k_vjp_axis(const double *__restrict__ u, const double *__restrict__ ubo,
                                                  double *ubi, double *acc, const StepRec *rec, int stage,
                                                  int nx, int ny, int nz, double d, double eps, int chunk,
                                                  int nchunks, AdjScalars *sc, double *partials) {
    const int n = AXIS == 0 ? nx : (AXIS == 1 ? ny : nz);
    const int na = AXIS == 0 ? ny : nx;
    const int nb = AXIS == 2 ? ny : nz;
    const long gid = static_cast<long>(blockIdx.x) * 256 + threadIdx.x;
    const long nlines = static_cast<long>(na) * nb;
    double sL = 0.0, sLa = 0.0;
    const double dt = rec->dt;
    const double alpha = stage == 1 ? rec->a1 : (stage == 2 ? rec->a2 : rec->a3);
    const double coef = (stage == 1 ? 1.0 : (stage == 2 ? 0.25 : 2.0 / 3.0)) * dt;
    const double rd = 1.0 / d;
    if (gid < nlines * nchunks) {
        const int a = static_cast<int>(gid % na);
        const long rest = gid / na;
        const int b = static_cast<int>(rest % nb);
        const int c = static_cast<int>(rest / nb);
        const unsigned plane = static_cast<unsigned>(nx) * ny;
        unsigned base, st;
        if (AXIS == 0) { base = static_cast<unsigned>(nx) * a + plane * b; st = 1; }
        else if (AXIS == 1) { base = static_cast<unsigned>(a) + plane * b; st = nx; }
        else { base = static_cast<unsigned>(a) + static_cast<unsigned>(nx) * b; st = plane; }
        const int n0 = c * chunk, n1 = min(n0 + chunk, n);
        auto wrapn = [&](int i) { i %= n; return i < 0 ? i + n : i; };
        // queues for nodes i-2 .. i+3 of face i (index 0..5)
        double fp[6], fm[6], uu[6], Ap[6], Am[6];
        for (int q = 0; q < 5; q++) {
            const double v = u[base + st * wrapn(n0 - 5 + q)];
            uu[q + 1] = v;
            const double f = (0.5 * v) * v, au = alpha * v;
            fp[q + 1] = 0.5 * (f + au);
            fm[q + 1] = 0.5 * (f - au);
        }
        for (int q = 0; q < 6; q++) Ap[q] = Am[q] = 0.0;
        double lam_i = coef * ubo[base + st * wrapn(n0 - 3)];
        for (int i = n0 - 3; i <= n1 + 1; i++) {
            // shift in node i+3
#pragma unroll
            for (int q = 0; q < 5; q++) { fp[q] = fp[q + 1]; fm[q] = fm[q + 1]; uu[q] = uu[q + 1]; }
            {
                const double v = u[base + st * wrapn(i + 3)];
                uu[5] = v;
                const double f = (0.5 * v) * v, au = alpha * v;
                fp[5] = 0.5 * (f + au);
                fm[5] = 0.5 * (f - au);
            }
            const double lam_n = coef * ubo[base + st * wrapn(i + 1)];
            const double Fb = (lam_n - lam_i) * rd;
            lam_i = lam_n;
            double gp[5], gm[5];
            const double F = side_grad<WENO, P, true>(fp[0], fp[1], fp[2], fp[3], fp[4], eps, gp) +
                             side_grad<WENO, P, false>(fm[1], fm[2], fm[3], fm[4], fm[5], eps, gm);
            if (i >= n0 && i < n1) sL += F * Fb;
#pragma unroll
            for (int q = 0; q < 5; q++) {
                Ap[q] += Fb * gp[q];
                Am[q + 1] += Fb * gm[q];
            }
            // node i-2 is complete
            const int j = i - 2;
            if (j >= n0 && j < n1) {
                const double uj = uu[0];
                const double contrib = Ap[0] * (0.5 * (uj + alpha)) + Am[0] * (0.5 * (uj - alpha));
                sLa += 0.5 * uj * (Ap[0] - Am[0]);
                const unsigned o = base + st * static_cast<unsigned>(j);
                if (FIRST) {
                    const double ub = ubo[o];
                    if (stage == 3) {
                        ubi[o] = (2.0 / 3.0) * ub + contrib;
                        acc[o] = ub * (1.0 / 3.0);
                    } else if (stage == 2) {
                        ubi[o] = 0.25 * ub + contrib;
                        acc[o] = acc[o] + 0.75 * ub;
                    } else {
                        ubi[o] = acc[o] + ub + contrib;
                    }
                    if (fabs(uj) == alpha) atomicMin(&sc->idx[stage - 1], o);
                } else {
                    ubi[o] += contrib;
                }
            }
#pragma unroll
            for (int q = 0; q < 5; q++) { Ap[q] = Ap[q + 1]; Am[q] = Am[q + 1]; }
            Ap[5] = Am[5] = 0.0;
        }
    }
    // block partial sums (fixed order)
    __shared__ double red[2][256];
    red[0][threadIdx.x] = sL;
    red[1][threadIdx.x] = sLa;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = red[0][0];
        partials[2 * blockIdx.x + 1] = red[1][0];
    }
}

// One thread: sums the stage's partials (x, y, z blocks in order), routes the
// alpha adjoint to its argmax node and, at stage 1, closes the dt chain.
__global__ void k_adj_scalars(const double *partials, int nparts, const StepRec *rec, int stage, double *ubi,
                              const double *uin, AdjScalars *sc, int clipped_last) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double L = 0.0, La = 0.0;
    for (int b = 0; b < nparts; b++) {
        L += partials[2 * b];
        La += partials[2 * b + 1];
    }
    sc->dtb += L / rec->dt;
    double ab = La;
    if (stage == 1) {
        const bool clipped = rec->dt != rec->dt_cfl;
        if (clipped && clipped_last) {  // only the run's last step can be clipped
            sc->carry = sc->dtb;
        } else {
            ab += -(sc->dtb - sc->carry) * rec->dt / rec->a1;
        }
        sc->dtb = 0.0;
    }
    const unsigned idx = sc->idx[stage - 1];
    if (idx != 0xffffffffu) ubi[idx] += ab * (uin[idx] >= 0.0 ? 1.0 : -1.0);
    sc->idx[stage - 1] = 0xffffffffu;
}

__global__ void k_adj_reset(AdjScalars *sc) {
    sc->dtb = 0.0;
    sc->carry = 0.0;
    sc->J = 0.0;
    sc->idx[0] = sc->idx[1] = sc->idx[2] = 0xffffffffu;
}

// ubar = u - target; block partial sums of 1/2 (u - target)^2.
__global__ __launch_bounds__(256) void k_misfit(const double *u, const double *tgt, double *ub, size_t n,
                                                double *partials) {
    double s = 0.0;
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
        const double e = u[i] - tgt[i];
        ub[i] = e;
        s += 0.5 * e * e;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

__global__ void k_sum_J(const double *partials, int n, AdjScalars *sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += partials[i];
        sc->J = s;
    }
}

__global__ void k_set_dtcfl(StepRec *rec, double cfl_dx) { rec->dt_cfl = cfl_dx / rec->a1; }

namespace {

constexpr int kAdjChunk = 32;
constexpr int kMisfitBlocks = 1024;

int vjp_blocks(int n, int na, int nb) {
    const int nch = (n + kAdjChunk - 1) / kAdjChunk;
    const long threads = static_cast<long>(na) * nb * nch;
    return static_cast<int>((threads + 255) / 256);
}

template <int WENO, int P>
int launch_stage_vjp(hipStream_t s, int stage, const double *uin, const double *ubo, double *ubi, double *acc,
                     const StepRec *rec, int nx, int ny, int nz, double dx, double dy, double dz, double eps,
                     AdjScalars *sc, double *partials, bool clipped_last) {
    const int bx = vjp_blocks(nx, ny, nz), by = vjp_blocks(ny, nx, nz), bz = vjp_blocks(nz, nx, ny);
    double *px = partials, *py = partials + 2 * bx, *pz = py + 2 * by;
    hipLaunchKernelGGL((k_vjp_axis<0, WENO, P, true>), dim3(bx), dim3(256), 0, s, uin, ubo, ubi, acc, rec, stage,
                       nx, ny, nz, dx, eps, kAdjChunk, (nx + kAdjChunk - 1) / kAdjChunk, sc, px);
    hipLaunchKernelGGL((k_vjp_axis<1, WENO, P, false>), dim3(by), dim3(256), 0, s, uin, ubo, ubi, acc, rec, stage,
                       nx, ny, nz, dy, eps, kAdjChunk, (ny + kAdjChunk - 1) / kAdjChunk, sc, py);
    hipLaunchKernelGGL((k_vjp_axis<2, WENO, P, false>), dim3(bz), dim3(256), 0, s, uin, ubo, ubi, acc, rec, stage,
                       nx, ny, nz, dz, eps, kAdjChunk, (nz + kAdjChunk - 1) / kAdjChunk, sc, pz);
    hipLaunchKernelGGL(k_adj_scalars, dim3(1), dim3(64), 0, s, partials, bx + by + bz, rec, stage, ubi, uin, sc,
                       clipped_last ? 1 : 0);
    return check_launch();
}

int stage_vjp(hipStream_t s, const mgad_scheme &sc, int stage, const double *uin, const double *ubo, double *ubi,
              double *acc, const StepRec *rec, int nx, int ny, int nz, double dx, double dy, double dz,
              AdjScalars *scal, double *partials, bool clipped_last) {
    if (sc.weno == MGAD_WENO_JS)
        return launch_stage_vjp<kWenoJS, 2>(s, stage, uin, ubo, ubi, acc, rec, nx, ny, nz, dx, dy, dz, sc.eps, scal,
                                            partials, clipped_last);
    if (sc.p == 2)
        return launch_stage_vjp<kWenoZ, 2>(s, stage, uin, ubo, ubi, acc, rec, nx, ny, nz, dx, dy, dz, sc.eps, scal,
                                           partials, clipped_last);
    return launch_stage_vjp<kWenoZ, 1>(s, stage, uin, ubo, ubi, acc, rec, nx, ny, nz, dx, dy, dz, sc.eps, scal,
                                       partials, clipped_last);
}

// Workspace layout (bytes, 256-aligned pieces).
struct AdjLayout {
    size_t n, K, nchk;
    size_t off_chk, off_seg, off_ub, off_rec, off_ctrl, off_sc, off_part, off_tchk, total;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

AdjLayout layout(unsigned nx, unsigned ny, unsigned nz, int max_steps, int every) {
    AdjLayout L{};
    L.n = static_cast<size_t>(nx) * ny * nz;
    L.K = static_cast<size_t>(every);
    L.nchk = (static_cast<size_t>(max_steps) + every - 1) / every + 1;
    const size_t f = align256(L.n * sizeof(double));
    size_t o = 0;
    L.off_chk = o; o += L.nchk * f;            // checkpoints u^{cK}
    L.off_seg = o; o += (3 * L.K + 1) * f;     // segment states S0[0..K], S1[0..K-1], S2[0..K-1]
    L.off_ub = o; o += 4 * f;                  // adjoint fields
    L.off_rec = o; o += align256(L.K * sizeof(StepRec));
    L.off_ctrl = o; o += align256(MGAD_CTRL_BYTES);
    L.off_sc = o; o += align256(sizeof(AdjScalars));
    const int maxb = vjp_blocks(std::max({nx, ny, nz}), std::max({nx, ny, nz}), std::max({nx, ny, nz}));
    L.off_part = o; o += align256(sizeof(double) * std::max<size_t>(6 * maxb, kMisfitBlocks));
    L.off_tchk = o; o += align256(sizeof(double) * L.nchk);
    L.total = o;
    return L;
}

}  // namespace
}  // namespace mgad

using namespace mgad;

extern "C" size_t mgad_adjoint_workspace_bytes(unsigned int nx, unsigned int ny, unsigned int nz, int max_steps,
                                               int checkpoint_every) {
    if (!valid_grid(nx, ny, nz) || max_steps <= 0 || checkpoint_every <= 0) return 0;
    return layout(nx, ny, nz, max_steps, checkpoint_every).total;
}

extern "C" int mgad_burgers_adjoint(hipStream_t stream, const mgad_scheme *scheme, unsigned int nx, unsigned int ny,
                                    unsigned int nz, double dx, double dy, double dz, double cfl, double t0,
                                    double tend, const double *u0, const double *target, double *grad,
                                    double *J_out, int max_steps, int checkpoint_every, void *workspace,
                                    size_t workspace_bytes, int *nsteps_out) {
    if (!scheme || !u0 || !target || !grad || !workspace || !valid_grid(nx, ny, nz) || max_steps <= 0 ||
        checkpoint_every <= 0 || nx % 16 || ny % 16 || !(tend > t0) || !(cfl > 0)) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    const AdjLayout L = layout(nx, ny, nz, max_steps, checkpoint_every);
    if (workspace_bytes < L.total) {
        set_error(MGAD_ERR_INVALID_ARG);
        return MGAD_ERR_INVALID_ARG;
    }
    char *ws = static_cast<char *>(workspace);
    const size_t f = align256(L.n * sizeof(double));
    auto chk = [&](size_t c) { return reinterpret_cast<double *>(ws + L.off_chk + c * f); };
    auto S0 = [&](size_t m) { return reinterpret_cast<double *>(ws + L.off_seg + m * f); };
    auto S1 = [&](size_t m) { return reinterpret_cast<double *>(ws + L.off_seg + (L.K + 1 + m) * f); };
    auto S2 = [&](size_t m) { return reinterpret_cast<double *>(ws + L.off_seg + (2 * L.K + 1 + m) * f); };
    auto UB = [&](int q) { return reinterpret_cast<double *>(ws + L.off_ub + q * f); };
    StepRec *rec = reinterpret_cast<StepRec *>(ws + L.off_rec);
    void *ctrl = ws + L.off_ctrl;
    AdjScalars *scal = reinterpret_cast<AdjScalars *>(ws + L.off_sc);
    double *partials = reinterpret_cast<double *>(ws + L.off_part);
    const size_t bytes = L.n * sizeof(double);
    const int K = checkpoint_every;

    // ---- forward pass, checkpoint every K steps ----
    int st = check_hip(hipMemcpyAsync(chk(0), u0, bytes, hipMemcpyDeviceToDevice, stream));
    if (!st) st = burgers_init_t<double>(stream, nx, ny, nz, chk(0), ctrl, t0, tend, cfl);
    if (st) return st;
    std::vector<double> tchk{t0};
    std::vector<int> seg_len;
    int total = 0;
    double *u = S0(0), *w1 = S1(0), *w2 = S2(0);
    st = check_hip(hipMemcpyAsync(u, u0, bytes, hipMemcpyDeviceToDevice, stream));
    if (st) return st;
    for (;;) {
        for (int s = 0; s < K && !st; s++) {
            st = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 1, u, u, w1, ctrl, 0, 0);
            if (!st) st = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 2, w1, u, w2, ctrl, 0, 0);
            if (!st) st = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 3, w2, u, u, ctrl, 0, 0);
        }
        if (st) return st;
        mgad_burgers_state cs;
        if ((st = mgad_burgers_read_ctrl(stream, ctrl, &cs))) return st;
        const int done = static_cast<int>(cs.steps) - total;
        seg_len.push_back(done);
        total = static_cast<int>(cs.steps);
        if (total > max_steps) {
            set_error(MGAD_ERR_INVALID_ARG);
            return MGAD_ERR_INVALID_ARG;
        }
        if (!(cs.t < tend)) break;
        if (tchk.size() >= L.nchk) {
            set_error(MGAD_ERR_INVALID_ARG);
            return MGAD_ERR_INVALID_ARG;
        }
        if ((st = check_hip(hipMemcpyAsync(chk(tchk.size()), u, bytes, hipMemcpyDeviceToDevice, stream)))) return st;
        tchk.push_back(cs.t);
    }
    // J and ubar^N = u^N - target
    double *ubN = UB(0), *ub2 = UB(1), *ub1 = UB(2), *acc = UB(3);
    hipLaunchKernelGGL(k_adj_reset, dim3(1), dim3(1), 0, stream, scal);
    hipLaunchKernelGGL(k_misfit, dim3(kMisfitBlocks), dim3(256), 0, stream, u, target, ubN, L.n, partials);
    hipLaunchKernelGGL(k_sum_J, dim3(1), dim3(1), 0, stream, partials, kMisfitBlocks, scal);
    if ((st = check_launch())) return st;

    // ---- reverse sweep, segment by segment ----
    double *ub_next = ubN;  // adjoint of the state after the step being reversed
    for (int c = static_cast<int>(seg_len.size()) - 1; c >= 0; c--) {
        const int len = seg_len[c];
        // recompute the segment from its checkpoint, keeping every stage input
        if ((st = check_hip(hipMemcpyAsync(S0(0), chk(c), bytes, hipMemcpyDeviceToDevice, stream)))) return st;
        if ((st = burgers_init_t<double>(stream, nx, ny, nz, S0(0), ctrl, tchk[c], tend, cfl))) return st;
        for (int m = 0; m < len && !st; m++) {
            st = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 1, S0(m), S0(m), S1(m), ctrl, 0, 0);
            if (!st) {
                hipLaunchKernelGGL(k_record, dim3(1), dim3(1), 0, stream, static_cast<const Ctrl *>(ctrl), rec + m, 1);
                hipLaunchKernelGGL(k_set_dtcfl, dim3(1), dim3(1), 0, stream, rec + m, cfl * dx);
                st = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 2, S1(m), S0(m), S2(m), ctrl, 0, 0);
            }
            if (!st) {
                hipLaunchKernelGGL(k_record, dim3(1), dim3(1), 0, stream, static_cast<const Ctrl *>(ctrl), rec + m, 2);
                st = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 3, S2(m), S0(m), S0(m + 1), ctrl, 0, 0);
            }
        }
        if (st) return st;
        for (int m = len - 1; m >= 0 && !st; m--) {
            const bool last = (c == static_cast<int>(seg_len.size()) - 1) && (m == len - 1);
            // stage 3: adjoint of u2 into ub2, acc = 1/3 ubar
            st = stage_vjp(stream, *scheme, 3, S2(m), ub_next, ub2, acc, rec + m, nx, ny, nz, dx, dy, dz, scal,
                           partials, last);
            if (!st)
                st = stage_vjp(stream, *scheme, 2, S1(m), ub2, ub1, acc, rec + m, nx, ny, nz, dx, dy, dz, scal,
                               partials, last);
            // stage 1 writes ubar^n in place of acc
            if (!st)
                st = stage_vjp(stream, *scheme, 1, S0(m), ub1, acc, acc, rec + m, nx, ny, nz, dx, dy, dz, scal,
                               partials, last);
            // rotate: acc holds ubar^n, which is the next ub_next
            std::swap(ub_next, acc);
        }
        if (st) return st;
    }
    st = check_hip(hipMemcpyAsync(grad, ub_next, bytes, hipMemcpyDeviceToDevice, stream));
    double Jh = 0.0;
    if (!st) st = check_hip(hipMemcpyAsync(&Jh, &scal->J, sizeof(double), hipMemcpyDeviceToHost, stream));
    if (!st) st = check_hip(hipStreamSynchronize(stream));
    if (st) return st;
    if (J_out) *J_out = Jh;
    if (nsteps_out) *nsteps_out = total;
    return MGAD_OK;
}
