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

// Records dt, alpha of stage 1 and the unclipped CFL step cfl*dx/a1 after
// stage 1 (a clipped last step is recognised by dt != dt_cfl), the alphas of
// stages 2 and 3 after stage 2.  A step past tend records dt = 0 (unused).
__global__ void k_record(const Ctrl *c, StepRec *rec, int after_stage, double cfl_dx) {
    if (after_stage == 1) {
        rec->dt = c->dt;
        rec->a1 = bits_double(c->amax[0]);
        rec->dt_cfl = cfl_dx / rec->a1;
    } else {
        rec->a2 = bits_double(c->amax[1]);
        rec->a3 = bits_double(c->amax[2]);
    }
}

// Contributions of one face to the six nodes of its stencil: the face
// i+1/2 reads f+ at nodes i-2..i+2 (v) and f- at i-1..i+3 (w); node t of
// i-2..i+3 receives Fb * (dF/df+_t * (u_t+alpha)/2 + dF/df-_t * (u_t-alpha)/2).
// Accumulates F*Fb (the face's share of <lam, L>) and Fb * sum_t u_t/2 *
// (dF/df+_t - dF/df-_t) (its share of <lam, dL/dalpha>).
template <int WENO, int P>
__device__ __forceinline__ void face_vjp(const double (&fp)[6], const double (&fm)[6], const double (&uu)[6],
                                         double Fb, double alpha, double eps, double (&c)[6], double &sL,
                                         double &sLa) {
    double gp[5], gm[5];
    const double F = side_grad<WENO, P, true>(fp[0], fp[1], fp[2], fp[3], fp[4], eps, gp) +
                     side_grad<WENO, P, false>(fm[1], fm[2], fm[3], fm[4], fm[5], eps, gm);
    sL += F * Fb;
    double la = 0.0;
#pragma unroll
    for (int t = 0; t < 6; t++) {
        const double Gp = t < 5 ? Fb * gp[t] : 0.0;
        const double Gm = t > 0 ? Fb * gm[t - 1] : 0.0;
        c[t] = Gp * (0.5 * (uu[t] + alpha)) + Gm * (0.5 * (uu[t] - alpha));
        la += 0.5 * uu[t] * (Gp - Gm);
    }
    sLa += la;
}

// Grid of the VJP launch: extents, spacings, planes per z chunk (the last
// chunk also takes the remainder, so every chunk has at least 5 planes).
struct VjpGrid {
    int nx, ny, nz, kchunk, nchunks;
    double dx, dy, dz;
};

// LDS of one VJP tile: split fluxes and u of the current plane with a halo of
// two nodes below and three above the tile in x and y, lambda with one node
// above, and the six node contributions of every x and y face of the tile.
template <int TX, int TY>
struct VjpSmem {
    double2 P[TY + 5][TX + 5];
    double U[TY + 5][TX + 5];
    double Lm[TY + 1][TX + 1];
    double Gx[6][TY][TX];  // Gx[t][ty][tx]: face x0+tx+1/2 to node x0+tx-2+t
    double Gy[6][TY][TX];  // Gy[t][ty][tx]: face y0+ty+1/2 to node y0+ty-2+t
    double red[2][TX * TY / 64];
};

// Stage VJP of the periodic RHS, all three axes in one pass: ubi = pass-through
// + sum over the faces of L of Fb * dF/du, with Fb = (lam_{i+1} - lam_i)/d and
// lam = coef * ubo.  A block owns a TX x TY column tile and a z chunk
// [z0, z1); it computes each face it owns once -- x faces x0..x0+TX-1 (+1/2),
// y faces likewise, z faces z0..z1-1 -- one of each per thread and plane.
// Contributions to nodes of the tile are gathered in a fixed order; those to
// nodes of a neighbouring tile or chunk (two below, three above in each axis)
// go to the exchange fields exx/exy/exz, which k_vjp_inflow adds afterwards.
// The z march finishes node k-2 once its last own z face, k+1/2, is done.
// Pass-through terms: stage 3 ubi = 2/3 ubo, acc = 1/3 ubo; stage 2
// ubi = 1/4 ubo, acc += 3/4 ubo; stage 1 ubi (= acc, in place) = acc + ubo.
template <int WENO, int P, int TX, int TY>
__global__ __launch_bounds__(TX * TY, 2) void k_vjp_tile(
    const double *__restrict__ u, const double *__restrict__ ubo, double *ubi, double *acc, const StepRec *rec,
    int stage, VjpGrid g, double eps, double *__restrict__ exx, double *__restrict__ exy, double *__restrict__ exz,
    AdjScalars *sc, double *partials) {
    constexpr int NT = TX * TY;
    constexpr int NW = NT / 64;
    static_assert(NT % 64 == 0 && TX >= 5 && TY >= 5, "tile shape");
    static_assert(5 * (TX + TY) + TX + TY <= NT, "one halo role per thread");
    __shared__ VjpSmem<TX, TY> sm;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int ntx = nx / TX, nty = ny / TY;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
    const int x0 = bx * TX, y0 = by * TY, z0 = bz * g.kchunk;
    const int z1 = bz == g.nchunks - 1 ? nz : z0 + g.kchunk;
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const double dt = rec->dt;
    const double alpha = stage == 1 ? rec->a1 : (stage == 2 ? rec->a2 : rec->a3);
    const double coef = (stage == 1 ? 1.0 : (stage == 2 ? 0.25 : 2.0 / 3.0)) * dt;
    const double rdx = 1.0 / g.dx, rdy = 1.0 / g.dy, rdz = 1.0 / g.dz;
    const unsigned plane = static_cast<unsigned>(nx) * static_cast<unsigned>(ny);
    const unsigned col = static_cast<unsigned>(x0 + tx) + static_cast<unsigned>(nx) * (y0 + ty);

    // u halo role: 5*TY x-halo nodes (x0-2, x0-1, x0+TX..x0+TX+2), then 5*TX y-halo nodes
    int hr = 0, hc = 0;
    unsigned hcol = 0;
    const bool has_halo = tid < 5 * (TX + TY);
    if (tid < 5 * TY) {
        const int r = tid / 5, c = tid % 5;
        hr = r + 2;
        hc = c < 2 ? c : TX + c;
        hcol = static_cast<unsigned>(wrapi(x0 + hc - 2, nx)) + static_cast<unsigned>(nx) * (y0 + r);
    } else if (has_halo) {
        const int h = tid - 5 * TY, c = h % TX, q = h / TX;
        hr = q < 2 ? q : TY + q;
        hc = c + 2;
        hcol = static_cast<unsigned>(x0 + c) + static_cast<unsigned>(nx) * wrapi(y0 + hr - 2, ny);
    }
    // lambda halo role: column x0+TX (TY nodes), row y0+TY (TX nodes), last threads
    const int lt = tid - (NT - TX - TY);
    const bool has_lh = lt >= 0;
    int lr = 0, lc = 0;
    unsigned lcol = 0;
    if (has_lh) {
        if (lt < TY) {
            lr = lt;
            lc = TX;
            lcol = static_cast<unsigned>(wrapi(x0 + TX, nx)) + static_cast<unsigned>(nx) * (y0 + lt);
        } else {
            lr = TY;
            lc = lt - TY;
            lcol = static_cast<unsigned>(x0 + lc) + static_cast<unsigned>(nx) * wrapi(y0 + TY, ny);
        }
    }
    auto zoff = [&](int k) { return plane * static_cast<unsigned>(wrapi(k, nz)); };

    // z queues: planes k-2..k+3 (index 0..5) of the own column; az: z-face
    // contributions to nodes k-2..k+3; dq: pass-through + x/y part of nodes k-2, k-1
    double fpq[6], fmq[6], uq[6], az[6], dq[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 5; q++) {
        uq[q] = u[col + zoff(z0 - 2 + q)];
        lf_split(uq[q], alpha, fpq[q], fmq[q]);
    }
#pragma unroll
    for (int q = 0; q < 6; q++) az[q] = 0.0;
    double ub0 = ubo[col + zoff(z0)];
    double u_next = u[col + zoff(z0 + 3)];
    double h_next = has_halo ? u[hcol + zoff(z0)] : 0.0;
    double ub_next = ubo[col + zoff(z0 + 1)];
    double lh_next = has_lh ? ubo[lcol + zoff(z0)] : 0.0;
    double ac_next = stage == 3 ? 0.0 : acc[col + zoff(z0)];
    double sL = 0.0, sLa = 0.0;

    // sum over own faces of the contributions to target node tt (tile-relative,
    // -2..TX+2) of row r, from Gx (AXIS 0) or Gy (AXIS 1), in a fixed order
    auto gather = [&](const double (&G)[6][TY][TX], int r, int tt, bool xdir) {
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < 6; t++) {
            const int f = tt + 2 - t;  // face index whose node t is tt
            const int lim = xdir ? TX : TY;
            if (f >= 0 && f < lim) s += xdir ? G[t][r][f] : G[t][f][r];
        }
        return s;
    };

    for (int k = z0; k < z1; k++) {
        const double un = u_next, hv = h_next, ub1 = ub_next, lh = lh_next, ac = ac_next;
        if (k + 1 < z1) {
            u_next = u[col + zoff(k + 4)];
            if (has_halo) h_next = u[hcol + zoff(k + 1)];
            ub_next = ubo[col + zoff(k + 2)];
            if (has_lh) lh_next = ubo[lcol + zoff(k + 1)];
            if (stage != 3) ac_next = acc[col + zoff(k + 1)];
        }
        uq[5] = un;
        lf_split(un, alpha, fpq[5], fmq[5]);
        sm.P[ty + 2][tx + 2] = make_double2(fpq[2], fmq[2]);
        sm.U[ty + 2][tx + 2] = uq[2];
        sm.Lm[ty][tx] = coef * ub0;
        if (has_halo) {
            double hp, hm;
            lf_split(hv, alpha, hp, hm);
            sm.P[hr][hc] = make_double2(hp, hm);
            sm.U[hr][hc] = hv;
        }
        if (has_lh) sm.Lm[lr][lc] = coef * lh;
        __syncthreads();

        double c[6];
        {  // x face x0+tx+1/2: nodes tile-x tx-2..tx+3 = LDS cols tx..tx+5
            double fp[6], fm[6], uu[6];
#pragma unroll
            for (int t = 0; t < 6; t++) {
                const double2 pv = sm.P[ty + 2][tx + t];
                fp[t] = pv.x;
                fm[t] = pv.y;
                uu[t] = sm.U[ty + 2][tx + t];
            }
            const double Fb = (sm.Lm[ty][tx + 1] - sm.Lm[ty][tx]) * rdx;
            face_vjp<WENO, P>(fp, fm, uu, Fb, alpha, eps, c, sL, sLa);
#pragma unroll
            for (int t = 0; t < 6; t++) sm.Gx[t][ty][tx] = c[t];
        }
        {  // y face y0+ty+1/2: LDS rows ty..ty+5
            double fp[6], fm[6], uu[6];
#pragma unroll
            for (int t = 0; t < 6; t++) {
                const double2 pv = sm.P[ty + t][tx + 2];
                fp[t] = pv.x;
                fm[t] = pv.y;
                uu[t] = sm.U[ty + t][tx + 2];
            }
            const double Fb = (sm.Lm[ty + 1][tx] - sm.Lm[ty][tx]) * rdy;
            face_vjp<WENO, P>(fp, fm, uu, Fb, alpha, eps, c, sL, sLa);
#pragma unroll
            for (int t = 0; t < 6; t++) sm.Gy[t][ty][tx] = c[t];
        }
        {  // z face k+1/2 from the register queues
            const double Fb = (coef * ub1 - coef * ub0) * rdz;
            face_vjp<WENO, P>(fpq, fmq, uq, Fb, alpha, eps, c, sL, sLa);
#pragma unroll
            for (int t = 0; t < 6; t++) az[t] += c[t];
        }
        __syncthreads();

        // own node (plane k): pass-through and the x/y faces of the tile
        const unsigned o = col + plane * static_cast<unsigned>(k);
        double pass;
        if (stage == 3) {
            pass = (2.0 / 3.0) * ub0;
            acc[o] = ub0 * (1.0 / 3.0);
        } else if (stage == 2) {
            pass = 0.25 * ub0;
            acc[o] = ac + 0.75 * ub0;
        } else {
            pass = ac + ub0;
        }
        const double pxy = pass + (gather(sm.Gx, ty, tx, true) + gather(sm.Gy, tx, ty, false));
        if (fabs(uq[2]) == alpha) atomicMin(&sc->idx[stage - 1], o);
        // x/y contributions leaving the tile (by the halo-role threads)
        if (tid < 5 * TY) {
            const int r = tid / 5, c5 = tid % 5;
            const int tt = c5 < 2 ? c5 - 2 : TX + c5 - 2;
            exx[static_cast<unsigned>(wrapi(x0 + tt, nx)) + static_cast<unsigned>(nx) * (y0 + r) +
                plane * static_cast<unsigned>(k)] = gather(sm.Gx, r, tt, true);
        } else if (has_halo) {
            const int h = tid - 5 * TY, cc = h % TX, q = h / TX;
            const int tt = q < 2 ? q - 2 : TY + q - 2;
            exy[static_cast<unsigned>(x0 + cc) + static_cast<unsigned>(nx) * wrapi(y0 + tt, ny) +
                plane * static_cast<unsigned>(k)] = gather(sm.Gy, cc, tt, false);
        }
        // z: node k-2 has all its own faces now
        if (k - 2 >= z0)
            ubi[col + plane * static_cast<unsigned>(k - 2)] = dq[0] + az[0];
        else
            exz[col + zoff(k - 2)] = az[0];
        dq[0] = dq[1];
        dq[1] = pxy;
#pragma unroll
        for (int q = 0; q < 5; q++) {
            az[q] = az[q + 1];
            fpq[q] = fpq[q + 1];
            fmq[q] = fmq[q + 1];
            uq[q] = uq[q + 1];
        }
        az[5] = 0.0;
        ub0 = ub1;
    }
    // nodes z1-2, z1-1 (their faces z1, z1+1 belong to the chunk above) and
    // the contributions to z1..z1+2
    ubi[col + plane * static_cast<unsigned>(z1 - 2)] = dq[0] + az[0];
    ubi[col + plane * static_cast<unsigned>(z1 - 1)] = dq[1] + az[1];
#pragma unroll
    for (int q = 2; q < 5; q++) exz[col + zoff(z1 + q - 2)] = az[q];

    // block partial sums of <lam, L> and <lam, dL/dalpha> (fixed order)
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sL += __shfl_xor(sL, off, 64);
        sLa += __shfl_xor(sLa, off, 64);
    }
    if (lane == 0) {
        sm.red[0][wave] = sL;
        sm.red[1][wave] = sLa;
    }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < NW; w++) {
            a += sm.red[0][w];
            b += sm.red[1][w];
        }
        partials[2 * blockIdx.x] = a;
        partials[2 * blockIdx.x + 1] = b;
    }
}

// Adds the exchange fields to the nodes that received contributions from a
// neighbouring tile or chunk: x offsets 0..2 and TX-2..TX-1 within the tile
// (exx), the same in y (exy), and planes z0..z0+2, z1-2..z1-1 of the chunk
// (exz).  Every such node's exchange value was written by exactly one block.
template <int TX, int TY>
__global__ __launch_bounds__(256) void k_vjp_inflow(double *ubi, const double *__restrict__ exx,
                                                    const double *__restrict__ exy, const double *__restrict__ exz,
                                                    VjpGrid g) {
    const unsigned plane = static_cast<unsigned>(g.nx) * static_cast<unsigned>(g.ny);
    const size_t n = static_cast<size_t>(plane) * g.nz;
    for (size_t o = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x; o < n; o += 256ull * gridDim.x) {
        const unsigned x = static_cast<unsigned>(o % g.nx), y = static_cast<unsigned>((o / g.nx) % g.ny);
        const int z = static_cast<int>(o / plane);
        const int xi = x % TX, yi = y % TY;
        const int bz = min(z / g.kchunk, g.nchunks - 1);
        const int z0 = bz * g.kchunk, z1 = bz == g.nchunks - 1 ? g.nz : z0 + g.kchunk;
        const bool ix = xi <= 2 || xi >= TX - 2, iy = yi <= 2 || yi >= TY - 2, iz = z - z0 <= 2 || z1 - z <= 2;
        if (!(ix || iy || iz)) continue;
        double s = 0.0;
        if (ix) s += exx[o];
        if (iy) s += exy[o];
        if (iz) s += exz[o];
        ubi[o] += s;
    }
}

// One block: sums the stage's partials (x, y, z blocks; each thread a fixed
// strided subset, then a fixed-order tree), routes the alpha adjoint to its
// argmax node and, at stage 1, closes the dt chain.
__global__ __launch_bounds__(256) void k_adj_scalars(const double *partials, int nparts, const StepRec *rec,
                                                     int stage, double *ubi, const double *uin, AdjScalars *sc,
                                                     int clipped_last) {
    __shared__ double red[2][256];
    double L = 0.0, La = 0.0;
    for (int b = threadIdx.x; b < nparts; b += 256) {
        L += partials[2 * b];
        La += partials[2 * b + 1];
    }
    red[0][threadIdx.x] = L;
    red[1][threadIdx.x] = La;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    L = red[0][0];
    La = red[1][0];
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

namespace {

constexpr int kMisfitBlocks = 1024;
constexpr int kVjpTX = 16, kVjpTY = 16, kVjpChunk = 16;

// z chunks of kVjpChunk planes, the last one taking the remainder (nz >= 5).
VjpGrid vjp_grid(int nx, int ny, int nz, double dx, double dy, double dz) {
    VjpGrid g{};
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.kchunk = kVjpChunk;
    g.nchunks = std::max(1, nz / kVjpChunk);
    g.dx = dx;
    g.dy = dy;
    g.dz = dz;
    return g;
}

int vjp_blocks(const VjpGrid &g) { return (g.nx / kVjpTX) * (g.ny / kVjpTY) * g.nchunks; }

template <int WENO, int P>
int launch_stage_vjp(hipStream_t s, int stage, const double *uin, const double *ubo, double *ubi, double *acc,
                     const StepRec *rec, const VjpGrid &g, double eps, double *ex, AdjScalars *sc, double *partials,
                     bool clipped_last) {
    const size_t n = static_cast<size_t>(g.nx) * g.ny * g.nz;
    double *exx = ex, *exy = ex + n, *exz = ex + 2 * n;
    const int nb = vjp_blocks(g);
    hipLaunchKernelGGL((k_vjp_tile<WENO, P, kVjpTX, kVjpTY>), dim3(nb), dim3(kVjpTX * kVjpTY), 0, s, uin, ubo, ubi,
                       acc, rec, stage, g, eps, exx, exy, exz, sc, partials);
    const unsigned ib = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 8192));
    hipLaunchKernelGGL((k_vjp_inflow<kVjpTX, kVjpTY>), dim3(ib), dim3(256), 0, s, ubi, exx, exy, exz, g);
    hipLaunchKernelGGL(k_adj_scalars, dim3(1), dim3(256), 0, s, partials, nb, rec, stage, ubi, uin, sc,
                       clipped_last ? 1 : 0);
    return check_launch();
}

int stage_vjp(hipStream_t s, const mgad_scheme &sc, int stage, const double *uin, const double *ubo, double *ubi,
              double *acc, const StepRec *rec, const VjpGrid &g, double *ex, AdjScalars *scal, double *partials,
              bool clipped_last) {
    if (sc.weno == MGAD_WENO_JS)
        return launch_stage_vjp<kWenoJS, 2>(s, stage, uin, ubo, ubi, acc, rec, g, sc.eps, ex, scal, partials,
                                            clipped_last);
    if (sc.p == 2)
        return launch_stage_vjp<kWenoZ, 2>(s, stage, uin, ubo, ubi, acc, rec, g, sc.eps, ex, scal, partials,
                                           clipped_last);
    return launch_stage_vjp<kWenoZ, 1>(s, stage, uin, ubo, ubi, acc, rec, g, sc.eps, ex, scal, partials,
                                       clipped_last);
}

// Workspace layout (bytes, 256-aligned pieces).
struct AdjLayout {
    size_t n, K, nchk;
    size_t off_chk, off_seg, off_ub, off_ex, off_rec, off_ctrl, off_sc, off_part, off_tchk, total;
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
    L.off_ex = o; o += align256(3 * L.n * sizeof(double));  // VJP exchange fields x, y, z
    L.off_rec = o; o += align256(L.K * sizeof(StepRec));
    L.off_ctrl = o; o += align256(MGAD_CTRL_BYTES);
    L.off_sc = o; o += align256(sizeof(AdjScalars));
    const int nb = vjp_blocks(vjp_grid(nx, ny, nz, 1.0, 1.0, 1.0));
    L.off_part = o; o += align256(sizeof(double) * std::max<size_t>(2 * nb, kMisfitBlocks));
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
        checkpoint_every <= 0 || nx % 16 || ny % 16 || nz < 5 || !(tend > t0) || !(cfl > 0)) {
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
    double *ex = reinterpret_cast<double *>(ws + L.off_ex);
    const VjpGrid vg = vjp_grid(static_cast<int>(nx), static_cast<int>(ny), static_cast<int>(nz), dx, dy, dz);
    const size_t bytes = L.n * sizeof(double);
    const int K = checkpoint_every;

    // Runs steps [0, len) of a segment from S0(0), keeping every stage input
    // (S0(m), S1(m), S2(m)) and the step record; steps past tend are no-ops.
    auto run_segment = [&](int len) {
        int r = 0;
        for (int m = 0; m < len && !r; m++) {
            r = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 1, S0(m), S0(m), S1(m), ctrl,
                                               0, 0);
            if (!r) {
                hipLaunchKernelGGL(k_record, dim3(1), dim3(1), 0, stream, static_cast<const Ctrl *>(ctrl), rec + m, 1,
                                   cfl * dx);
                r = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 2, S1(m), S0(m), S2(m),
                                                   ctrl, 0, 0);
            }
            if (!r) {
                hipLaunchKernelGGL(k_record, dim3(1), dim3(1), 0, stream, static_cast<const Ctrl *>(ctrl), rec + m, 2,
                                   cfl * dx);
                r = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 3, S2(m), S0(m), S0(m + 1),
                                                   ctrl, 0, 0);
            }
        }
        return r;
    };

    // ---- forward pass: segments of K steps, checkpoint at each start ----
    // The segment buffers keep the stage inputs of the segment being run, so
    // the last segment is already in place for the reverse sweep.
    int st = check_hip(hipMemcpyAsync(chk(0), u0, bytes, hipMemcpyDeviceToDevice, stream));
    if (!st) st = check_hip(hipMemcpyAsync(S0(0), u0, bytes, hipMemcpyDeviceToDevice, stream));
    if (!st) st = burgers_init_t<double>(stream, nx, ny, nz, S0(0), ctrl, t0, tend, cfl);
    if (st) return st;
    std::vector<double> tchk{t0};
    std::vector<int> seg_len;
    int total = 0;
    for (;;) {
        if ((st = run_segment(K))) return st;
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
        if ((st = check_hip(hipMemcpyAsync(chk(tchk.size()), S0(K), bytes, hipMemcpyDeviceToDevice, stream))))
            return st;
        if ((st = check_hip(hipMemcpyAsync(S0(0), S0(K), bytes, hipMemcpyDeviceToDevice, stream)))) return st;
        tchk.push_back(cs.t);
    }
    const int nseg = static_cast<int>(seg_len.size());
    // J and ubar^N = u^N - target
    double *ubN = UB(0), *ub2 = UB(1), *ub1 = UB(2), *acc = UB(3);
    const double *uN = S0(seg_len.back());
    hipLaunchKernelGGL(k_adj_reset, dim3(1), dim3(1), 0, stream, scal);
    hipLaunchKernelGGL(k_misfit, dim3(kMisfitBlocks), dim3(256), 0, stream, uN, target, ubN, L.n, partials);
    hipLaunchKernelGGL(k_sum_J, dim3(1), dim3(1), 0, stream, partials, kMisfitBlocks, scal);
    if ((st = check_launch())) return st;

    // ---- reverse sweep, segment by segment ----
    double *ub_next = ubN;  // adjoint of the state after the step being reversed
    for (int c = nseg - 1; c >= 0; c--) {
        const int len = seg_len[c];
        if (c != nseg - 1) {
            // recompute the segment from its checkpoint, keeping every stage input
            if ((st = check_hip(hipMemcpyAsync(S0(0), chk(c), bytes, hipMemcpyDeviceToDevice, stream)))) return st;
            if ((st = burgers_init_t<double>(stream, nx, ny, nz, S0(0), ctrl, tchk[c], tend, cfl))) return st;
            if ((st = run_segment(len))) return st;
        }
        for (int m = len - 1; m >= 0 && !st; m--) {
            const bool last = (c == nseg - 1) && (m == len - 1);
            // stage 3: adjoint of u2 into ub2, acc = 1/3 ubar
            st = stage_vjp(stream, *scheme, 3, S2(m), ub_next, ub2, acc, rec + m, vg, ex, scal, partials, last);
            if (!st)
                st = stage_vjp(stream, *scheme, 2, S1(m), ub2, ub1, acc, rec + m, vg, ex, scal, partials, last);
            // stage 1 writes ubar^n in place of acc
            if (!st)
                st = stage_vjp(stream, *scheme, 1, S0(m), ub1, acc, acc, rec + m, vg, ex, scal, partials, last);
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
