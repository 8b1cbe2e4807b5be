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
// Each face is recomputed once in reverse mode, seeded with its flux adjoint
// Fbar, which gives Fbar times the gradient of the face flux with respect to
// its ten split-flux inputs.  These terms are summed per node, separately for
// f+ and f-, in a fixed order (through LDS for the x and y faces of a tile,
// in register queues along a z march), and the split's chain rule is applied
// once per node.  Every node is written by its owner only, so the result does
// not depend on thread timing.  Block partial sums of <lam, L> and
// <lam, dL/dalpha> give the dt and alpha adjoints, folded by one block of a
// short launch per stage, all on the device.
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

// 1/x for normal x: the hardware reciprocal (relative error below 2^-24 on
// gfx950) and one Newton step, which squares it (below 2^-48).
__device__ __forceinline__ double frcp(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

// Value h of one reconstruction side (same formulas as the forward
// weno_side_s) and, in reverse mode with the seed on h, seed * dh/dv_t for
// t = 0..4: stored in g[t] (ACC false) or added to it (ACC true).
template <int WENO, int P, bool PLUS, bool ACC>
__device__ __forceinline__ double side_vjp(double v0, double v1, double v2, double v3, double v4, double eps,
                                           double seed, double *g) {
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
    const double rS6 = rS * (1.0 / 6.0);
    const double h = (a0 * c0 + a1 * c1 + a2 * c2) * rS6;
    // reverse
    const double Nb = seed * rS6, Sb = -(h * seed) * rS;
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
    double e[5];
    e[0] = s0b + t0b;
    e[1] = -2.0 * s0b - 4.0 * t0b + s1b + t1b;
    e[2] = s0b + 3.0 * t0b - 2.0 * s1b + s2b + 3.0 * t2b;
    e[3] = s1b - t1b - 2.0 * s2b - 4.0 * t2b;
    e[4] = s2b + t2b;
    if (PLUS) {
        e[0] += 2.0 * cb0; e[1] += -7.0 * cb0 - cb1; e[2] += 11.0 * cb0 + 5.0 * cb1 + 2.0 * cb2;
        e[3] += 2.0 * cb1 + 5.0 * cb2; e[4] += -cb2;
    } else {
        e[0] += -cb0; e[1] += 5.0 * cb0 + 2.0 * cb1; e[2] += 2.0 * cb0 + 5.0 * cb1 + 11.0 * cb2;
        e[3] += -cb1 - 7.0 * cb2; e[4] += 2.0 * cb2;
    }
#pragma unroll
    for (int t = 0; t < 5; t++) g[t] = ACC ? g[t] + e[t] : e[t];
    return h;
}

// The split fluxes f+- = (u^2/2 +- alpha u)/2 have du-derivatives hu +- ha
// and dalpha-derivatives +-hu, with hu = u/2 and ha = alpha/2.  A node whose
// split-flux adjoints are Pb (of f+) and Mb (of f-) therefore receives
// hu (Pb + Mb) + ha (Pb - Mb) in its u adjoint and adds hu (Pb - Mb) to
// <lam, dL/dalpha>.  Applying this once per node, after summing the face
// terms, replaces applying it to each of the ten stencil terms of a face.
__device__ __forceinline__ double split_chain(double hu, double ha, double Pb, double Mb, double &sLa) {
    const double d = Pb - Mb;
    sLa += hu * d;
    return hu * (Pb + Mb) + ha * d;
}

// Device-side scalars of the reverse sweep.
struct AdjScalars {
    double dtb;                // accumulated dJ/d(dt) of the current step
    double carry;              // dJ/d(dt_last) of a clipped last step
    double J;
    unsigned idx[3];           // first argmax|u| of the stage inputs (stage 1..3)
    unsigned pad;
};

// Points the run's step record at rec, starting with the current step.
__global__ void k_ctrl_rec(Ctrl *c, StepRec *rec) {
    c->rec = rec;
    c->rec_step0 = c->steps;
}

// Reverse of the face i+1/2 whose split fluxes are Pp[t*st], Pm[t*st] for
// nodes t = 0..5 (i-2..i+3), seeded with Fb: stores Fb * dF/df+ of nodes
// 0..4 at G[t*ga] (t = 0..4) and Fb * dF/df- of nodes 1..5 at
// G[(5+t)*ga]; returns the face flux.  The two sides are evaluated one
// after the other (a scheduling barrier keeps the compiler from
// interleaving them), which lowers the live registers.
template <int WENO, int P>
__device__ __forceinline__ double face_vjp_lds(const double *Pp, const double *Pm, int st, double Fb, double eps,
                                               double *G, int ga) {
    double g[5];
    const double hp = side_vjp<WENO, P, true, false>(Pp[0], Pp[st], Pp[2 * st], Pp[3 * st], Pp[4 * st], eps, Fb, g);
#pragma unroll
    for (int t = 0; t < 5; t++) G[t * ga] = g[t];
    __builtin_amdgcn_sched_barrier(0);
    const double hm =
        side_vjp<WENO, P, false, false>(Pm[st], Pm[2 * st], Pm[3 * st], Pm[4 * st], Pm[5 * st], eps, Fb, g);
#pragma unroll
    for (int t = 0; t < 5; t++) G[(5 + t) * ga] = g[t];
    return hp + hm;
}

// Grid of the VJP launches: extents, spacings, the number of plane ranges of
// the x/y blocks (nxy) and of the z march (nzc).  Range c of n covers planes
// [c*nz/n, (c+1)*nz/n); z ranges have at least 5 planes.
struct VjpGrid {
    int nx, ny, nz, nxy, nzc;
    double dx, dy, dz;
};

__host__ __device__ inline int range_lo(int c, int n, int nz) {
    return static_cast<int>(static_cast<long long>(c) * nz / n);
}

// LDS of one x/y VJP tile: split fluxes of the current plane with a halo of
// two nodes below and three above the tile in x and y, lambda with one node
// above, and the face terms of one axis at a time.  G holds ten arrays
// (Fb dF/df+ of stencil nodes 0..4, then Fb dF/df- of nodes 1..5) of
// (TY+5) rows by (TX+5) columns; the face at tile offset f of tile row or
// column r is stored at row r+5, column f+5 for x faces and row f+5, column
// r+5 for y faces.  Only rows and columns 5..T+4 of each array are ever
// written: the rest stays zero, so a gather over faces f-5..f+4 of either
// axis reads zeros for faces outside the tile and needs no bounds tests.
template <int TX, int TY>
struct VjpSmem {
    static constexpr int GW = TX + 5, GA = (TY + 5) * GW, GN = 10 * GA + 5 * GW;
    double G[GN];
    double Pp[TY + 5][TX + 5];  // f+
    double Pm[TY + 5][TX + 5];  // f-
    double Lm[TY + 1][TX + 1];
    double red[2][TX * TY / 64];
};

// Block partial sums of <lam, L> and <lam, dL/dalpha> (fixed order).
template <int NT>
__device__ __forceinline__ void block_partials(double sL, double sLa, double (&red)[2][NT / 64], double *out) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        sL += __shfl_xor(sL, off, 64);
        sLa += __shfl_xor(sLa, off, 64);
    }
    if (lane == 0) {
        red[0][wave] = sL;
        red[1][wave] = sLa;
    }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < NT / 64; w++) {
            a += red[0][w];
            b += red[1][w];
        }
        out[2 * blockIdx.x] = a;
        out[2 * blockIdx.x + 1] = b;
    }
}

// x and y part of the stage VJP of the periodic RHS: for every node,
// ubi = pass-through + the contributions of the x and y faces of its tile,
// with Fb = (lam_{i+1} - lam_i)/d and lam = coef * ubo.  A block owns a
// TX x TY column tile over a range of planes and computes the tile's own
// faces x0..x0+TX-1 (+1/2) and y0..y0+TY-1 (+1/2) once each, one x and one y
// face per thread and plane; the x faces' terms are gathered before the y
// faces reuse the LDS arrays.  Contributions to nodes of a neighbouring tile
// (two below, three above in each axis) go to the exchange fields exx / exy;
// k_vjp_z adds them.  Pass-through terms: stage 3 ubi = 2/3 ubo,
// acc = 1/3 ubo; stage 2 ubi = 1/4 ubo, acc += 3/4 ubo; stage 1 ubi (= acc,
// in place) = acc + ubo.
template <int WENO, int P, int TX, int TY>
__global__ __launch_bounds__(TX * TY, 3) void k_vjp_xy(
    const double *__restrict__ u, const double *__restrict__ ubo, double *ubi, double *acc, const StepRec *rec,
    int stage, VjpGrid g, double eps, double *__restrict__ exx, double *__restrict__ exy,
    AdjScalars *sc, double *partials) {
    using Smem = VjpSmem<TX, TY>;
    constexpr int NT = TX * TY, GW = Smem::GW, GA = Smem::GA;
    static_assert(NT % 64 == 0 && TX >= 5 && TY >= 5, "tile shape");
    static_assert(5 * (TX + TY) + TX + TY <= NT, "one halo role per thread");
    __shared__ Smem sm;
    // the argmax slot of the stage processed before this one (reverse order
    // 3, 2, 1, 3, ...) has been used by k_vjp_tail: clear it for its next use
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->idx[stage == 3 ? 0 : stage] = 0xffffffffu;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int ntx = nx / TX, nty = ny / TY;
    const unsigned tile = xcd_tile(blockIdx.x, gridDim.x);
    const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
    const int x0 = bx * TX, y0 = by * TY, z0 = range_lo(bz, g.nxy, nz), z1 = range_lo(bz + 1, g.nxy, nz);
    const int tid = threadIdx.x, tx = tid % TX, ty = tid / TX;
    const double dt = rec->dt;
    const double alpha = stage == 1 ? rec->a1 : (stage == 2 ? rec->a2 : rec->a3);
    const double ha = 0.5 * alpha;
    const double coef = (stage == 1 ? 1.0 : (stage == 2 ? 0.25 : 2.0 / 3.0)) * dt;
    const double rdx = 1.0 / g.dx, rdy = 1.0 / g.dy;
    const unsigned plane = static_cast<unsigned>(nx) * static_cast<unsigned>(ny);
    const unsigned col = static_cast<unsigned>(x0 + tx) + static_cast<unsigned>(nx) * (y0 + ty);
    for (int i = tid; i < Smem::GN; i += NT) sm.G[i] = 0.0;

    // u halo role: 5*TY x-halo nodes (x0-2, x0-1, x0+TX..x0+TX+2), then 5*TX
    // y-halo nodes; the same thread later gathers the tile's contributions to
    // that node (its exchange value), tile offset ot in the halo's axis
    int hr = 0, hc = 0, ot = 0, orr = 0;
    unsigned hcol = 0;
    const bool xrole = tid < 5 * TY, has_halo = tid < 5 * (TX + TY);
    if (xrole) {
        const int r = tid / 5, c = tid % 5;
        hr = r + 2;
        hc = c < 2 ? c : TX + c;
        ot = hc - 2;
        orr = r;
        hcol = static_cast<unsigned>(wrapi(x0 + ot, nx)) + static_cast<unsigned>(nx) * (y0 + r);
    } else if (has_halo) {
        const int h = tid - 5 * TY, c = h % TX, q = h / TX;
        hr = q < 2 ? q : TY + q;
        hc = c + 2;
        ot = hr - 2;
        orr = c;
        hcol = static_cast<unsigned>(x0 + c) + static_cast<unsigned>(nx) * wrapi(y0 + ot, ny);
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

    // split-flux adjoints of the node at tile offset n of row / column r from
    // the faces n-3..n+2 of the axis whose terms G holds
    auto gather_x = [&](int r, int n, double &Pb, double &Mb) {
        const double *q = &sm.G[(r + 5) * GW + n + 5];
        double p = 0.0, m = 0.0;
#pragma unroll
        for (int t = 0; t < 5; t++) {
            p += q[t * GA + 2 - t];
            m += q[(5 + t) * GA + 1 - t];
        }
        Pb = p;
        Mb = m;
    };
    auto gather_y = [&](int r, int n, double &Pb, double &Mb) {
        const double *q = &sm.G[(n + 5) * GW + r + 5];
        double p = 0.0, m = 0.0;
#pragma unroll
        for (int t = 0; t < 5; t++) {
            p += q[t * GA + (2 - t) * GW];
            m += q[(5 + t) * GA + (1 - t) * GW];
        }
        Pb = p;
        Mb = m;
    };

    double sL = 0.0, sLa = 0.0;
    double u_next = u[col + plane * static_cast<unsigned>(z0)];
    double h_next = has_halo ? u[hcol + plane * static_cast<unsigned>(z0)] : 0.0;
    double ub_next = ubo[col + plane * static_cast<unsigned>(z0)];
    double lh_next = has_lh ? ubo[lcol + plane * static_cast<unsigned>(z0)] : 0.0;
    double ac_next = stage == 3 ? 0.0 : acc[col + plane * static_cast<unsigned>(z0)];
    double *const Gf = &sm.G[(ty + 5) * GW + tx + 5];  // this thread's face slot
    for (int k = z0; k < z1; k++) {
        const double uv = u_next, hv = h_next, ub = ub_next, lh = lh_next, ac = ac_next;
        if (k + 1 < z1) {
            const unsigned pk = plane * static_cast<unsigned>(k + 1);
            u_next = u[col + pk];
            if (has_halo) h_next = u[hcol + pk];
            ub_next = ubo[col + pk];
            if (has_lh) lh_next = ubo[lcol + pk];
            if (stage != 3) ac_next = acc[col + pk];
        }
        {
            double fp, fm;
            lf_split(uv, alpha, fp, fm);
            sm.Pp[ty + 2][tx + 2] = fp;
            sm.Pm[ty + 2][tx + 2] = fm;
            sm.Lm[ty][tx] = coef * ub;
        }
        if (has_halo) {
            double hp, hm;
            lf_split(hv, alpha, hp, hm);
            sm.Pp[hr][hc] = hp;
            sm.Pm[hr][hc] = hm;
        }
        if (has_lh) sm.Lm[lr][lc] = coef * lh;
        __syncthreads();

        const unsigned pk = plane * static_cast<unsigned>(k);
        {  // x face x0+tx+1/2: nodes tile-x tx-2..tx+3 = LDS cols tx..tx+5
            const double Fb = (sm.Lm[ty][tx + 1] - sm.Lm[ty][tx]) * rdx;
            sL += face_vjp_lds<WENO, P>(&sm.Pp[ty + 2][tx], &sm.Pm[ty + 2][tx], 1, Fb, eps, Gf, GA) * Fb;
        }
        __syncthreads();
        double Px, Mx;
        gather_x(ty, tx, Px, Mx);
        if (xrole) {
            double Po, Mo;
            gather_x(orr, ot, Po, Mo);
            exx[hcol + pk] = split_chain(0.5 * hv, ha, Po, Mo, sLa);
        }
        __syncthreads();
        {  // y face y0+ty+1/2: LDS rows ty..ty+5
            const double Fb = (sm.Lm[ty + 1][tx] - sm.Lm[ty][tx]) * rdy;
            sL += face_vjp_lds<WENO, P>(&sm.Pp[ty][tx + 2], &sm.Pm[ty][tx + 2], TX + 5, Fb, eps, Gf, GA) * Fb;
        }
        __syncthreads();
        double Py, My;
        gather_y(tx, ty, Py, My);
        if (has_halo && !xrole) {
            double Po, Mo;
            gather_y(orr, ot, Po, Mo);
            exy[hcol + pk] = split_chain(0.5 * hv, ha, Po, Mo, sLa);
        }

        const unsigned o = col + pk;
        double pass;
        if (stage == 3) {
            pass = (2.0 / 3.0) * ub;
            acc[o] = ub * (1.0 / 3.0);
        } else if (stage == 2) {
            pass = 0.25 * ub;
            acc[o] = ac + 0.75 * ub;
        } else {
            pass = ac + ub;
        }
        ubi[o] = pass + split_chain(0.5 * uv, ha, Px + Py, Mx + My, sLa);
        if (fabs(uv) == alpha) atomicMin(&sc->idx[stage - 1], o);
    }
    block_partials<NT>(sL, sLa, sm.red, partials);
}

// z part of the stage VJP, one thread per column and z chunk (lanes along x):
// marches the chunk's own z faces z0..z1-1 (+1/2) with register queues of the
// split fluxes and u/2 of planes k-2..k+3 and split-flux adjoint
// accumulators, and adds, for each node of the chunk, the contribution of
// its own z faces and the x/y exchange values k_vjp_xy left for it (x
// offsets 0..2 and TX-2..TX-1 within its tile, the same in y).  Node k-2 is
// finished once its last own face, k+1/2, is done; contributions to nodes
// of the chunks below (z0-2, z0-1) and above (z1..z1+2) go to exz, which
// k_vjp_tail adds.
template <int WENO, int P, int TX, int TY>
__global__ __launch_bounds__(256, 3) void k_vjp_z(const double *__restrict__ u, const double *__restrict__ ubo,
                                               double *ubi, const StepRec *rec, int stage, VjpGrid g, double eps,
                                               const double *__restrict__ exx, const double *__restrict__ exy,
                                               double *__restrict__ exz, double *partials) {
    __shared__ double red[2][4];
    const unsigned plane = static_cast<unsigned>(g.nx) * static_cast<unsigned>(g.ny);
    const size_t gid = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    double sL = 0.0, sLa = 0.0;
    if (gid < static_cast<size_t>(plane) * g.nzc) {
        const unsigned col = static_cast<unsigned>(gid % plane);
        const int bz = static_cast<int>(gid / plane);
        const int nz = g.nz, z0 = range_lo(bz, g.nzc, nz), z1 = range_lo(bz + 1, g.nzc, nz);
        const int x = static_cast<int>(col % g.nx), y = static_cast<int>(col / g.nx);
        const bool ix = x % TX <= 2 || x % TX >= TX - 2, iy = y % TY <= 2 || y % TY >= TY - 2;
        const double dt = rec->dt;
        const double alpha = stage == 1 ? rec->a1 : (stage == 2 ? rec->a2 : rec->a3);
        const double ha = 0.5 * alpha;
        const double coef = (stage == 1 ? 1.0 : (stage == 2 ? 0.25 : 2.0 / 3.0)) * dt;
        const double rdz = 1.0 / g.dz;
        auto zo = [&](int k) { return col + plane * static_cast<unsigned>(wrapi(k, nz)); };
        // ubi, exx and exy of the next node to finish, loaded one face ahead
        double fb = 0.0, fx = 0.0, fy = 0.0;
        auto fetch = [&](int k) {
            const unsigned o = col + plane * static_cast<unsigned>(k);
            fb = ubi[o];
            fx = ix ? exx[o] : 0.0;
            fy = iy ? exy[o] : 0.0;
        };
        auto finish = [&](int k, double zc) {
            ubi[col + plane * static_cast<unsigned>(k)] = fb + ((zc + fx) + fy);
        };
        // queues over nodes k-2..k+3; ap / am: adjoints of f+ / f- of nodes k-2..k+2
        double fpq[6], fmq[6], hq[6], ap[5], am[5];
#pragma unroll
        for (int q = 0; q < 5; q++) {
            const double v = u[zo(z0 - 2 + q)];
            lf_split(v, alpha, fpq[q], fmq[q]);
            hq[q] = 0.5 * v;
            ap[q] = 0.0;
            am[q] = 0.0;
        }
        double l0 = coef * ubo[zo(z0)];
        double u_next = u[zo(z0 + 3)], ub_next = ubo[zo(z0 + 1)];
        for (int k = z0; k < z1; k++) {
            const double v = u_next;
            const double l1 = coef * ub_next;
            if (k + 1 < z1) {
                u_next = u[zo(k + 4)];
                ub_next = ubo[zo(k + 2)];
            }
            lf_split(v, alpha, fpq[5], fmq[5]);
            hq[5] = 0.5 * v;
            const double Fb = (l1 - l0) * rdz;
            double gm[5];
            const double hp = side_vjp<WENO, P, true, true>(fpq[0], fpq[1], fpq[2], fpq[3], fpq[4], eps, Fb, ap);
            const double hm =
                side_vjp<WENO, P, false, false>(fmq[1], fmq[2], fmq[3], fmq[4], fmq[5], eps, Fb, gm);
            sL += (hp + hm) * Fb;
            const double c = split_chain(hq[0], ha, ap[0], am[0], sLa);
            if (k - 2 >= z0)
                finish(k - 2, c);
            else
                exz[zo(k - 2)] = c;
            if (k - 1 >= z0) fetch(k - 1);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                ap[q] = ap[q + 1];
                am[q] = am[q + 1] + gm[q];
            }
            ap[4] = 0.0;
            am[4] = gm[4];
#pragma unroll
            for (int q = 0; q < 5; q++) {
                fpq[q] = fpq[q + 1];
                fmq[q] = fmq[q + 1];
                hq[q] = hq[q + 1];
            }
            l0 = l1;
        }
        finish(z1 - 2, split_chain(hq[0], ha, ap[0], am[0], sLa));
        fetch(z1 - 1);
        finish(z1 - 1, split_chain(hq[1], ha, ap[1], am[1], sLa));
#pragma unroll
        for (int q = 2; q < 5; q++) exz[zo(z1 + q - 2)] = split_chain(hq[q], ha, ap[q], am[q], sLa);
    }
    block_partials<256>(sL, sLa, red, partials);
}

// The end of a stage VJP, one launch:
//  - block 0 sums the stage's partials (x, y, z blocks; each thread a fixed
//    strided subset, then a fixed-order tree), routes the alpha adjoint to
//    the argmax node and, at stage 1, closes the dt chain; being dispatched
//    first, it runs alongside the others;
//  - the other blocks add exz to the planes that received z contributions
//    from a neighbouring chunk: z0..z0+2 and z1-2..z1-1 of every chunk (5
//    planes per chunk, each written by exactly one chunk).
// The argmax node is left to block 0, which adds its exz term first, so its
// value is (ubi + exz) + alpha-term as with separate launches.  Its slot is
// cleared by the next stage's k_vjp_xy, not here, since the other blocks
// read it.
__device__ __forceinline__ bool inflow_plane(int z, const VjpGrid &g) {
    for (int c = 0; c < g.nzc; c++) {
        const int z0 = range_lo(c, g.nzc, g.nz), z1 = range_lo(c + 1, g.nzc, g.nz);
        if (z >= z0 && z < z1) return z < z0 + 3 || z >= z1 - 2;
    }
    return false;
}

__global__ __launch_bounds__(256) void k_vjp_tail(double *ubi, const double *__restrict__ exz, VjpGrid g,
                                                  const double *partials, int nparts, const StepRec *rec, int stage,
                                                  const double *uin, AdjScalars *sc, int clipped_last) {
    const unsigned plane = static_cast<unsigned>(g.nx) * static_cast<unsigned>(g.ny);
    const unsigned idx = sc->idx[stage - 1];
    if (blockIdx.x > 0) {
        const size_t n = static_cast<size_t>(plane) * 5 * g.nzc;
        for (size_t i = static_cast<size_t>(blockIdx.x - 1) * 256 + threadIdx.x; i < n;
             i += 256ull * (gridDim.x - 1)) {
            const unsigned col = static_cast<unsigned>(i % plane);
            const int p = static_cast<int>(i / plane), bz = p / 5, j = p % 5;
            const int z0 = range_lo(bz, g.nzc, g.nz), z1 = range_lo(bz + 1, g.nzc, g.nz);
            const int z = j < 3 ? z0 + j : z1 - 5 + j;
            const unsigned o = col + plane * static_cast<unsigned>(z);
            if (o != idx) ubi[o] += exz[o];
        }
        return;
    }
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
    if (idx != 0xffffffffu) {
        double v = ubi[idx];
        if (inflow_plane(static_cast<int>(idx / plane), g)) v += exz[idx];
        ubi[idx] = v + ab * (uin[idx] >= 0.0 ? 1.0 : -1.0);
    }
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
constexpr int kVjpTX = 16, kVjpTY = 16;  // x/y tile

// Number of plane ranges for `units` blocks per range: the smallest count in
// [1, max_n] whose block total fills the resident capacity R to 95 % or more
// (whole waves of R), else the best filling one.
int fill_ranges(size_t units, int max_n, size_t R) {
    int best = 1;
    double best_eff = 0.0;
    for (int n = 1; n <= max_n; n++) {
        const size_t b = units * n;
        const double eff = static_cast<double>(b) / (static_cast<double>((b + R - 1) / R) * R);
        if (eff >= 0.95) return n;
        if (eff > best_eff + 1e-9) {
            best_eff = eff;
            best = n;
        }
    }
    return best;
}

template <int WENO, int P>
size_t resident_blocks(const void *kernel, int threads) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, threads, 0) != hipSuccess || per < 1)
        return 1024;  // not known: ranges then simply fill 1024 blocks
    return static_cast<size_t>(cus) * per;
}

// Plane ranges sized so each launch fills the device's resident blocks.
template <int WENO, int P>
VjpGrid vjp_grid_t(int nx, int ny, int nz, double dx, double dy, double dz) {
    VjpGrid g{};
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.dx = dx;
    g.dy = dy;
    g.dz = dz;
    const size_t tiles = static_cast<size_t>(nx / kVjpTX) * (ny / kVjpTY);
    const size_t Rxy = resident_blocks<WENO, P>(reinterpret_cast<const void *>(&k_vjp_xy<WENO, P, kVjpTX, kVjpTY>),
                                                kVjpTX * kVjpTY);
    g.nxy = fill_ranges(tiles, nz, Rxy);
    const size_t cols = static_cast<size_t>(nx) * ny;  // z blocks of 256 columns
    const size_t Rz = resident_blocks<WENO, P>(reinterpret_cast<const void *>(&k_vjp_z<WENO, P, kVjpTX, kVjpTY>), 256);
    g.nzc = cols % 256 == 0 ? fill_ranges(cols / 256, std::max(1, nz / 5), Rz) : std::max(1, nz / 8);
    return g;
}

VjpGrid vjp_grid(const mgad_scheme &sc, int nx, int ny, int nz, double dx, double dy, double dz) {
    if (sc.weno == MGAD_WENO_JS) return vjp_grid_t<kWenoJS, 2>(nx, ny, nz, dx, dy, dz);
    if (sc.p == 2) return vjp_grid_t<kWenoZ, 2>(nx, ny, nz, dx, dy, dz);
    return vjp_grid_t<kWenoZ, 1>(nx, ny, nz, dx, dy, dz);
}

int vjp_xy_blocks(const VjpGrid &g) { return (g.nx / kVjpTX) * (g.ny / kVjpTY) * g.nxy; }
int vjp_z_blocks(const VjpGrid &g) {
    return static_cast<int>((static_cast<size_t>(g.nx) * g.ny * g.nzc + 255) / 256);
}
// Upper bound of the partial-sum slots over every choice of ranges.
size_t vjp_max_partials(int nx, int ny, int nz) {
    const size_t xy = static_cast<size_t>(nx / kVjpTX) * (ny / kVjpTY) * nz;
    const size_t z = (static_cast<size_t>(nx) * ny * std::max(1, nz / 5) + 255) / 256;
    return 2 * (xy + z);
}

template <int WENO, int P>
int launch_stage_vjp(hipStream_t s, int stage, const double *uin, const double *ubo, double *ubi, double *acc,
                     const StepRec *rec, const VjpGrid &g, double eps, double *ex, AdjScalars *sc, double *partials,
                     bool clipped_last) {
    const size_t n = static_cast<size_t>(g.nx) * g.ny * g.nz;
    double *exx = ex, *exy = ex + n, *exz = ex + 2 * n;
    const int nxy = vjp_xy_blocks(g), nzb = vjp_z_blocks(g);
    hipLaunchKernelGGL((k_vjp_xy<WENO, P, kVjpTX, kVjpTY>), dim3(nxy), dim3(kVjpTX * kVjpTY), 0, s, uin, ubo, ubi,
                       acc, rec, stage, g, eps, exx, exy, sc, partials);
    hipLaunchKernelGGL((k_vjp_z<WENO, P, kVjpTX, kVjpTY>), dim3(nzb), dim3(256), 0, s, uin, ubo, ubi, rec, stage, g,
                       eps, exx, exy, exz, partials + 2 * nxy);
    const size_t nin = static_cast<size_t>(g.nx) * g.ny * 5 * g.nzc;
    const unsigned tail_blocks = static_cast<unsigned>(std::min<size_t>((nin + 255) / 256, 8192)) + 1;
    hipLaunchKernelGGL(k_vjp_tail, dim3(tail_blocks), dim3(256), 0, s, ubi, exz, g, partials, nxy + nzb, rec, stage,
                       uin, sc, clipped_last ? 1 : 0);
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
    L.off_part = o; o += align256(sizeof(double) * std::max<size_t>(vjp_max_partials(nx, ny, nz), kMisfitBlocks));
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
    const VjpGrid vg = vjp_grid(*scheme, static_cast<int>(nx), static_cast<int>(ny), static_cast<int>(nz), dx, dy, dz);
    const size_t bytes = L.n * sizeof(double);
    const int K = checkpoint_every;

    // Runs steps [0, len) of a segment from S0(0), keeping every stage input
    // (S0(m), S1(m), S2(m)) and the step record; steps past tend are no-ops.
    auto run_segment = [&](int len) {
        hipLaunchKernelGGL(k_ctrl_rec, dim3(1), dim3(1), 0, stream, static_cast<Ctrl *>(ctrl), rec);
        int r = check_launch();
        for (int m = 0; m < len && !r; m++) {
            r = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 1, S0(m), S0(m), S1(m), ctrl,
                                               0, 0);
            if (!r)
                r = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 2, S1(m), S0(m), S2(m),
                                                   ctrl, 0, 0);
            if (!r)
                r = burgers_stage_launch_t<double>(stream, *scheme, nx, ny, nz, dx, dy, dz, 3, S2(m), S0(m), S0(m + 1),
                                                   ctrl, 0, 0);
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
