/*
 * oracle.c -- CPU restatement of the reference stencil routines
 * (TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load this file's library, and only as the checker).
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared (oracle/Makefile).
 * No -ffast-math, no contraction: every expression below is evaluated as the
 * IEEE binary64 expression tree that the cited CUDA source spells out after
 * macro expansion (e.g. C1312*(s)*(s) is ((13./12.)*s)*s, D0N*(x) is
 * (1./10.)*x, 6*alphasum is 6.0*alphasum).
 *
 * What pins each function (tests/test_oracle_*.py):
 *  - orc_flux, orc_split ............ closed forms: f+ + f- = f(u), f+ - f- = a*u,
 *                                       local-LF u>=0 gives f+ = 3/4 u^2, f- = -1/4 u^2;
 *                                       monotone split for a >= max|u|.
 *  - orc_face (all variants) ........ constant data gives f(c); quadratic data makes
 *                                       all candidate stencils agree; plus/minus mirror
 *                                       symmetry; 5th-order convergence of the 1-D
 *                                       periodic derivative (a swapped linear weight
 *                                       or coefficient drops the order to <= 3);
 *                                       brute force against a long-double restatement.
 *  - orc_adv_x/y/z .................. constant field gives Lu = 0 on the written set;
 *                                       x/y/z permutation of the data permutes Lu;
 *                                       telescoping of the written differences.
 *  - orc_rhs_periodic ............... sum(Lu) = 0 (periodic conservation); permutation
 *                                       invariance; 1-D data reduces to orc_rhs_1d.
 *  - orc_mb3_adv, orc_mb3_rk ........ x/y on planes 3..nz-4 with pitch nx equal orc_adv_x/y;
 *                                       z equals orc_adv_z on planes 3..nz-4 and splits over
 *                                       plane ranges; RK equals rk_node on every node.
 *  - orc_weno7_face ................. constant data gives hn = f+(c), hp = f-(c); every
 *                                       smoothness form vanishes on constants and its
 *                                       coefficients are a positive semidefinite form;
 *                                       cubic data makes the four candidates agree; mirror
 *                                       symmetry of the two halves; 7th-order convergence.
 *  - orc_weno7_rhs .................. constant field: 0 except node 0 of each line
 *                                       (face below it taken as 0); telescoping.
 *  - orc_lsrk_stage, orc_weno7_run .. u' = lambda*u converges at 4th order; the stage
 *                                       coefficients give b-sums of the classical
 *                                       conditions.
 *  - orc_ssprk_b5, orc_rk3d,
 *    orc_rk_stage_all ............... linear ODE u' = lambda*u gives the SSP-RK3
 *                                       amplification 1 + z + z^2/2 + z^3/6.
 *  - orc_lu3d_o2, orc_heat7_fe ...... Dirichlet sine eigenmodes decay by exactly
 *                                       (1 - dt*lambda)^n; max principle.
 *  - orc_lu3d_13, orc_diff13_b5 ..... exact (to rounding) on polynomials of degree <= 5
 *                                       in each direction (the 5-point 4th-order second
 *                                       difference), and agree with each other.
 *  - orc_burgers_step/run ........... composition of the above; conservation of
 *                                       sum(u) over a periodic run.
 *  - orc_side_grad .................. central finite differences of orc_face per
 *                                       stencil entry (test_side_gradient_fd).
 *  - orc_rhs_vjp .................... dot-product test <lam, J v> = <J^T lam, v> with
 *                                       J v from central differences of orc_rhs_periodic
 *                                       at fixed alpha, and the alpha derivative against
 *                                       central differences in alpha (test_rhs_vjp_dot_product).
 *  - orc_burgers_adjoint ............ central finite differences of J along seeded
 *                                       directions, before and after shocks
 *                                       (test_adjoint_fd).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* B5/Burgers.h:20-28 (same values in MB3/BurgersMPI.h:21-29) */
#define D0N (1. / 10.)
#define D1N (6. / 10.)
#define D2N (3. / 10.)
#define D0P (3. / 10.)
#define D1P (6. / 10.)
#define D2P (1. / 10.)
#define C1312 (13. / 12.)
#define C14 (1. / 4.)
#define C1O6 (1. / 6.) /* B5S/Burgers.h:29 */

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* B5/kernels.cu:37-40:  0.5*u*u  ==  (0.5*u)*u */
double orc_flux(double u) { return 0.5 * u * u; }

/* B5/kernels.cu:224-232 (local alpha = |u|);  MLB/WENO5resAdv1d.m:58-60 (global). */
void orc_split(double u, double a, int lf, double *fp, double *fm) {
    double f = orc_flux(u);
    double au = (lf == ORC_LF_LOCAL) ? fabs(u) * u : a * u;
    *fp = 0.5 * (f + au);
    *fm = 0.5 * (f - au);
}

/* Smoothness indicators, B5/kernels.cu:157-159 (identical in MB3/Kernels.cu:176-178).
 * a..e are the five stencil values (vmm, vm, v, vp, vpp). */
static inline double beta0(double a, double b, double c) {
    return C1312 * (a - 2 * b + c) * (a - 2 * b + c) + C14 * (a - 4 * b + 3 * c) * (a - 4 * b + 3 * c);
}
static inline double beta1(double b, double c, double d) {
    return C1312 * (b - 2 * c + d) * (b - 2 * c + d) + C14 * (b - d) * (b - d);
}
static inline double beta2(double c, double d, double e) {
    return C1312 * (c - 2 * d + e) * (c - 2 * d + e) + C14 * (3 * c - 4 * d + e) * (3 * c - 4 * d + e);
}

/* Candidate polynomials for the left-biased (f+) side, B5/kernels.cu:169-171. */
static inline double pp0(const double *v) { return 2 * v[0] - 7 * v[1] + 11 * v[2]; }
static inline double pp1(const double *v) { return -v[1] + 5 * v[2] + 2 * v[3]; }
static inline double pp2(const double *v) { return 2 * v[2] + 5 * v[3] - v[4]; }
/* Right-biased (f-) side, B5/kernels.cu:186-188. */
static inline double pm0(const double *w) { return -w[0] + 5 * w[1] + 2 * w[2]; }
static inline double pm1(const double *w) { return 2 * w[1] + 5 * w[2] - w[3]; }
static inline double pm2(const double *w) { return 11 * w[2] - 7 * w[3] + 2 * w[4]; }

static inline double zpow(double q, int p) { return p == 2 ? q * q : q; }

/* f+ half of the face flux: v = f+ at nodes i-2..i+2. */
double orc_weno_plus(const double *v, const orc_scheme *s) {
    const double eps = s->eps;
    double B0 = beta0(v[0], v[1], v[2]);
    double B1 = beta1(v[1], v[2], v[3]);
    double B2 = beta2(v[2], v[3], v[4]);
    switch (s->weno) {
    case ORC_WENO_Z: { /* B5/kernels.cu:162-171 */
        double tau5 = fabs(B0 - B2);
        double a0 = D0N * (1. + zpow(tau5 / (B0 + eps), s->p));
        double a1 = D1N * (1. + zpow(tau5 / (B1 + eps), s->p));
        double a2 = D2N * (1. + zpow(tau5 / (B2 + eps), s->p));
        double alphasum = a0 + a1 + a2;
        return (a0 * pp0(v) + a1 * pp1(v) + a2 * pp2(v)) / (6 * alphasum);
    }
    case ORC_WENO_Z_B5S: { /* B5S/kernels.cu:166-176 */
        double tau5 = fabs(B0 - B2);
        double a0 = D0N * (1. + zpow(tau5 / (B0 + eps), s->p));
        double a1 = D1N * (1. + zpow(tau5 / (B1 + eps), s->p));
        double a2 = D2N * (1. + zpow(tau5 / (B2 + eps), s->p));
        double alphasum = 1. / (a0 + a1 + a2);
        return (a0 * pp0(v) + a1 * pp1(v) + a2 * pp2(v)) * (C1O6 * alphasum);
    }
    case ORC_WENO_JS: { /* MB3/Kernels.cu:181-195 */
        double a0 = D0N / ((eps + B0) * (eps + B0));
        double a1 = D1N / ((eps + B1) * (eps + B1));
        double a2 = D2N / ((eps + B2) * (eps + B2));
        double alphasum = a0 + a1 + a2;
        double w0 = a0 / alphasum, w1 = a1 / alphasum, w2 = a2 / alphasum;
        return (w0 * pp0(v) + w1 * pp1(v) + w2 * pp2(v)) / 6;
    }
    default: { /* ORC_WENO_JS_B5, B5/kernels.cu:112-120 */
        double a0 = D0N / ((eps + B0) * (eps + B0));
        double a1 = D1N / ((eps + B1) * (eps + B1));
        double a2 = D2N / ((eps + B2) * (eps + B2));
        double alphasum = a0 + a1 + a2;
        return (a0 * pp0(v) + a1 * pp1(v) + a2 * pp2(v)) / (6 * alphasum);
    }
    }
}

/* f- half of the face flux: w = f- at nodes i-1..i+3. */
double orc_weno_minus(const double *w, const orc_scheme *s) {
    const double eps = s->eps;
    double B0 = beta0(w[0], w[1], w[2]);
    double B1 = beta1(w[1], w[2], w[3]);
    double B2 = beta2(w[2], w[3], w[4]);
    switch (s->weno) {
    case ORC_WENO_Z: { /* B5/kernels.cu:179-188 */
        double tau5 = fabs(B0 - B2);
        double a0 = D0P * (1. + zpow(tau5 / (B0 + eps), s->p));
        double a1 = D1P * (1. + zpow(tau5 / (B1 + eps), s->p));
        double a2 = D2P * (1. + zpow(tau5 / (B2 + eps), s->p));
        double alphasum = a0 + a1 + a2;
        return (a0 * pm0(w) + a1 * pm1(w) + a2 * pm2(w)) / (6 * alphasum);
    }
    case ORC_WENO_Z_B5S: { /* B5S/kernels.cu:189-200 */
        double tau5 = fabs(B0 - B2);
        double a0 = D0P * (1. + zpow(tau5 / (B0 + eps), s->p));
        double a1 = D1P * (1. + zpow(tau5 / (B1 + eps), s->p));
        double a2 = D2P * (1. + zpow(tau5 / (B2 + eps), s->p));
        double alphasum = 1. / (a0 + a1 + a2);
        return (a0 * pm0(w) + a1 * pm1(w) + a2 * pm2(w)) * (C1O6 * alphasum);
    }
    case ORC_WENO_JS: { /* MB3/Kernels.cu:202-216 */
        double a0 = D0P / ((eps + B0) * (eps + B0));
        double a1 = D1P / ((eps + B1) * (eps + B1));
        double a2 = D2P / ((eps + B2) * (eps + B2));
        double alphasum = a0 + a1 + a2;
        double w0 = a0 / alphasum, w1 = a1 / alphasum, w2 = a2 / alphasum;
        return (w0 * pm0(w) + w1 * pm1(w) + w2 * pm2(w)) / 6;
    }
    default: { /* ORC_WENO_JS_B5, B5/kernels.cu:128-136 */
        double a0 = D0P / ((eps + B0) * (eps + B0));
        double a1 = D1P / ((eps + B1) * (eps + B1));
        double a2 = D2P / ((eps + B2) * (eps + B2));
        double alphasum = a0 + a1 + a2;
        return (a0 * pm0(w) + a1 * pm1(w) + a2 * pm2(w)) / (6 * alphasum);
    }
    }
}

/* Face flux F_{i+1/2} = f^+_{i+1/2} + f^-_{i+1/2}:
 * B5/kernels.cu:168-191 (dflux = plus; dflux += minus), MB3/Kernels.cu:218 (hn+hp). */
double orc_face(const double *v, const double *w, const orc_scheme *s) {
    return orc_weno_plus(v, s) + orc_weno_minus(w, s);
}

/* ------------------------------------------------------------------------ */
/* Reference drop-in sweeps, frozen 3-cell ghosts.                           */
/* ------------------------------------------------------------------------ */

/* One line march shared by the three sweeps: B5/kernels.cu:224-258 (x),
 * :291-325 (y), :358-392 (z).  base = linear index of node 0 of the line,
 * st = stride between consecutive nodes, n = line length, d = spacing,
 * mode 0 writes Lu = -(fu-fu_old)/d, mode 1 does Lu -= (fu-fu_old)/d. */
static void march_line(const double *u, double *Lu, size_t base, size_t st,
                       unsigned n, double d, int mode, const orc_scheme *s) {
    double fp[5], fm[5], dummy;
    double fu_old = 0;
    /* f+ at nodes 0..3, f- at nodes 1..4 (B5/kernels.cu:224-232) */
    for (int q = 0; q < 4; q++) orc_split(u[base + st * q], s->alpha, s->lf, &fp[q], &dummy);
    for (int q = 0; q < 4; q++) orc_split(u[base + st * (q + 1)], s->alpha, s->lf, &dummy, &fm[q]);
    for (unsigned i = 2; i < n - 3; i++) {
        orc_split(u[base + st * (i + 2)], s->alpha, s->lf, &fp[4], &dummy);
        orc_split(u[base + st * (i + 3)], s->alpha, s->lf, &dummy, &fm[4]);
        double fu = orc_face(fp, fm, s);
        size_t o = base + st * i;
        if (mode == 0)
            Lu[o] = -(fu - fu_old) / d;
        else
            Lu[o] -= (fu - fu_old) / d;
        fu_old = fu;
        for (int q = 0; q < 4; q++) { fp[q] = fp[q + 1]; fm[q] = fm[q + 1]; }
    }
}

/* Compute_dF, B5/kernels.cu:197-260: rows 2<j<ny-3, 2<k<nz-3, i = 2..nx-4. */
void orc_adv_x(unsigned nx, unsigned ny, unsigned nz, double dx, const double *u,
               double *Lu, const orc_scheme *s) {
    if (nx < 7 || ny < 7 || nz < 7) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 3; k < nz - 3; k++)
        for (unsigned j = 3; j < ny - 3; j++)
            march_line(u, Lu, (size_t)nx * j + (size_t)nx * ny * k, 1, nx, dx, 0, s);
}

/* Compute_dG, B5/kernels.cu:265-327: columns 2<i<nx-3, 2<k<nz-3, j = 2..ny-4. */
void orc_adv_y(unsigned nx, unsigned ny, unsigned nz, double dy, const double *u,
               double *Lu, const orc_scheme *s) {
    if (nx < 7 || ny < 7 || nz < 7) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 3; k < nz - 3; k++)
        for (unsigned i = 3; i < nx - 3; i++)
            march_line(u, Lu, i + (size_t)nx * ny * k, nx, ny, dy, 1, s);
}

/* Compute_dH, B5/kernels.cu:332-394: columns 2<i<nx-3, 2<j<ny-3, k = 2..nz-4. */
void orc_adv_z(unsigned nx, unsigned ny, unsigned nz, double dz, const double *u,
               double *Lu, const orc_scheme *s) {
    if (nx < 7 || ny < 7 || nz < 7) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned j = 3; j < ny - 3; j++)
        for (unsigned i = 3; i < nx - 3; i++)
            march_line(u, Lu, i + (size_t)nx * j, (size_t)nx * ny, nz, dz, 1, s);
}

/* Compute_Laplace_Async, B5/kernels.cu:444-494, launched by Call_Diff_
 * (:560-565) with kstart=3, kstop=nz-2: Lu += 13-point operator on
 * 2<i<nx-3, 2<j<ny-3 and planes k = 3..nz-3, each plane once. */
void orc_diff13_b5(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                   double kz, const double *u, double *Lu) {
    if (nx < 7 || ny < 7 || nz < 7) return;
    const size_t px = nx, px2 = 2 * (size_t)nx, XY = (size_t)nx * ny, XY2 = 2 * XY;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 3; k < nz - 2; k++)
        for (unsigned j = 3; j < ny - 3; j++)
            for (unsigned i = 3; i < nx - 3; i++) {
                size_t o = i + px * j + XY * k;
                double center = u[o];
                double below2 = u[o - XY2], below = u[o - XY], above = u[o + XY], above2 = u[o + XY2];
                Lu[o] += kx * (-u[o - 2] + 16 * u[o - 1] - 30 * center + 16 * u[o + 1] - u[o + 2]) +
                         ky * (-u[o - px2] + 16 * u[o - px] - 30 * center + 16 * u[o + px] - u[o + px2]) +
                         kz * (-below2 + 16 * below - 30 * center + 16 * above - above2);
            }
}

/* One SSP-RK3 stage at one node: B5/kernels.cu:520-527 and D3/kernels.cu:129-139. */
static inline double rk_node(unsigned step, double dt, double u, double uo, double Lu) {
    switch (step) {
    case 1: return uo + dt * (Lu);
    case 2: return 0.75 * uo + 0.25 * (u + dt * (Lu));
    case 3: return (uo + 2 * (u + dt * (Lu))) / 3;
    default: return u;
    }
}

/* Compute_RK, B5/kernels.cu:499-530: interior 2<i<nx-3 (3-cell ghosts frozen). */
void orc_ssprk_b5(unsigned nx, unsigned ny, unsigned nz, unsigned step, double dt,
                  double *u, const double *uo, const double *Lu) {
    if (nx < 7 || ny < 7 || nz < 7) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 3; k < nz - 3; k++)
        for (unsigned j = 3; j < ny - 3; j++)
            for (unsigned i = 3; i < nx - 3; i++) {
                size_t o = i + (size_t)nx * j + (size_t)nx * ny * k;
                u[o] = rk_node(step, dt, u[o], uo[o], Lu[o]);
            }
}

/* Compute_Laplace3d, D3/kernels.cu:66-99: 13-point operator on 1<i<nx-2. */
void orc_lu3d_13(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                 double kz, const double *u, double *Lu) {
    if (nx < 5 || ny < 5 || nz < 5) return;
    const size_t xy = (size_t)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 2; k < nz - 2; k++)
        for (unsigned j = 2; j < ny - 2; j++)
            for (unsigned i = 2; i < nx - 2; i++) {
                size_t o = i + (size_t)nx * j + xy * k;
                size_t nn = o + nx + nx, n = o + nx, s = o - nx, ss = o - nx - nx;
                size_t ee = o + 2, e = o + 1, w = o - 1, ww = o - 2;
                size_t tt = o + xy + xy, t = o + xy, b = o - xy, bb = o - xy - xy;
                Lu[o] = kx * (-u[ee] + 16 * u[e] - 30 * u[o] + 16 * u[w] - u[ww]) +
                        ky * (-u[nn] + 16 * u[n] - 30 * u[o] + 16 * u[s] - u[ss]) +
                        kz * (-u[tt] + 16 * u[t] - 30 * u[o] + 16 * u[b] - u[bb]);
            }
}

/* Compute_Laplace3d_O2, D3/kernels.cu:37-64: 7-point operator on 0<i<nx-1. */
void orc_lu3d_o2(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                 double kz, const double *u, double *Lu) {
    if (nx < 3 || ny < 3 || nz < 3) return;
    const size_t xy = (size_t)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 1; k < nz - 1; k++)
        for (unsigned j = 1; j < ny - 1; j++)
            for (unsigned i = 1; i < nx - 1; i++) {
                size_t o = i + (size_t)nx * j + xy * k;
                size_t n = o + nx, s = o - nx, e = o + 1, w = o - 1, t = o + xy, b = o - xy;
                Lu[o] = 12 * kx * (u[e] - 2 * u[o] + u[w]) +
                        12 * ky * (u[n] - 2 * u[o] + u[s]) +
                        12 * kz * (u[t] - 2 * u[o] + u[b]);
            }
}

/* Compute_RK, D3/kernels.cu:104-142: interior 1<i<nx-2 (2-cell ring frozen). */
void orc_rk3d(unsigned step, unsigned nx, unsigned ny, unsigned nz, double dt,
              double *u, const double *uo, const double *Lu) {
    if (nx < 5 || ny < 5 || nz < 5) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 2; k < nz - 2; k++)
        for (unsigned j = 2; j < ny - 2; j++)
            for (unsigned i = 2; i < nx - 2; i++) {
                size_t o = i + (size_t)nx * j + (size_t)nx * ny * k;
                u[o] = rk_node(step, dt, u[o], uo[o], Lu[o]);
            }
}

/* ------------------------------------------------------------------------ */
/* The blocked, shared-memory and MPI drivers' variants of the operators.    */
/* ------------------------------------------------------------------------ */

/* Compute_Laplace3d_Async as Call_Lu3d of SingleGPU/Diffusion3d_Blocking
 * launches it (kernels.cu:37-88, wrapper :187-191 with kstart = 3,
 * kstop = nz-2): Lu = the 13-point operator with each axis summed from the
 * low side (-u_{-2} + 16 u_{-1} - 30 u + 16 u_{+1} - u_{+2}, :69-71) on
 * 2<i<nx-3, 2<j<ny-3, planes 3..nz-3.  Other nodes keep their Lu. */
void orc_lu3d_blk(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                  double kz, const double *u, double *Lu) {
    if (nx < 7 || ny < 7 || nz < 6) return;
    const size_t px = nx, px2 = 2 * (size_t)nx, XY = (size_t)nx * ny, XY2 = 2 * XY;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 3; k < nz - 2; k++)
        for (unsigned j = 3; j < ny - 3; j++)
            for (unsigned i = 3; i < nx - 3; i++) {
                size_t o = i + px * j + XY * k;
                double center = u[o];
                double below2 = u[o - XY2], below = u[o - XY], above = u[o + XY], above2 = u[o + XY2];
                Lu[o] = kx * (-u[o - 2] + 16 * u[o - 1] - 30 * center + 16 * u[o + 1] - u[o + 2]) +
                        ky * (-u[o - px2] + 16 * u[o - px] - 30 * center + 16 * u[o + px] - u[o + px2]) +
                        kz * (-below2 + 16 * below - 30 * center + 16 * above - above2);
            }
}

/* Node q of a line (base, stride st, length n), 0 outside 0..n-1: the
 * zero pad Compute_dF/dG/dH of Burgers3d_WENO5_SharedMem put beyond the
 * array (kernels.cu:248-260, :313-325, :378-390). */
static inline double line_node(const double *u, size_t base, size_t st, unsigned n, long q) {
    return (q < 0 || q >= (long)n) ? 0.0 : u[base + st * (size_t)q];
}

/* WENO5Zreconstruction of the shared-memory code (kernels.cu:153-205) for
 * the face below node l: f+ of nodes l-3..l+1 and f- of nodes l-2..l+2
 * (its stencil s_u[si-3..si+2], :264), local LF split (:159-163). */
static double b5s_face(const double *u, size_t base, size_t st, unsigned n, long l, const orc_scheme *s) {
    double fp[5], fm[5], dummy;
    for (int q = 0; q < 5; q++) {
        orc_split(line_node(u, base, st, n, l - 3 + q), 0.0, ORC_LF_LOCAL, &fp[q], &dummy);
        orc_split(line_node(u, base, st, n, l - 2 + q), 0.0, ORC_LF_LOCAL, &dummy, &fm[q]);
    }
    return orc_face(fp, fm, s);
}

/* Compute_dF (axis 0), Compute_dG (1), Compute_dH (2) of
 * Burgers3d_WENO5_SharedMem (kernels.cu:212-402), as its driver launches
 * them (main.cpp:94-112: every node of the array): Lu = -(F_{i+1/2} -
 * F_{i-1/2})/dx for axis 0 (:269), Lu -= (..)/d for axes 1 and 2
 * (:334, :399).  s->weno is ORC_WENO_Z_B5S (p = 1, eps = 1e-6 there,
 * Burgers.h:26). */
void orc_adv_b5s(int axis, unsigned nx, unsigned ny, unsigned nz, double d,
                 const double *u, double *Lu, const orc_scheme *s) {
    const size_t xy = (size_t)nx * ny;
    const unsigned n = axis == 0 ? nx : (axis == 1 ? ny : nz);
    const unsigned na = axis == 0 ? ny : nx, nb = axis == 2 ? ny : nz;
    const size_t st = axis == 0 ? 1 : (axis == 1 ? nx : xy);
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned b = 0; b < nb; b++)
        for (unsigned a = 0; a < na; a++) {
            size_t base = axis == 0 ? (size_t)nx * a + xy * b
                        : axis == 1 ? a + xy * b
                                    : a + (size_t)nx * b;
            for (unsigned i = 0; i < n; i++) {
                const double flo = b5s_face(u, base, st, n, (long)i, s);
                const double fhi = b5s_face(u, base, st, n, (long)i + 1, s);
                const size_t o = base + st * i;
                if (axis == 0)
                    Lu[o] = -(fhi - flo) / d;
                else
                    Lu[o] -= (fhi - flo) / d;
            }
        }
}

/* LaplaceO2_async (order 2) and LaplaceO4_async (order 4, the one
 * Call_Diff_ launches) of MultiGPU/Diffusion3d_Baseline (Kernels.cu:152-201,
 * :207-260, wrapper :408-414), on a slab of nz planes with row pitch px:
 *   O2: Lu = dx_ (u_w - 2u + u_e) + dy_ (u_s - 2u + u_n) + dz_ (u_b - 2u + u_t)
 *   O4: Lu = dx_/12 (-u_ww + 16u_w + 16u_e - u_ee - u_ss + 16u_s - 90u
 *                    + 16u_n - u_nn - u_bb + 16u_b + 16u_t - u_tt)
 * (dy_, dz_ unused by O4) on 2<i<nx-3, 2<j<ny-3, planes kstart..kstop-1. */
void orc_md3_lap(int order, unsigned px, unsigned nx, unsigned ny, unsigned nz,
                 unsigned kstart, unsigned kstop, double dxc, double dyc,
                 double dzc, const double *u, double *Lu) {
    const size_t XY = (size_t)px * ny, XY2 = 2 * XY, px2 = 2 * (size_t)px;
    if (nx < 7 || ny < 7 || kstop > nz) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = kstart; k < kstop; k++)
        for (unsigned j = 3; j < ny - 3; j++)
            for (unsigned i = 3; i < nx - 3; i++) {
                const size_t o = i + (size_t)px * j + XY * k;
                const double center = u[o];
                if (order == 2) {
                    const double below = u[o - XY], above = u[o + XY];
                    Lu[o] = dxc * (u[o - 1] - 2 * center + u[o + 1]) +
                            dyc * (u[o - px] - 2 * center + u[o + px]) +
                            dzc * (below - 2 * center + above);
                } else {
                    const double below2 = u[o - XY2], below = u[o - XY], above = u[o + XY], above2 = u[o + XY2];
                    Lu[o] = dxc / 12 * (
                        - u[o - 2] + 16 * u[o - 1] + 16 * u[o + 1] - u[o + 2]
                        - u[o - px2] + 16 * u[o - px] - 90 * center + 16 * u[o + px] - u[o + px2]
                        - below2 + 16 * below + 16 * above - above2);
                }
            }
}

/* Compute_RK of MultiGPU/Diffusion3d_Baseline (Kernels.cu:266-297, Call_sspRK
 * :401-406): SSP-RK3 stage on 2<i<nx-3, 2<j<ny-3 and every plane k < nz
 * (the slab's ghost planes advance too). */
void orc_md3_rk(unsigned step, unsigned px, unsigned nx, unsigned ny, unsigned nz,
                double dt, double *u, const double *uo, const double *Lu) {
    if (nx < 7 || ny < 7) return;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 3; j < ny - 3; j++)
            for (unsigned i = 3; i < nx - 3; i++) {
                const size_t o = i + (size_t)px * j + (size_t)px * ny * k;
                u[o] = rk_node(step, dt, u[o], uo[o], Lu[o]);
            }
}

/* ------------------------------------------------------------------------ */
/* MultiGPU/Burgers3d_Baseline: one rank's z slab (pitched)                 */
/* ------------------------------------------------------------------------ */

/* Compute_dH's line form, MB3/Kernels.cu:376-452: nodes k0..k1-1 of a line,
 * the face below k0 computed from nodes k0-3..k0+2 (not taken as 0). */
static void march_line_from(const double *u, double *Lu, size_t base, size_t st, unsigned k0,
                            unsigned k1, double d, const orc_scheme *s) {
    double fp[5], fm[5], dummy;
    for (int q = 0; q < 5; q++) orc_split(u[base + st * (k0 - 3 + q)], s->alpha, s->lf, &fp[q], &dummy);
    for (int q = 0; q < 5; q++) orc_split(u[base + st * (k0 - 2 + q)], s->alpha, s->lf, &dummy, &fm[q]);
    double fu_old = orc_face(fp, fm, s);
    for (unsigned k = k0; k < k1; k++) {
        for (int q = 0; q < 4; q++) { fp[q] = fp[q + 1]; fm[q] = fm[q + 1]; }
        orc_split(u[base + st * (k + 2)], s->alpha, s->lf, &fp[4], &dummy);
        orc_split(u[base + st * (k + 3)], s->alpha, s->lf, &dummy, &fm[4]);
        double fu = orc_face(fp, fm, s);
        Lu[base + st * k] -= (fu - fu_old) / d;
        fu_old = fu;
    }
}

/* Compute_dF (MB3/Kernels.cu:225-293, axis 0: rows 3<=j<ny-3, Lu =),
 * Compute_dG (:298-371, axis 1: columns 3<=i<nx-3, Lu -=), both on planes
 * kstart..kstop-1 with i (j) = 2..n-4 from fu_old = 0; Compute_dH (:376-452,
 * axis 2: columns 3<=i<nx-3, 3<=j<ny-3, planes kstart..kstop-1).  pitch is the
 * row pitch of u and Lu. */
void orc_mb3_adv(int axis, unsigned pitch, unsigned nx, unsigned ny, unsigned nz, unsigned kstart,
                 unsigned kstop, double d, const double *u, double *Lu, const orc_scheme *s) {
    if (nx < 7 || ny < 7 || nz < 7 || kstart >= kstop) return;
    const size_t XY = (size_t)pitch * ny;
    if (axis == 0) {
#pragma omp parallel for collapse(2) schedule(static)
        for (unsigned k = kstart; k < kstop; k++)
            for (unsigned j = 3; j < ny - 3; j++)
                march_line(u, Lu, (size_t)pitch * j + XY * k, 1, nx, d, 0, s);
    } else if (axis == 1) {
#pragma omp parallel for collapse(2) schedule(static)
        for (unsigned k = kstart; k < kstop; k++)
            for (unsigned i = 3; i < nx - 3; i++)
                march_line(u, Lu, i + XY * k, pitch, ny, d, 1, s);
    } else {
#pragma omp parallel for collapse(2) schedule(static)
        for (unsigned j = 3; j < ny - 3; j++)
            for (unsigned i = 3; i < nx - 3; i++)
                march_line_from(u, Lu, i + (size_t)pitch * j, XY, kstart, kstop, d, s);
    }
}

/* Compute_RK, MB3/Kernels.cu:457-490: every node of the slab, ghosts included. */
void orc_mb3_rk(unsigned step, unsigned pitch, unsigned nx, unsigned ny, unsigned nz, double dt,
                double *q, const double *qo, const double *Lq) {
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++)
            for (unsigned i = 0; i < nx; i++) {
                size_t o = i + (size_t)pitch * j + (size_t)pitch * ny * k;
                q[o] = rk_node(step, dt, q[o], qo[o], Lq[o]);
            }
}

/* ------------------------------------------------------------------------ */
/* WENO7 Burgers prototype (Matlab_Prototipes/InviscidBurgersNd)             */
/* ------------------------------------------------------------------------ */

/* Smoothness indicators of the four 4-point stencils of a 7-point window
 * s[0..6] (the Balsara-Shu quadratic forms), written as WENO7resAdv_X.m:
 * 25-32 writes them for v(i-3)..v(i+3); the U side (:58-65) applies the same
 * forms to u(i-2)..u(i+4). */
static void w7_beta(const double *s, double *B) {
    const double a = s[0], b = s[1], c = s[2], d = s[3], e = s[4], f = s[5], g = s[6];
    B[0] = c * (134241 * c - 114894 * d) + a * (56694 * c - 47214 * b + 6649 * a - 22778 * d) +
           25729 * (d * d) + b * (-210282 * c + 85641 * b + 86214 * d);
    B[1] = d * (41001 * d - 30414 * e) + b * (-19374 * c + 3169 * b + 19014 * d - 5978 * e) +
           6649 * (e * e) + c * (33441 * c - 70602 * d + 23094 * e);
    B[2] = e * (33441 * e - 19374 * f) + c * (6649 * c - 30414 * d + 23094 * e - 5978 * f) +
           3169 * (f * f) + d * (41001 * d - 70602 * e + 19014 * f);
    B[3] = f * (85641 * f - 47214 * g) + d * (25729 * d - 114894 * e + 86214 * f - 22778 * g) +
           6649 * (g * g) + e * (134241 * e - 210282 * f + 56694 * g);
}

/* Nonlinear weights alpha_k = g_k/(eps + B_k)^2, w_k = alpha_k/sum
 * (WENO7resAdv_X.m:35-47, :68-80), eps = 1e-6. */
static void w7_weights(const double *B, const double *g, double *w) {
    const double eps = 1e-6;
    double al[4];
    for (int k = 0; k < 4; k++) al[k] = g[k] / ((eps + B[k]) * (eps + B[k]));
    const double sum = al[0] + al[1] + al[2] + al[3];
    for (int k = 0; k < 4; k++) w[k] = al[k] / sum;
}

/* The two halves of the face flux at i+1/2 from the split fluxes of nodes
 * i-3..i+4: hn from V = f+ at i-3..i+3 (WENO7resAdv_X.m:14-52), hp from
 * U = f- at i-2..i+4 (:55-85); candidate terms (w_k * poly_k)/12 summed left
 * to right. */
void orc_weno7_face(const double *V, const double *U, double *hn, double *hp) {
    static const double gn[4] = {1. / 35, 12. / 35, 18. / 35, 4. / 35};
    static const double gp[4] = {4. / 35, 18. / 35, 12. / 35, 1. / 35};
    double B[4], w[4];
    const double *v = V, *u = U + 1;  /* v: i-3..i+3, u: i-2..i+4 */
    w7_beta(v, B);
    w7_weights(B, gn, w);
    *hn = w[0] * (-3 * v[0] + 13 * v[1] - 23 * v[2] + 25 * v[3]) / 12 +
          w[1] * (1 * v[1] - 5 * v[2] + 13 * v[3] + 3 * v[4]) / 12 +
          w[2] * (-1 * v[2] + 7 * v[3] + 7 * v[4] - 1 * v[5]) / 12 +
          w[3] * (3 * v[3] + 13 * v[4] - 5 * v[5] + 1 * v[6]) / 12;
    w7_beta(u, B);
    w7_weights(B, gp, w);
    *hp = w[0] * (1 * u[0] - 5 * u[1] + 13 * u[2] + 3 * u[3]) / 12 +
          w[1] * (-1 * u[1] + 7 * u[2] + 7 * u[3] - 1 * u[4]) / 12 +
          w[2] * (3 * u[2] + 13 * u[3] - 5 * u[4] + 1 * u[5]) / 12 +
          w[3] * (25 * u[3] - 23 * u[4] + 13 * u[5] - 3 * u[6]) / 12;
}

/* Local LF split of the prototype (WENO7resAdv_X.m:8-10 with flux = w.^2/2,
 * dflux = w): V = 0.5*(w^2/2 + |w| w), U = 0.5*(w^2/2 - |w| w). */
static inline void w7_split(double w, double *V, double *U) {
    const double f = (w * w) / 2, a = fabs(w);
    *V = 0.5 * (f + a * w);
    *U = 0.5 * (f - a * w);
}

/* df/dx of one line of n nodes (stride st), WENO7resAdv_X.m: ghost nodes
 * repeat the end values (3 at the start, 4 at the end, :4), faces
 * 1/2 .. n-1/2 are computed and the face below node 0 is 0 (hn, hp are zero
 * outside I), res = (hp(I) + hn(I) - hp(I-1) - hn(I-1))/dx. */
static void w7_line(const double *w, double *res, size_t base, size_t st, unsigned n, double d, int mode) {
    double V[8], U[8];
    double hn_old = 0.0, hp_old = 0.0;
    for (unsigned i = 0; i < n; i++) {
        for (int q = 0; q < 8; q++) {
            long j = (long)i - 3 + q;
            if (j < 0) j = 0;
            if (j > (long)n - 1) j = (long)n - 1;
            w7_split(w[base + st * (size_t)j], &V[q], &U[q]);
        }
        double hn, hp;
        orc_weno7_face(V, U, &hn, &hp);
        const double r = (hp + hn - hp_old - hn_old) / d;
        if (mode == 0)
            res[base + st * i] = r;
        else
            res[base + st * i] += r;
        hn_old = hn;
        hp_old = hp;
    }
}

/* dF + dG + dH of LFWENO7FDM3d.m:101-103 (x, then + y, then + z). */
void orc_weno7_rhs(unsigned nx, unsigned ny, unsigned nz, double dx, double dy, double dz,
                   const double *u, double *res) {
    const size_t XY = (size_t)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++) w7_line(u, res, (size_t)nx * j + XY * k, 1, nx, dx, 0);
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 0; k < nz; k++)
        for (unsigned i = 0; i < nx; i++) w7_line(u, res, i + XY * k, nx, ny, dy, 1);
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned j = 0; j < ny; j++)
        for (unsigned i = 0; i < nx; i++) w7_line(u, res, i + (size_t)nx * j, XY, nz, dz, 1);
}

/* Low-storage 5-stage 4th-order Runge-Kutta of LFWENO7FDM3d.m:66-82. */
const double orc_lsrk_a[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                              -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
const double orc_lsrk_b[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                              1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                              2277821191437.0 / 14882151754819.0};

/* Stage s (0..4) at one node: res = a_s*res + dt*L; u = u - b_s*res
 * (LFWENO7FDM3d.m:104-105). */
void orc_lsrk_stage(uint64_t n, int s, double dt, double *u, double *res, const double *L) {
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < n; i++) {
        res[i] = orc_lsrk_a[s] * res[i] + dt * L[i];
        u[i] = u[i] - orc_lsrk_b[s] * res[i];
    }
}

/* The time loop of LFWENO7FDM3d.m:88-108: dt = cfl*dx/max|u| clipped to tend,
 * five stages per step; res carries over between steps as there.  Advances
 * u and res in place; returns the steps taken and leaves the time in *t. */
int orc_weno7_run(unsigned nx, unsigned ny, unsigned nz, double dx, double dy, double dz, double cfl,
                  double *t, double tend, int nsteps, double *u, double *res) {
    const size_t N = (size_t)nx * ny * nz;
    double *L = (double *)malloc(sizeof(double) * N);
    int it = 0;
    while (*t < tend && (nsteps < 0 || it < nsteps)) {
        double dt = cfl * dx / orc_maxabs(N, u);
        if (*t + dt > tend) dt = tend - *t;
        *t = *t + dt;
        for (int s = 0; s < 5; s++) {
            orc_weno7_rhs(nx, ny, nz, dx, dy, dz, u, L);
            orc_lsrk_stage(N, s, dt, u, res, L);
        }
        it++;
    }
    free(L);
    return it;
}

/* ------------------------------------------------------------------------ */
/* Periodic global-LF Burgers                                               */
/* ------------------------------------------------------------------------ */

double orc_maxabs(uint64_t n, const double *u) {
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (int64_t q = 0; q < (int64_t)n; q++) {
        double a = fabs(u[q]);
        if (a > m) m = a;
    }
    return m;
}

static inline unsigned wrap(long i, unsigned n) {
    long r = i % (long)n;
    return (unsigned)(r < 0 ? r + (long)n : r);
}

/* Face F at (i+1/2) along axis `ax` for every node, periodic wrap.
 * fp/fm are the split fluxes of every node. */
static void faces_periodic(unsigned nx, unsigned ny, unsigned nz, int ax,
                           const double *fp, const double *fm, double *F,
                           const orc_scheme *s) {
    const size_t xy = (size_t)nx * ny;
    const unsigned n = ax == 0 ? nx : (ax == 1 ? ny : nz);
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++)
            for (unsigned i = 0; i < nx; i++) {
                unsigned c = ax == 0 ? i : (ax == 1 ? j : k);
                double v[5], w[5];
                for (int q = 0; q < 5; q++) {
                    unsigned a = wrap((long)c - 2 + q, n); /* f+ at c-2..c+2 */
                    unsigned b = wrap((long)c - 1 + q, n); /* f- at c-1..c+3 */
                    size_t oa = ax == 0 ? a + (size_t)nx * j + xy * k
                              : ax == 1 ? i + (size_t)nx * a + xy * k
                                        : i + (size_t)nx * j + xy * a;
                    size_t ob = ax == 0 ? b + (size_t)nx * j + xy * k
                              : ax == 1 ? i + (size_t)nx * b + xy * k
                                        : i + (size_t)nx * j + xy * b;
                    v[q] = fp[oa];
                    w[q] = fm[ob];
                }
                F[i + (size_t)nx * j + xy * k] = orc_face(v, w, s);
            }
}

/* Periodic restatement of the Call_Adv_x -> Call_Adv_y -> Call_Adv_z sequence
 * (B5/main.cpp:144-146): Lu = -(F_{i+1/2}-F_{i-1/2})/dx (overwrite, :244),
 * then Lu -= dG/dy (:311), then Lu -= dH/dz (:378), every node updated. */
void orc_rhs_periodic(unsigned nx, unsigned ny, unsigned nz, double dx, double dy,
                      double dz, const double *u, double *Lu, const orc_scheme *s) {
    const size_t N = (size_t)nx * ny * nz, xy = (size_t)nx * ny;
    double *fp = (double *)malloc(sizeof(double) * N);
    double *fm = (double *)malloc(sizeof(double) * N);
    double *F = (double *)malloc(sizeof(double) * N);
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (int64_t)N; q++) orc_split(u[q], s->alpha, s->lf, &fp[q], &fm[q]);

    faces_periodic(nx, ny, nz, 0, fp, fm, F, s);
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++)
            for (unsigned i = 0; i < nx; i++) {
                size_t o = i + (size_t)nx * j + xy * k;
                size_t om = wrap((long)i - 1, nx) + (size_t)nx * j + xy * k;
                Lu[o] = -(F[o] - F[om]) / dx;
            }
    faces_periodic(nx, ny, nz, 1, fp, fm, F, s);
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++)
            for (unsigned i = 0; i < nx; i++) {
                size_t o = i + (size_t)nx * j + xy * k;
                size_t om = i + (size_t)nx * wrap((long)j - 1, ny) + xy * k;
                Lu[o] -= (F[o] - F[om]) / dy;
            }
    faces_periodic(nx, ny, nz, 2, fp, fm, F, s);
#pragma omp parallel for collapse(2) schedule(static)
    for (unsigned k = 0; k < nz; k++)
        for (unsigned j = 0; j < ny; j++)
            for (unsigned i = 0; i < nx; i++) {
                size_t o = i + (size_t)nx * j + xy * k;
                size_t om = i + (size_t)nx * j + xy * wrap((long)k - 1, nz);
                Lu[o] -= (F[o] - F[om]) / dz;
            }
    free(fp);
    free(fm);
    free(F);
}

/* Stage update on every node (periodic grid has no ghosts). */
void orc_rk_stage_all(uint64_t n, unsigned stage, double dt, double *u,
                      const double *uo, const double *Lu) {
#pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < (int64_t)n; q++) u[q] = rk_node(stage, dt, u[q], uo[q], Lu[q]);
}

/* One SSP-RK3 step, MLB/LFWENO5FDM3d.m:70-95 with the CUDA sign convention
 * (Lu = -div F, B5/main.cpp:143-149): dt = CFL*dx/max|u^n| clipped to tEnd
 * (:71); each stage splits with alpha = max|u| of its own input
 * (MLB/WENO5resAdv1d.m:58).  work holds 2*N doubles (uo, Lu). */
double orc_burgers_step(unsigned nx, unsigned ny, unsigned nz, double dx, double dy,
                        double dz, double cfl, double t, double tend,
                        const orc_scheme *s0, double *u, double *work) {
    const size_t N = (size_t)nx * ny * nz;
    double *uo = work, *Lu = work + N;
    orc_scheme s = *s0;
    double amax = orc_maxabs(N, u);
    double dt = cfl * dx / amax;
    if (t + dt > tend) dt = tend - t;
    memcpy(uo, u, sizeof(double) * N);
    for (unsigned stage = 1; stage <= 3; stage++) {
        if (s.lf == ORC_LF_GLOBAL) s.alpha = stage == 1 ? amax : orc_maxabs(N, u);
        orc_rhs_periodic(nx, ny, nz, dx, dy, dz, u, Lu, &s);
        orc_rk_stage_all(N, stage, dt, u, uo, Lu);
    }
    return dt;
}

int orc_burgers_run(unsigned nx, unsigned ny, unsigned nz, double dx, double dy,
                    double dz, double cfl, double *t, double tend, int nsteps,
                    const orc_scheme *s, double *u) {
    const size_t N = (size_t)nx * ny * nz;
    double *work = (double *)malloc(sizeof(double) * 2 * N);
    int it = 0;
    while (*t < tend && (nsteps < 0 || it < nsteps)) {
        double dt = orc_burgers_step(nx, ny, nz, dx, dy, dz, cfl, *t, tend, s, u, work);
        *t = *t + dt;
        it++;
    }
    free(work);
    return it;
}

/* ------------------------------------------------------------------------ */
/* Heat                                                                      */
/* ------------------------------------------------------------------------ */

/* Forward Euler with the 7-point operator: Compute_Laplace3d_O2
 * (D3/kernels.cu:37-64) composed with stage 1 of Compute_RK (:130-132) on the
 * Dirichlet interior 0<i<n-1; boundary nodes are held. */
void orc_heat7_fe(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                  double kz, double dt, int nsteps, double *u) {
    const size_t N = (size_t)nx * ny * nz, xy = (size_t)nx * ny;
    double *Lu = (double *)calloc(N, sizeof(double));
    double *uo = (double *)malloc(sizeof(double) * N);
    for (int it = 0; it < nsteps; it++) {
        memcpy(uo, u, sizeof(double) * N);
        orc_lu3d_o2(nx, ny, nz, kx, ky, kz, uo, Lu);
#pragma omp parallel for collapse(2) schedule(static)
        for (unsigned k = 1; k < nz - 1; k++)
            for (unsigned j = 1; j < ny - 1; j++)
                for (unsigned i = 1; i < nx - 1; i++) {
                    size_t o = i + (size_t)nx * j + xy * k;
                    u[o] = rk_node(1, dt, u[o], uo[o], Lu[o]);
                }
    }
    free(Lu);
    free(uo);
}

/* 1-D periodic derivative -(F_{i+1/2}-F_{i-1/2})/dx (order-of-accuracy checks). */
void orc_rhs_1d_periodic(unsigned n, double dx, const double *u, double *Lu,
                         const orc_scheme *s) {
    double *fp = (double *)malloc(sizeof(double) * n), *fm = (double *)malloc(sizeof(double) * n);
    double *F = (double *)malloc(sizeof(double) * n);
    for (unsigned i = 0; i < n; i++) orc_split(u[i], s->alpha, s->lf, &fp[i], &fm[i]);
    for (unsigned i = 0; i < n; i++) {
        double v[5], w[5];
        for (int q = 0; q < 5; q++) {
            v[q] = fp[wrap((long)i - 2 + q, n)];
            w[q] = fm[wrap((long)i - 1 + q, n)];
        }
        F[i] = orc_face(v, w, s);
    }
    for (unsigned i = 0; i < n; i++) Lu[i] = -(F[i] - F[wrap((long)i - 1, n)]) / dx;
    free(fp);
    free(fm);
    free(F);
}

/* ------------------------------------------------------------------------ */
/* Discrete adjoint (reverse mode) of the periodic global-LF Burgers step.  */
/* ------------------------------------------------------------------------ */
/* Derivative of the face-flux side function of orc_weno_plus / orc_weno_minus
 * (B5/kernels.cu:142-192 for WENO-Z, MB3/Kernels.cu:112-220 for WENO-JS),
 * reverse mode with unit output seed.  The JS normalisation w = alpha/sum,
 * (sum w c)/6, is the same function as (sum alpha c)/(6 sum). */
double orc_side_grad(const double *v, int plus, const orc_scheme *s, double *g) {
    const double eps = s->eps;
    const double s0 = (v[0] - 2 * v[1]) + v[2], t0 = (v[0] - 4 * v[1]) + 3 * v[2];
    const double s1 = (v[1] - 2 * v[2]) + v[3], t1 = v[1] - v[3];
    const double s2 = (v[2] - 2 * v[3]) + v[4], t2 = (3 * v[2] - 4 * v[3]) + v[4];
    const double B[3] = {C1312 * s0 * s0 + C14 * t0 * t0, C1312 * s1 * s1 + C14 * t1 * t1,
                         C1312 * s2 * s2 + C14 * t2 * t2};
    const double d[3] = {plus ? D0N : D0P, plus ? D1N : D1P, plus ? D2N : D2P};
    double c[3];
    if (plus) {
        c[0] = 2 * v[0] - 7 * v[1] + 11 * v[2];
        c[1] = -v[1] + 5 * v[2] + 2 * v[3];
        c[2] = 2 * v[2] + 5 * v[3] - v[4];
    } else {
        c[0] = -v[0] + 5 * v[1] + 2 * v[2];
        c[1] = 2 * v[1] + 5 * v[2] - v[3];
        c[2] = 11 * v[2] - 7 * v[3] + 2 * v[4];
    }
    const int Z = (s->weno == ORC_WENO_Z || s->weno == ORC_WENO_Z_B5S);
    double a[3], b[3], q[3], tau = 0;
    if (Z) tau = fabs(B[0] - B[2]);
    for (int k = 0; k < 3; k++) {
        b[k] = B[k] + eps;
        if (Z) {
            q[k] = tau / b[k];
            a[k] = d[k] * (1. + zpow(q[k], s->p));
        } else {
            a[k] = d[k] / (b[k] * b[k]);
        }
    }
    const double S = a[0] + a[1] + a[2], N = a[0] * c[0] + a[1] * c[1] + a[2] * c[2];
    const double h = N / (6 * S);
    /* reverse */
    const double Nb = 1.0 / (6 * S), Sb = -h / S;
    double ab[3], cb[3], Bb[3] = {0, 0, 0}, taub = 0;
    for (int k = 0; k < 3; k++) {
        ab[k] = Nb * c[k] + Sb;
        cb[k] = Nb * a[k];
        if (Z) {
            const double qb = s->p == 2 ? ab[k] * d[k] * 2 * q[k] : ab[k] * d[k];
            taub += qb / b[k];
            Bb[k] += -qb * q[k] / b[k];
        } else {
            Bb[k] += ab[k] * (-2 * a[k] / b[k]);
        }
    }
    if (Z) {
        const double sg = (B[0] - B[2]) >= 0 ? 1.0 : -1.0;
        Bb[0] += sg * taub;
        Bb[2] -= sg * taub;
    }
    const double s0b = 2 * C1312 * s0 * Bb[0], t0b = 2 * C14 * t0 * Bb[0];
    const double s1b = 2 * C1312 * s1 * Bb[1], t1b = 2 * C14 * t1 * Bb[1];
    const double s2b = 2 * C1312 * s2 * Bb[2], t2b = 2 * C14 * t2 * Bb[2];
    g[0] = s0b + t0b;
    g[1] = -2 * s0b - 4 * t0b + s1b + t1b;
    g[2] = s0b + 3 * t0b - 2 * s1b + s2b + 3 * t2b;
    g[3] = s1b - t1b - 2 * s2b - 4 * t2b;
    g[4] = s2b + t2b;
    if (plus) {
        g[0] += 2 * cb[0];
        g[1] += -7 * cb[0] - cb[1];
        g[2] += 11 * cb[0] + 5 * cb[1] + 2 * cb[2];
        g[3] += 2 * cb[1] + 5 * cb[2];
        g[4] += -cb[2];
    } else {
        g[0] += -cb[0];
        g[1] += 5 * cb[0] + 2 * cb[1];
        g[2] += 2 * cb[0] + 5 * cb[1] + 11 * cb[2];
        g[3] += -cb[1] - 7 * cb[2];
        g[4] += 2 * cb[2];
    }
    return h;
}

/* ubar += (dL/du)^T lam along one axis (line by line, serial per line so the
 * accumulation order is fixed). */
static void vjp_axis(unsigned nx, unsigned ny, unsigned nz, int ax, double d, const double *u,
                     const double *lam, double *ubar, const orc_scheme *s, double *lamL, double *lamLa) {
    const size_t xy = (size_t)nx * ny;
    const unsigned n = ax == 0 ? nx : (ax == 1 ? ny : nz);
    const unsigned na = ax == 0 ? ny : nx, nb = ax == 2 ? ny : nz;
    double sL = 0, sLa = 0;
#pragma omp parallel for collapse(2) reduction(+ : sL, sLa) schedule(static)
    for (unsigned b = 0; b < nb; b++)
        for (unsigned a = 0; a < na; a++) {
            size_t base, st;
            if (ax == 0) { base = (size_t)nx * a + xy * b; st = 1; }
            else if (ax == 1) { base = a + xy * b; st = nx; }
            else { base = a + (size_t)nx * b; st = xy; }
            double *fp = (double *)malloc(sizeof(double) * n * 4);
            double *fm = fp + n, *fpb = fp + 2 * n, *fmb = fp + 3 * n;
            for (unsigned i = 0; i < n; i++) {
                orc_split(u[base + st * i], s->alpha, ORC_LF_GLOBAL, &fp[i], &fm[i]);
                fpb[i] = fmb[i] = 0;
            }
            for (unsigned i = 0; i < n; i++) {
                double v[5], w[5], gv[5], gw[5];
                for (int q = 0; q < 5; q++) {
                    v[q] = fp[wrap((long)i - 2 + q, n)];
                    w[q] = fm[wrap((long)i - 1 + q, n)];
                }
                const double F = orc_side_grad(v, 1, s, gv) + orc_side_grad(w, 0, s, gw);
                const double Fb = (lam[base + st * wrap((long)i + 1, n)] - lam[base + st * i]) / d;
                sL += F * Fb;
                for (int q = 0; q < 5; q++) {
                    fpb[wrap((long)i - 2 + q, n)] += Fb * gv[q];
                    fmb[wrap((long)i - 1 + q, n)] += Fb * gw[q];
                }
            }
            for (unsigned j = 0; j < n; j++) {
                const double uj = u[base + st * j];
                ubar[base + st * j] += fpb[j] * (0.5 * (uj + s->alpha)) + fmb[j] * (0.5 * (uj - s->alpha));
                sLa += 0.5 * uj * (fpb[j] - fmb[j]);
            }
            free(fp);
        }
    *lamL += sL;
    *lamLa += sLa;
}

void orc_rhs_vjp(unsigned nx, unsigned ny, unsigned nz, double dx, double dy, double dz,
                 const double *u, const double *lam, double *ubar, const orc_scheme *s,
                 double *lam_dot_L, double *lam_dot_La) {
    double L = 0, La = 0;
    vjp_axis(nx, ny, nz, 0, dx, u, lam, ubar, s, &L, &La);
    vjp_axis(nx, ny, nz, 1, dy, u, lam, ubar, s, &L, &La);
    vjp_axis(nx, ny, nz, 2, dz, u, lam, ubar, s, &L, &La);
    *lam_dot_L = L;
    *lam_dot_La = La;
}

/* index of the first node with the largest |u| (the subgradient used for
 * alpha = max|u|) */
static size_t argmax_abs(size_t n, const double *u) {
    size_t im = 0;
    double m = -1;
    for (size_t q = 0; q < n; q++)
        if (fabs(u[q]) > m) { m = fabs(u[q]); im = q; }
    return im;
}

double orc_burgers_adjoint(unsigned nx, unsigned ny, unsigned nz, double dx, double dy, double dz,
                           double cfl, double t0, double tend, const orc_scheme *s0,
                           const double *u0, const double *target, double *grad, int *nsteps) {
    const size_t N = (size_t)nx * ny * nz;
    orc_scheme s = *s0;
    s.lf = ORC_LF_GLOBAL;
    /* forward, keeping every stage input */
    int cap = 64, n = 0;
    double **st = (double **)malloc(sizeof(double *) * cap * 3);
    double *dts = (double *)malloc(sizeof(double) * cap), *al = (double *)malloc(sizeof(double) * cap * 3);
    int *clip = (int *)malloc(sizeof(int) * cap);
    double *u = (double *)malloc(sizeof(double) * N), *Lu = (double *)malloc(sizeof(double) * N);
    memcpy(u, u0, sizeof(double) * N);
    double t = t0;
    while (t < tend) {
        if (n == cap) {
            cap *= 2;
            st = (double **)realloc(st, sizeof(double *) * cap * 3);
            dts = (double *)realloc(dts, sizeof(double) * cap);
            al = (double *)realloc(al, sizeof(double) * cap * 3);
            clip = (int *)realloc(clip, sizeof(int) * cap);
        }
        const double a1 = orc_maxabs(N, u);
        double dt = cfl * dx / a1;
        clip[n] = 0;
        if (t + dt > tend) { dt = tend - t; clip[n] = 1; }
        dts[n] = dt;
        double *un = (double *)malloc(sizeof(double) * N);
        memcpy(un, u, sizeof(double) * N);
        for (unsigned stage = 1; stage <= 3; stage++) {
            double *in = (double *)malloc(sizeof(double) * N);
            memcpy(in, u, sizeof(double) * N);
            st[3 * n + stage - 1] = in;
            s.alpha = stage == 1 ? a1 : orc_maxabs(N, u);
            al[3 * n + stage - 1] = s.alpha;
            orc_rhs_periodic(nx, ny, nz, dx, dy, dz, u, Lu, &s);
            orc_rk_stage_all(N, stage, dt, u, un, Lu);
        }
        free(un);
        t = t + dt;
        n++;
    }
    double J = 0;
    double *ub = (double *)malloc(sizeof(double) * N), *ub2 = (double *)malloc(sizeof(double) * N);
    double *ub1 = (double *)malloc(sizeof(double) * N), *acc = (double *)malloc(sizeof(double) * N);
    double *lam = (double *)malloc(sizeof(double) * N);
    for (size_t q = 0; q < N; q++) {
        const double e = u[q] - target[q];
        J += 0.5 * e * e;
        ub[q] = e;
    }
    double carry = 0;
    for (int m = n - 1; m >= 0; m--) {
        const double dt = dts[m];
        double dtb = 0, L, La;
        /* stage 3: u3 = (un + 2 (u2 + dt L(u2)))/3 */
        for (size_t q = 0; q < N; q++) { lam[q] = (2.0 / 3.0) * dt * ub[q]; ub2[q] = (2.0 / 3.0) * ub[q]; acc[q] = ub[q] / 3.0; }
        s.alpha = al[3 * m + 2];
        orc_rhs_vjp(nx, ny, nz, dx, dy, dz, st[3 * m + 2], lam, ub2, &s, &L, &La);
        dtb += L / dt;
        { size_t im = argmax_abs(N, st[3 * m + 2]); ub2[im] += La * (st[3 * m + 2][im] >= 0 ? 1.0 : -1.0); }
        /* stage 2: u2 = 0.75 un + 0.25 (u1 + dt L(u1)) */
        for (size_t q = 0; q < N; q++) { lam[q] = 0.25 * dt * ub2[q]; ub1[q] = 0.25 * ub2[q]; acc[q] += 0.75 * ub2[q]; }
        s.alpha = al[3 * m + 1];
        orc_rhs_vjp(nx, ny, nz, dx, dy, dz, st[3 * m + 1], lam, ub1, &s, &L, &La);
        dtb += L / dt;
        { size_t im = argmax_abs(N, st[3 * m + 1]); ub1[im] += La * (st[3 * m + 1][im] >= 0 ? 1.0 : -1.0); }
        /* stage 1: u1 = un + dt L(un) */
        for (size_t q = 0; q < N; q++) { lam[q] = dt * ub1[q]; ub[q] = acc[q] + ub1[q]; }
        s.alpha = al[3 * m];
        orc_rhs_vjp(nx, ny, nz, dx, dy, dz, st[3 * m], lam, ub, &s, &L, &La);
        dtb += L / dt;
        double a1b = La;
        if (clip[m]) {
            carry = dtb;  /* dt_last = tend - sum of earlier dt */
        } else {
            a1b += -(dtb - carry) * dt / al[3 * m];
        }
        { size_t im = argmax_abs(N, st[3 * m]); ub[im] += a1b * (st[3 * m][im] >= 0 ? 1.0 : -1.0); }
    }
    if (grad) memcpy(grad, ub, sizeof(double) * N);
    for (int q = 0; q < 3 * n; q++) free(st[q]);
    free(st); free(dts); free(al); free(clip); free(u); free(Lu);
    free(ub); free(ub2); free(ub1); free(acc); free(lam);
    if (nsteps) *nsteps = n;
    return J;
}
