/*
 * oracle.h -- CPU restatement of the reference stencil routines.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, always as the checker and never as the thing
 * measured or shipped.
 *
 * Every routine keeps the reference's floating-point expression tree
 * (operand order and association as written in the CUDA source after macro
 * expansion) and is compiled with -ffp-contract=off, so that each result is
 * the IEEE-754 binary64 evaluation of that source expression.
 *
 * Reference files are cited relative to the reference repository root:
 *   B5  = SingleGPU/Burgers3d_WENO5
 *   B5S = SingleGPU/Burgers3d_WENO5_SharedMem
 *   D3  = SingleGPU/Diffusion3d_baselineCode
 *   MB3 = MultiGPU/Burgers3d_Baseline
 *   MD3 = MultiGPU/Diffusion3d_Baseline
 *   MLB = Matlab_Prototipes/InviscidBurgersNd
 *   MLD = Matlab_Prototipes/DiffusionNd
 *
 * Pinning (which independent check covers each function) is recorded in the
 * header comment of oracle.c.
 */
#ifndef MGAD_ORACLE_H
#define MGAD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* WENO variants */
#define ORC_WENO_Z      0  /* B5/kernels.cu:142-192, tau5 = |B0-B2|, alpha = d(1+(tau5/(B+eps))^p) */
#define ORC_WENO_JS     1  /* MB3/Kernels.cu:112-220 (Reconstruct1d), w = alpha/sum, then /6 */
#define ORC_WENO_JS_B5  2  /* B5/kernels.cu:92-140 (WENO5reconstruction), /(6*sum) */
#define ORC_WENO_Z_B5S  3  /* B5S/kernels.cu:153-207, 1./sum then *(1/6 * inv) */

/* Lax-Friedrichs splitting */
#define ORC_LF_LOCAL    0  /* alpha = |u_i|   (B5/kernels.cu:224-238) */
#define ORC_LF_GLOBAL   1  /* alpha = max|u|  (MLB/WENO5resAdv1d.m:58-60) */

typedef struct {
    int    weno;   /* ORC_WENO_* */
    int    p;      /* WENO-Z power (1 in the reference, 2 in Borges et al. 2008) */
    double eps;    /* regularisation epsilon */
    int    lf;     /* ORC_LF_* */
    double alpha;  /* global LF alpha (used when lf == ORC_LF_GLOBAL) */
} orc_scheme;

/* ---- scalar building blocks ---- */
double orc_flux(double u);
void   orc_split(double u, double a, int lf, double *fp, double *fm);
double orc_face(const double *v, const double *w, const orc_scheme *s);
double orc_weno_plus(const double *v, const orc_scheme *s);
double orc_weno_minus(const double *w, const orc_scheme *s);

/* ---- reference drop-in sweeps (frozen 3-cell ghosts, x-fastest layout) ---- */
void orc_adv_x(unsigned nx, unsigned ny, unsigned nz, double dx,
               const double *u, double *Lu, const orc_scheme *s);
void orc_adv_y(unsigned nx, unsigned ny, unsigned nz, double dy,
               const double *u, double *Lu, const orc_scheme *s);
void orc_adv_z(unsigned nx, unsigned ny, unsigned nz, double dz,
               const double *u, double *Lu, const orc_scheme *s);
void orc_diff13_b5(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                   double kz, const double *u, double *Lu);
void orc_ssprk_b5(unsigned nx, unsigned ny, unsigned nz, unsigned step,
                  double dt, double *u, const double *uo, const double *Lu);
void orc_lu3d_13(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                 double kz, const double *u, double *Lu);
void orc_lu3d_o2(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                 double kz, const double *u, double *Lu);
void orc_rk3d(unsigned step, unsigned nx, unsigned ny, unsigned nz, double dt,
              double *u, const double *uo, const double *Lu);

/* ---- MultiGPU/Burgers3d_Baseline: one rank's slab, row pitch `pitch` ---- */
void orc_mb3_adv(int axis, unsigned pitch, unsigned nx, unsigned ny, unsigned nz,
                 unsigned kstart, unsigned kstop, double d, const double *u,
                 double *Lu, const orc_scheme *s);
void orc_mb3_rk(unsigned step, unsigned pitch, unsigned nx, unsigned ny,
                unsigned nz, double dt, double *q, const double *qo,
                const double *Lq);

/* ---- WENO7 Burgers prototype (Matlab_Prototipes/InviscidBurgersNd) ---- */
void orc_weno7_face(const double *V, const double *U, double *hn, double *hp);
void orc_weno7_rhs(unsigned nx, unsigned ny, unsigned nz, double dx, double dy,
                   double dz, const double *u, double *res);
extern const double orc_lsrk_a[5], orc_lsrk_b[5];
void orc_lsrk_stage(uint64_t n, int s, double dt, double *u, double *res,
                    const double *L);
int orc_weno7_run(unsigned nx, unsigned ny, unsigned nz, double dx, double dy,
                  double dz, double cfl, double *t, double tend, int nsteps,
                  double *u, double *res);

/* ---- periodic global-LF Burgers (the fused path's specification) ---- */
double orc_maxabs(uint64_t n, const double *u);
void   orc_rhs_periodic(unsigned nx, unsigned ny, unsigned nz, double dx,
                        double dy, double dz, const double *u, double *Lu,
                        const orc_scheme *s);
void   orc_rk_stage_all(uint64_t n, unsigned stage, double dt, double *u,
                        const double *uo, const double *Lu);
/* One SSP-RK3 step; returns dt used.  work must hold 2*nx*ny*nz doubles. */
double orc_burgers_step(unsigned nx, unsigned ny, unsigned nz, double dx,
                        double dy, double dz, double cfl, double t,
                        double tend, const orc_scheme *s, double *u,
                        double *work);
/* Steps until t >= tend or nsteps steps are done; returns steps taken. */
int    orc_burgers_run(unsigned nx, unsigned ny, unsigned nz, double dx,
                       double dy, double dz, double cfl, double *t,
                       double tend, int nsteps, const orc_scheme *s,
                       double *u);

/* ---- blocked / shared-memory / MPI driver variants ---- */
void orc_lu3d_blk(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                  double kz, const double *u, double *Lu);
void orc_adv_b5s(int axis, unsigned nx, unsigned ny, unsigned nz, double d,
                 const double *u, double *Lu, const orc_scheme *s);
void orc_md3_lap(int order, unsigned px, unsigned nx, unsigned ny, unsigned nz,
                 unsigned kstart, unsigned kstop, double dxc, double dyc,
                 double dzc, const double *u, double *Lu);
void orc_md3_rk(unsigned step, unsigned px, unsigned nx, unsigned ny, unsigned nz,
                double dt, double *u, const double *uo, const double *Lu);

/* ---- heat ---- */
void orc_heat7_fe(unsigned nx, unsigned ny, unsigned nz, double kx, double ky,
                  double kz, double dt, int nsteps, double *u);

/* ---- 1-D periodic WENO derivative (order-of-accuracy checks) ---- */
void orc_rhs_1d_periodic(unsigned n, double dx, const double *u, double *Lu,
                         const orc_scheme *s);

/* ---- discrete adjoint of the periodic global-LF Burgers step ---- */
/* Value and gradient of one side of the reconstruction w.r.t. its 5 inputs
 * (plus: f+ side, weights (1,6,3)/10; else the f- side). */
double orc_side_grad(const double *v, int plus, const orc_scheme *s, double *g);
/* VJP of L(u; alpha) (orc_rhs_periodic) applied to lam, ADDED into ubar;
 * returns <lam, L(u)> in *lam_dot_L and <lam, dL/dalpha> in *lam_dot_La. */
void orc_rhs_vjp(unsigned nx, unsigned ny, unsigned nz, double dx, double dy,
                 double dz, const double *u, const double *lam, double *ubar,
                 const orc_scheme *s, double *lam_dot_L, double *lam_dot_La);
/* J = 1/2 |u(T) - target|^2 after a run from t0 to tend (CFL rule,
 * global alpha per stage); returns J, writes dJ/du0 into grad (may be NULL)
 * and the number of steps into *nsteps. */
double orc_burgers_adjoint(unsigned nx, unsigned ny, unsigned nz, double dx,
                           double dy, double dz, double cfl, double t0,
                           double tend, const orc_scheme *s, const double *u0,
                           const double *target, double *grad, int *nsteps);

int  orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
