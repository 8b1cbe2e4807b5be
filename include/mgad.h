/*
 * mgad.h -- C ABI of the MI355X (gfx950) advection-diffusion stencil library.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in replacements for the reference's "device wrappers", same names,
 *     same argument lists and meaning (REAL = double, cudaStream_t ->
 *     hipStream_t, CUDA dim3 -> HIP dim3, both three 32-bit unsigned ints).
 *     They return void like the reference; failures are reported through the
 *     sticky code of mgad_get_last_error() instead of printf + exit(-1).
 *     The launch-geometry arguments are accepted and ignored: the library
 *     always covers the whole update region with its own launch shape.
 *
 *  2. mgad_* extensions for the periodic global-Lax-Friedrichs Burgers path
 *     (fused SSP-RK3 stages with an on-device time step), the 7-point heat
 *     forward-Euler step, and host helpers.  They return an mgad_status.
 *
 * All device pointers are caller-owned (allocated with hipMalloc or any
 * allocator of the same device); no entry point allocates device memory.
 * Arrays are x-fastest: index o = i + nx*j + nx*ny*k.
 */
#ifndef MGAD_H
#define MGAD_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes                                                              */
/* ------------------------------------------------------------------------ */
typedef enum {
    MGAD_OK = 0,
    MGAD_ERR_INVALID_ARG = 1,   /* bad size, null pointer, bad stage number */
    MGAD_ERR_UNSUPPORTED = 2,   /* grid shape not supported by the kernel set */
    MGAD_ERR_HIP = 3,           /* a HIP runtime call or a launch failed */
    MGAD_ERR_NONFINITE = 4,     /* max|u| read back as NaN or Inf */
    MGAD_ERR_SCRATCH_FULL = 5   /* input needs more than the library's static
                                   device scratch (mgad_batch_init_static) */
} mgad_status;

/* Returns the first error recorded since the previous call and clears it. */
int mgad_get_last_error(void);
const char *mgad_error_string(int code);
/* Library version string. */
const char *mgad_version(void);

/* ------------------------------------------------------------------------ */
/* 1. Reference drop-in entry points                                         */
/* ------------------------------------------------------------------------ */

/* Replaces CopyToConstantMemory, SingleGPU/Burgers3d_WENO5/Burgers.h:72 and
 * SingleGPU/Diffusion3d_baselineCode/acoustics3d.h:57 (same signature in
 * both): diffusion coefficients kx = C/(12 dx^2) used by Call_Diff_ and
 * Call_Lu3d.  Process-global, set once before launches. */
void CopyToConstantMemory(double kx, double ky, double kz);

/* Replaces Call_Adv_x, SingleGPU/Burgers3d_WENO5/Burgers.h:73-74
 * (kernel Compute_dF, kernels.cu:197-260): Lu = -(F_{i+1/2}-F_{i-1/2})/dx on
 * rows 2<j<Ny-3, 2<k<Nz-3 for i = 2..Nx-4 (i = 2 with F_{3/2} := 0). */
void Call_Adv_x(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                unsigned int Nx, unsigned int Ny, unsigned int Nz, double dx,
                double *u, double *Lu);
/* Replaces Call_Adv_y, Burgers.h:75-76 (Compute_dG, kernels.cu:265-327):
 * Lu -= (G_{j+1/2}-G_{j-1/2})/dy on 2<i<Nx-3, 2<k<Nz-3, j = 2..Ny-4. */
void Call_Adv_y(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                unsigned int Nx, unsigned int Ny, unsigned int Nz, double dy,
                double *v, double *Lv);
/* Replaces Call_Adv_z, Burgers.h:77-78 (Compute_dH, kernels.cu:332-394):
 * Lu -= (H_{k+1/2}-H_{k-1/2})/dz on 2<i<Nx-3, 2<j<Ny-3, k = 2..Nz-4. */
void Call_Adv_z(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                unsigned int Nx, unsigned int Ny, unsigned int Nz, double dz,
                double *w, double *Lw);
/* Replaces Call_Diff_, Burgers.h:79-80 (Compute_Laplace_Async,
 * kernels.cu:444-494): Lq += 13-point 4th-order Laplacian with the
 * CopyToConstantMemory coefficients on 2<i<Nx-3, 2<j<Ny-3, k = 3..Nz-3. */
void Call_Diff_(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                unsigned int Nx, unsigned int Ny, unsigned int Nz, double *q,
                double *Lq);
/* Replaces Call_sspRK, Burgers.h:81-82 (Compute_RK, kernels.cu:499-530):
 * SSP-RK3 stage `step` (1..3) on 2<i<Nx-3, 2<j<Ny-3, 2<k<Nz-3. */
void Call_sspRK(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                unsigned int Nx, unsigned int Ny, unsigned int Nz,
                unsigned int step, double dt, double *u, double *uo,
                double *Lu);
/* Replaces Call_Lu3d, SingleGPU/Diffusion3d_baselineCode/acoustics3d.h:58
 * (Compute_Laplace3d, kernels.cu:66-99): Lu = 13-point Laplacian on
 * 1<i<nx-2, 1<j<ny-2, 1<k<nz-2. */
void Call_Lu3d(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
               unsigned int nx, unsigned int ny, unsigned int nz, double *u,
               double *Lu);
/* Replaces Call_RK3d, acoustics3d.h:59 (Compute_RK, kernels.cu:104-142):
 * SSP-RK3 stage `step` on 1<i<nx-2, 1<j<ny-2, 1<k<nz-2. */
void Call_RK3d(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
               unsigned int step, unsigned int nx, unsigned int ny,
               unsigned int nz, double dt, double *u, double *uo, double *Lu);

/* ------------------------------------------------------------------------ */
/* Pitched drop-ins                                                          */
/* ------------------------------------------------------------------------ */
/* The reference's pitched-memory variants take the same names with an extra
 * `pitch` (row length in elements, >= Nx; cudaMallocPitch bytes / sizeof
 * REAL): node (i, j, k) is at i + pitch*j + pitch*Ny*k.  One library cannot
 * export two signatures under one name, so they carry a _pitched suffix.
 * Same update sets and arithmetic as the unpitched entry points above;
 * padding columns i >= Nx are neither read nor written.
 *   Burgers: SingleGPU/Burgers3d_WENO5_PitchedMem/Burgers.h:73-77; its
 *   Call_Diff_ and Call_sspRK are also those of Burgers3d_WENO5_SharedMem
 *   (Burgers.h:82-83, kernels.cu:570-581), whose Call_Adv_* differ (below).
 *   Heat: SingleGPU/Diffusion3d_PitchedMem/acoustics3d.h:58-59; its Call_RK3d
 *   is also Diffusion3d_Blocking's (acoustics3d.h:60, kernels.cu:193-198),
 *   whose Call_Lu3d differs (below). */
void Call_Adv_x_pitched(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                        unsigned int pitch, unsigned int Nx, unsigned int Ny,
                        unsigned int Nz, double dx, double *u, double *Lu);
void Call_Adv_y_pitched(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                        unsigned int pitch, unsigned int Nx, unsigned int Ny,
                        unsigned int Nz, double dy, double *v, double *Lv);
void Call_Adv_z_pitched(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                        unsigned int pitch, unsigned int Nx, unsigned int Ny,
                        unsigned int Nz, double dz, double *w, double *Lw);
void Call_Diff_pitched(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                       unsigned int pitch, unsigned int Nx, unsigned int Ny,
                       unsigned int Nz, double *q, double *Lq);
void Call_sspRK_pitched(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                        unsigned int pitch, unsigned int Nx, unsigned int Ny,
                        unsigned int Nz, unsigned int step, double dt, double *u,
                        double *uo, double *Lu);
void Call_Lu3d_pitched(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                       unsigned int pitch, unsigned int nx, unsigned int ny,
                       unsigned int nz, double *u, double *Lu);
void Call_RK3d_pitched(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                       unsigned int step, unsigned int pitch, unsigned int nx,
                       unsigned int ny, unsigned int nz, double dt, double *u,
                       double *uo, double *Lu);

/* ------------------------------------------------------------------------ */
/* Drivers whose wrappers differ from the base code's                         */
/* ------------------------------------------------------------------------ */
/* Replace Call_Adv_x/y/z of SingleGPU/Burgers3d_WENO5_SharedMem
 * (Burgers.h:79-81; kernels.cu:552-568 launching Compute_dF/dG/dH :212-402
 * with WENO5Zreconstruction :153-205), pitched arrays.  Unlike the base
 * sweeps they update every node of the array: the face below node i reads
 * nodes i-3..i+2 of its line with nodes outside the array taken as 0
 * (:248-260), and each WENO-Z side is scaled as *(1/6 * (1/sum)) instead of
 * /(6*sum).  x: Lu = -(F_{i+1/2}-F_{i-1/2})/dx; y, z: Lu -= (..)/d.  Local
 * LF split; p and eps from mgad_set_adv_scheme (defaults p = 1, eps = 1e-6:
 * the reference's Burgers.h:26).  Ghost nodes near the high end, where the
 * reference reads shared memory no thread wrote, get the zero-padded value
 * (the SSP-RK stage never reads them). */
void Call_Adv_x_sharedmem(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                          unsigned int pitch, unsigned int Nx, unsigned int Ny,
                          unsigned int Nz, double dx, double *u, double *Lu);
void Call_Adv_y_sharedmem(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                          unsigned int pitch, unsigned int Nx, unsigned int Ny,
                          unsigned int Nz, double dy, double *v, double *Lv);
void Call_Adv_z_sharedmem(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                          unsigned int pitch, unsigned int Nx, unsigned int Ny,
                          unsigned int Nz, double dz, double *w, double *Lw);
/* Replaces Call_Lu3d of SingleGPU/Diffusion3d_Blocking (acoustics3d.h:59;
 * kernels.cu:187-191 launching Compute_Laplace3d_Async :37-88 with
 * kstart = 3, kstop = nz-2): Lu = the 13-point operator, each axis summed
 * from its low side, on 2<i<nx-3, 2<j<ny-3 and planes 3..nz-3 only; other
 * nodes keep their Lu (the driver zeroes Lu once, main.cpp:101, so the
 * ring i = 2 stays at its initial value under Call_RK3d_pitched).  Planes
 * the reference also writes when nz % 16 is 1 or 2 (a block's unconditional
 * first plane, reading past the array) are not written. */
void Call_Lu3d_blocking(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                        unsigned int pitch, unsigned int nx, unsigned int ny,
                        unsigned int nz, double *u, double *Lu);
/* Replace Call_Diff_ and Call_sspRK of MultiGPU/Diffusion3d_Baseline
 * (DiffusionMPICUDA.h:118-123, Kernels.cu:401-414), whose signatures clash
 * with the single-GPU ones.  One rank's slab of _Nz planes (ghost planes
 * included), row pitch `pitch`:
 *   Call_Diff_mpi: LaplaceO4_async (Kernels.cu:207-260), the kernel the
 *     wrapper launches: Lu = diff_x/12 * (one left-to-right sum of the 12
 *     neighbours and -90 u) on 2<i<Nx-3, 2<j<Ny-3, planes kstart..kstop-1
 *     (2 <= kstart <= kstop <= _Nz-2; diff_y, diff_z unused, as there);
 *   Call_Diff_O2_mpi: LaplaceO2_async (Kernels.cu:152-201), the wrapper's
 *     commented-out alternative: Lu = diff_x (u_w - 2u + u_e) + diff_y (..)
 *     + diff_z (..), 1 <= kstart <= kstop <= _Nz-1;
 *   Call_sspRK_mpi: Compute_RK (Kernels.cu:266-297) on 2<i<Nx-3, 2<j<Ny-3
 *     and every plane of the slab.
 * The planes written are exactly kstart..kstop-1 (the reference's blocks
 * of 16 planes also write their first plane when it lies past kstop). */
void Call_Diff_mpi(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                   unsigned int pitch, unsigned int Nx, unsigned int Ny,
                   unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                   double diff_x, double diff_y, double diff_z, double *q,
                   double *Lq);
void Call_Diff_O2_mpi(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                      unsigned int pitch, unsigned int Nx, unsigned int Ny,
                      unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                      double diff_x, double diff_y, double diff_z, double *q,
                      double *Lq);
void Call_sspRK_mpi(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                    unsigned int step, unsigned int pitch, unsigned int Nx,
                    unsigned int Ny, unsigned int _Nz, double dt, double *q,
                    double *qo, double *Lq);

/* ------------------------------------------------------------------------ */
/* MultiGPU/Burgers3d_Baseline device wrappers: one rank's z slab            */
/* ------------------------------------------------------------------------ */
/* Same names, arguments and update sets as BurgersMPI.h:112-132 /
 * Kernels.cu:528-580 (launch geometry advisory, errors through
 * mgad_get_last_error).  A slab holds _Nz planes (3 ghost planes at each end)
 * of row pitch `pitch`.  The scheme is the one that code is built with:
 * Reconstruct1d (WENO5-JS, w = alpha/sum then /6), EPS 1e-6, local LF. */

/* Compute_dF / dG (Kernels.cu:225-371): Lu = (x) or Lu -= (y) on lines over
 * 3..N-4 and planes kstart..kstop-1, nodes 2..N-4 along the line (the face
 * below node 2 taken as 0). */
void Compute_Adv_x_Async(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                         unsigned int pitch, unsigned int Nx, unsigned int Ny,
                         unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                         double dx, double *d_s_u, double *d_s_Lu);
void Compute_Adv_y_Async(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                         unsigned int pitch, unsigned int Nx, unsigned int Ny,
                         unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                         double dy, double *d_s_u, double *d_s_Lu);
/* Compute_dH (Kernels.cu:376-452): Lu -= on columns 3..N-4 in x and y and
 * planes kstart..kstop-1 (kstart >= 3, kstop <= _Nz-3), both faces of every
 * plane computed. */
void Compute_Adv_z_Async(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                         unsigned int pitch, unsigned int Nx, unsigned int Ny,
                         unsigned int _Nz, unsigned int kstart, unsigned int kstop,
                         double dz, double *d_s_u, double *d_s_Lu);
/* Compute_RK (Kernels.cu:457-490): the SSP-RK3 stage on every node of the
 * slab, ghosts included. */
void Compute_sspRK3Async(dim3 numBlocks, dim3 threadsPerBlock, hipStream_t aStream,
                         unsigned int step, unsigned int pitch, unsigned int Nx,
                         unsigned int Ny, unsigned int _Nz, double dt, double *q,
                         double *qo, double *Lq);
/* copy_br_to_gc (Kernels.cu:40-72): planes _Nz-6.._Nz-4 (side 0) or 3..5
 * (side 1) into a 3-plane buffer of row pitch gc_pitch. */
void CopyBoundaryRegionToGhostCellAsync(dim3 numBlocks, dim3 threadsPerBlock,
                                        hipStream_t aStream, double *d_s_q,
                                        double *d_send_buffer, unsigned int Nx,
                                        unsigned int Ny, unsigned int _Nz,
                                        unsigned int pitch, unsigned int gc_pitch,
                                        unsigned int side);
/* copy_gc_to_br (Kernels.cu:77-107): a 3-plane buffer into planes
 * _Nz-3.._Nz-1 (side 0) or 0..2 (side 1). */
void CopyGhostCellToBoundaryRegionAsync(dim3 numBlocks, dim3 threadsPerBlock,
                                        hipStream_t aStream, double *d_s_q,
                                        double *d_recv_buffer, unsigned int Nx,
                                        unsigned int Ny, unsigned int _Nz,
                                        unsigned int pitch, unsigned int gc_pitch,
                                        unsigned int side);

/* ------------------------------------------------------------------------ */
/* Scheme selection for the Call_Adv_* sweeps                                */
/* ------------------------------------------------------------------------ */
#define MGAD_WENO_Z     0  /* SingleGPU/Burgers3d_WENO5/kernels.cu:142-192 */
#define MGAD_WENO_JS    1  /* MultiGPU/Burgers3d_Baseline/Kernels.cu:112-220 */
#define MGAD_LF_LOCAL   0  /* alpha = |u_i|, SingleGPU/Burgers3d_WENO5/kernels.cu:224-238 */
#define MGAD_LF_GLOBAL  1  /* alpha = given constant (max|u|) */

/* Extra parameters of the Call_Adv_* sweeps (the reference fixes them at
 * compile time, Burgers.h:20-28): WENO variant, WENO-Z power p (1 or 2),
 * epsilon, LF mode and the global alpha.  Defaults reproduce the reference:
 * WENO-Z, p = 1, eps = 1e-6, local LF.  Not thread-safe; set before launch. */
int mgad_set_adv_scheme(int weno, int p, double eps, int lf, double alpha);

/* Compute_Laplace3d_O2, SingleGPU/Diffusion3d_baselineCode/kernels.cu:37-64
 * (defined there, never launched): Lu = 7-point Laplacian
 * 12kx(ue-2uo+uw)+12ky(..)+12kz(..) on 0<i<nx-1 with the
 * CopyToConstantMemory coefficients. */
int mgad_lu3d_o2(hipStream_t stream, unsigned int nx, unsigned int ny,
                 unsigned int nz, const double *u, double *Lu);

/* ------------------------------------------------------------------------ */
/* 2. Fused periodic Burgers path                                            */
/* ------------------------------------------------------------------------ */

/* Device-resident step control.  The caller allocates MGAD_CTRL_BYTES of
 * device memory and passes it to every call of one run.  Fields (read with
 * mgad_burgers_read_ctrl): current time t, dt of the last step, step count,
 * and max|u| of the stage inputs for the global LF split and the CFL rule. */
#define MGAD_CTRL_BYTES 256

typedef struct {
    double t;        /* time at the start of the next step */
    double dt;       /* dt of the step in flight / last step (0 once done) */
    double tend;     /* final time */
    double cfl;      /* CFL number: dt = cfl*dx/max|u^n| */
    double amax;     /* max|u^n| of the current solution */
    uint64_t steps;  /* steps taken */
} mgad_burgers_state;

/* Scheme of the fused path. */
typedef struct {
    int weno;      /* MGAD_WENO_Z or MGAD_WENO_JS */
    int p;         /* WENO-Z power: 1 or 2 */
    double eps;    /* epsilon (JS 1e-6; Z p=2 uses 1e-40) */
} mgad_scheme;

/* Initialise ctrl for a run starting at time t0 from u (device, nx*ny*nz):
 * computes max|u| on the device.  Asynchronous on stream. */
int mgad_burgers_init(hipStream_t stream, unsigned int nx, unsigned int ny,
                      unsigned int nz, const double *u, void *ctrl, double t0,
                      double tend, double cfl);

/* One SSP-RK3 stage on a periodic grid with global LF alpha = max|u_in|:
 *   stage 1: u_out = u_n + dt*L(u_n)             (u_in == u_n)
 *   stage 2: u_out = 0.75 u_n + 0.25 (u_in + dt*L(u_in))
 *   stage 3: u_out = (u_n + 2 (u_in + dt*L(u_in)))/3
 * with L(u) = -dF/dx - dG/dy - dH/dz (WENO5 face fluxes, the order of
 * Call_Adv_x -> Call_Adv_y -> Call_Adv_z).  Stage 1 sets dt =
 * cfl*dx/max|u_n| clipped to tend; the step is a no-op once t >= tend, and
 * the run stops (dt = 0; t, the step count and u as after the last completed
 * step) when a stage input's max|u| is NaN or infinite, which
 * mgad_burgers_read_ctrl then reports as MGAD_ERR_NONFINITE.  Any nx, ny, nz >= 4 with fewer than 2^31 points.
 * u_out may alias u_n (stage 3 runs in place), never u_in. */
int mgad_burgers_stage(hipStream_t stream, const mgad_scheme *scheme,
                       unsigned int nx, unsigned int ny, unsigned int nz,
                       double dx, double dy, double dz, int stage,
                       const double *u_in, const double *u_n, double *u_out,
                       void *ctrl);

/* nsteps full SSP-RK3 steps: u (in/out) plus two scratch fields w1, w2 of
 * nx*ny*nz doubles.  use_graph != 0 captures one step in a hipGraph and
 * replays it.  Asynchronous on stream. */
int mgad_burgers_steps(hipStream_t stream, const mgad_scheme *scheme,
                       unsigned int nx, unsigned int ny, unsigned int nz,
                       double dx, double dy, double dz, double *u, double *w1,
                       double *w2, void *ctrl, int nsteps, int use_graph);

/* Storage precision of the fused path's fields. */
#define MGAD_F64 0  /* double storage, double arithmetic */
#define MGAD_F32 1  /* float storage; split fluxes, smoothness indicators,
                       weights and the RK combine in double, one rounding to
                       float per stored value */

/* mgad_burgers_init / mgad_burgers_steps with the storage precision chosen
 * by `precision`; u, w1, w2 are float* for MGAD_F32, double* for MGAD_F64. */
int mgad_burgers_init_p(int precision, hipStream_t stream, unsigned int nx,
                        unsigned int ny, unsigned int nz, const void *u,
                        void *ctrl, double t0, double tend, double cfl);
int mgad_burgers_steps_p(int precision, hipStream_t stream,
                         const mgad_scheme *scheme, unsigned int nx,
                         unsigned int ny, unsigned int nz, double dx,
                         double dy, double dz, void *u, void *w1, void *w2,
                         void *ctrl, int nsteps, int use_graph);

/* kchunk of a stage launch whose z bands shrink towards the top of the grid
 * (big ones first, 4- or 8-plane ones last). */
#define MGAD_BANDS_SHRINK (-1)

/* Autotuner: times stage-1 launches (u -> w1) of the default launch and of
 * every candidate tile width (16, 32) and z band rule on the caller's buffers
 * and records the fastest for this (precision, scheme, grid); later launches
 * of that grid use it.  A stage launch runs one workgroup per column tile and
 * band of z planes.  The default launch (tile 0, kchunk 0) uses 32x32 tiles,
 * one workgroup per CU, when one round of them covers the grid (bands of >= 4
 * planes filling >= 90 % of the CUs; where 32 does not divide nx or ny, the
 * tiles may cover at most 7 % more columns than the grid), and 32x16 tiles
 * otherwise; kchunk > 0 makes the bands kchunk planes deep
 * (4..128 and nz are tried), MGAD_BANDS_SHRINK makes them shrink towards the
 * top of the grid.  Overwrites w1 and ctrl: call mgad_burgers_init[_p]
 * afterwards.  Results do not depend on the configuration (bit-identical). */
int mgad_burgers_autotune(hipStream_t stream, int precision,
                          const mgad_scheme *scheme, unsigned int nx,
                          unsigned int ny, unsigned int nz, double dx,
                          double dy, double dz, const void *u, void *w1,
                          void *ctrl, int *tile, int *kchunk);
/* Sets the launch configuration of a grid explicitly (tile 0 / kchunk 0 =
 * the built-in default; kchunk > 0 = bands of that depth, MGAD_BANDS_SHRINK
 * = shrinking bands). */
int mgad_burgers_set_launch(int precision, const mgad_scheme *scheme,
                            unsigned int nx, unsigned int ny, unsigned int nz,
                            int tile, int kchunk);

/* ------------------------------------------------------------------------ */
/* Batched periodic Burgers: many independent problems, one launch per stage */
/* ------------------------------------------------------------------------ */

/* One problem of a batch: its three fields (precision of the batch), its own
 * MGAD_CTRL_BYTES control buffer, extents (multiples of 16 in x and y) and
 * spacings. */
typedef struct {
    void *u, *w1, *w2;
    void *ctrl;
    unsigned int nx, ny, nz;
    double dx, dy, dz;
} mgad_problem;

/* Filled by mgad_batch_init; passed to the step calls. */
typedef struct {
    void *workspace;        /* device buffer holding descriptors + work list */
    size_t workspace_bytes;
    int nprob, nitems;      /* problems, (problem, tile) work items */
    int grid;               /* persistent blocks per stage launch */
    int kchunk;             /* planes per work item, or MGAD_BATCH_SPLIT */
    int precision;          /* MGAD_F64 or MGAD_F32 */
    mgad_scheme scheme;
} mgad_batch;

/* kchunk of the split work lists: one item per column tile (all its planes),
 * the items' planes cut into shares that shrink towards the end of the list,
 * which the resident blocks draw from a counter in the workspace.  Tiles are
 * 16x16 columns, or 32x16 for the wide list (where nx % 32 == 16, the last
 * tile of a row reaches past nx and stores only its first 16 columns). */
#define MGAD_BATCH_SPLIT      (-1)
#define MGAD_BATCH_SPLIT_WIDE (-2)

/* Device bytes the workspace of this batch needs (upper bound over the
 * work lists the autotuner tries). */
size_t mgad_batch_workspace_bytes(int nprob, const mgad_problem *probs);

/* Builds the work list in the caller's workspace and initialises every
 * problem's ctrl at t0 (u of each problem must hold its initial field).
 * kchunk > 0 fixes the planes per work item, MGAD_BATCH_SPLIT(_WIDE) picks
 * a split list, and 0 autotunes: it times the batched stages with both split
 * lists and with 8, 16 and 32 planes per item and keeps the fastest.
 * Synchronises the stream. */
int mgad_batch_init(hipStream_t stream, int precision,
                    const mgad_scheme *scheme, int nprob,
                    const mgad_problem *probs, void *workspace,
                    size_t workspace_bytes, int kchunk, double t0,
                    double tend, double cfl, mgad_batch *out);

/* nsteps SSP-RK3 steps of every problem: three persistent launches per step
 * (each walks the whole work list); use_graph captures one step in a
 * hipGraph and replays it.  Each problem keeps its own dt and LF alpha, and
 * its results are bit-identical to advancing it alone. */
int mgad_batch_steps(hipStream_t stream, const mgad_batch *batch, int nsteps,
                     int use_graph);

/* The same steps with the single-problem kernel launched problem by problem
 * (3 launches per problem per step, each with its tuned configuration). */
int mgad_batch_steps_serial(hipStream_t stream, const mgad_batch *batch,
                            const mgad_problem *probs, int nsteps,
                            int use_graph);

/* Size of the library's static device scratch that holds the work list of
 * mgad_batch_init_static (descriptors, items, shares): 1 MiB per GPU, held
 * from library load until the process ends. */
#define MGAD_BATCH_SCRATCH_BYTES (1u << 20)

/* No counterpart in the reference.  mgad_batch_init for callers that cannot
 * allocate device memory (code under HIP graph capture set-up, serving loops,
 * harnesses that forbid hipMalloc): the same work list, written into the
 * library's static scratch of the current device instead of a caller
 * workspace, and no device allocation (no hipMalloc, hipFree, events or
 * graphs).  kchunk must name the list (MGAD_BATCH_SPLIT, MGAD_BATCH_SPLIT_WIDE
 * or planes per item > 0): 0 (autotune) is rejected, since timing needs HIP
 * events.  A work list above MGAD_BATCH_SCRATCH_BYTES returns
 * MGAD_ERR_SCRATCH_FULL.  out->workspace points into the scratch, so
 * mgad_batch_steps(use_graph = 0) then runs without allocating either.  One
 * scratch per GPU: the next mgad_batch_init_static on the same device
 * replaces the work list, so a batch from it is valid until then, and calls
 * that overlap on one device (other host threads, or steps of an earlier
 * static batch still queued on another stream) race.  Synchronises the
 * stream. */
int mgad_batch_init_static(hipStream_t stream, int precision,
                           const mgad_scheme *scheme, int nprob,
                           const mgad_problem *probs, int kchunk, double t0,
                           double tend, double cfl, mgad_batch *out);

/* ------------------------------------------------------------------------ */
/* z-slab decomposition of the periodic Burgers run (FP64)                   */
/* ------------------------------------------------------------------------ */

/* The layout and halo exchange of the z-slab decomposition of
 * MultiGPU/Burgers3d_Baseline (main.c:64-66, 3 ghost layers BurgersMPI.h:50),
 * with the fused stage kernel on each slab.  A slab set holds nslab slabs of
 * nzs planes of a periodic nx x ny x (nslab*nzs) grid one after another;
 * slab s stores nzs + 6 planes: 3 ghost planes, its own planes
 * s*nzs .. s*nzs+nzs-1, 3 ghost planes (nx*ny*(nzs+6) doubles per slab).
 * Requires nzs >= 3 and nx*ny*(nzs+6) < 2^31. */

/* Ghost planes of every slab from its neighbours' boundary planes (slab
 * indices periodic; one slab wraps onto itself).  Each side is 3 contiguous
 * planes, the data one GPU of the decomposition would send to and receive
 * from each neighbour per stage.  Replaces the CopyBoundaryRegionToGhostCell
 * + MPI + CopyGhostCellToBoundaryRegion sequence of MultiGPU/
 * Burgers3d_Baseline/main.c:207-294.  Asynchronous on stream. */
int mgad_slab_exchange(hipStream_t stream, unsigned int nx, unsigned int ny,
                       unsigned int nzs, unsigned int nslab, double *u);

/* mgad_burgers_init for a slab set: exchanges u's ghosts, then max|u| over
 * the own planes of all slabs.  ctrl: MGAD_CTRL_BYTES of device memory. */
int mgad_burgers_slab_init(hipStream_t stream, unsigned int nx,
                           unsigned int ny, unsigned int nzs,
                           unsigned int nslab, double *u, void *ctrl,
                           double t0, double tend, double cfl);

/* nsteps SSP-RK3 steps of the periodic run on a slab set: per stage one
 * launch over all slabs (one control block: max|u| and dt are global) and
 * one exchange of the stage's output.  u, w1, w2: slab sets of the same
 * shape; nx, ny multiples of 16.  The own planes equal mgad_burgers_steps
 * on the unsplit grid bit for bit.  Asynchronous on stream. */
int mgad_burgers_slab_steps(hipStream_t stream, const mgad_scheme *scheme,
                            unsigned int nx, unsigned int ny,
                            unsigned int nzs, unsigned int nslab, double dx,
                            double dy, double dz, double *u, double *w1,
                            double *w2, void *ctrl, int nsteps,
                            int use_graph);

/* ------------------------------------------------------------------------ */
/* WENO7 Burgers (Matlab_Prototipes/InviscidBurgersNd/LFWENO7FDM3d.m)        */
/* ------------------------------------------------------------------------ */
/* u_t + (u^2/2)_x + (u^2/2)_y + (u^2/2)_z = 0 on a node grid: local LF split,
 * WENO7 faces (Balsara-Shu smoothness forms, eps 1e-6, WENO7resAdv_X.m), ends
 * extended by repeating the boundary value, the face below node 0 of each
 * line taken as 0, and the low-storage 5-stage 4th-order Runge-Kutta with
 * dt = cfl*dx/max|u| clipped to tend (LFWENO7FDM3d.m:66-108).  Fields are
 * nx*ny*nz doubles, x fastest; nx*ny*nz < 2^31. */

/* Starts a run at t0: ctrl (MGAD_CTRL_BYTES) from max|u|, res zeroed. */
int mgad_weno7_init(hipStream_t stream, unsigned int nx, unsigned int ny,
                    unsigned int nz, const double *u, double *res, void *ctrl,
                    double t0, double tend, double cfl);

/* The right-hand side L = dF + dG + dH of one field (res = L), as the
 * prototype's WENO7resAdv_X + _Y + _Z. */
int mgad_weno7_rhs(hipStream_t stream, unsigned int nx, unsigned int ny,
                   unsigned int nz, double dx, double dy, double dz,
                   const double *u, double *L);

/* nsteps steps, five stage launches each: u (in/out), scratch fields w1, w2
 * and the carried residual res; steps after tend are no-ops on the device.
 * use_graph captures one step in a hipGraph and replays it. */
int mgad_weno7_steps(hipStream_t stream, unsigned int nx, unsigned int ny,
                     unsigned int nz, double dx, double dy, double dz,
                     double *u, double *w1, double *w2, double *res,
                     void *ctrl, int nsteps, int use_graph);

/* t, dt of the last step and the step count (synchronises the stream). */
int mgad_weno7_read_ctrl(hipStream_t stream, const void *ctrl, double *t,
                         double *dt, unsigned long long *steps);

/* ------------------------------------------------------------------------ */
/* Discrete adjoint of the periodic Burgers run (FP64)                       */
/* ------------------------------------------------------------------------ */

/* Device bytes mgad_burgers_adjoint needs for a grid, a bound on the number
 * of steps and a checkpoint interval (sqrt(max_steps) keeps both the
 * checkpoints and one recomputed segment at about sqrt(N) fields). */
size_t mgad_adjoint_workspace_bytes(unsigned int nx, unsigned int ny,
                                    unsigned int nz, int max_steps,
                                    int checkpoint_every);

/* J = 1/2 |u(tend) - target|^2 for the run of mgad_burgers_steps from u0 at
 * t0 (CFL rule, global LF alpha per stage) and its gradient dJ/du0 by the
 * discrete adjoint: the forward pass stores u every checkpoint_every steps;
 * the reverse sweep recomputes each segment from its checkpoint (keeping the
 * stage inputs) and applies the stage VJPs in reverse.  The alpha = max|u|
 * dependence is differentiated at the first node attaining the maximum; dt
 * = cfl*dx/max|u^n| and the clipped last step dt = tend - t are included.
 * u0, target, grad: device, nx*ny*nz doubles; nx, ny multiples of 16.
 * Synchronises the stream; returns MGAD_ERR_INVALID_ARG if the run needs
 * more than max_steps steps. */
int mgad_burgers_adjoint(hipStream_t stream, const mgad_scheme *scheme,
                         unsigned int nx, unsigned int ny, unsigned int nz,
                         double dx, double dy, double dz, double cfl,
                         double t0, double tend, const double *u0,
                         const double *target, double *grad, double *J,
                         int max_steps, int checkpoint_every,
                         void *workspace, size_t workspace_bytes,
                         int *nsteps);

/* Copies ctrl to the host (synchronises the stream). */
int mgad_burgers_read_ctrl(hipStream_t stream, const void *ctrl,
                           mgad_burgers_state *out);

/* ------------------------------------------------------------------------ */
/* The reference's own Burgers step, fused                                   */
/* ------------------------------------------------------------------------ */

/* nsteps SSP-RK3 steps of the reference driver's loop, one launch per stage:
 * the same u as, per step, copying u to uo and running for rk = 1..3
 *   Call_Adv_x, Call_Adv_y, Call_Adv_z, Call_Diff_ (if viscous),
 *   Call_sspRK(rk, dt, u, uo, Lu)
 * (SingleGPU/Burgers3d_WENO5/main.cpp:139-149) with
 * mgad_set_adv_scheme(scheme->weno, scheme->p, scheme->eps, MGAD_LF_LOCAL)
 * and the CopyToConstantMemory coefficients, bit for bit.  Interior nodes
 * 3..n-4 of every axis are updated, the three ghost layers are left as they
 * are; dt is the caller's fixed step.  w1 and w2 are scratch fields of the
 * same size (each call first copies u into both, for the ghost layers).
 * Requires n >= 7 per axis.  No Lu array is produced. */
int mgad_burgers_ref_steps(hipStream_t stream, const mgad_scheme *scheme,
                           int viscous, unsigned int nx, unsigned int ny,
                           unsigned int nz, double dx, double dy, double dz,
                           double dt, double *u, double *w1, double *w2,
                           int nsteps, int use_graph);

/* ------------------------------------------------------------------------ */
/* Heat: 7-point forward Euler                                               */
/* ------------------------------------------------------------------------ */

/* nsteps forward-Euler steps u <- u + dt*L7(u) on the Dirichlet interior
 * 0<i<n-1 (boundary nodes held), L7 as mgad_lu3d_o2 with coefficients
 * kx, ky, kz (= C/(12 dx^2)).  u and w are two device fields; the result is
 * left in u.  Requires n >= 3 per axis, nx*ny*nz < 2^31 and nx*ny < 2^29.
 * Asynchronous on stream. */
int mgad_heat7_fe_steps(hipStream_t stream, unsigned int nx, unsigned int ny,
                        unsigned int nz, double kx, double ky, double kz,
                        double dt, double *u, double *w, int nsteps,
                        int use_graph);

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                               */
/* ------------------------------------------------------------------------ */

/* No counterpart in the reference.  Evaluates on the device, for n operand
 * pairs (a[i], b[i]), the shortened division the stage kernels use for the
 * WENO weight divisions when max|u| and eps allow it, and the IEEE division
 * '/' compiles to: mode 0 gives a/b, mode 1 and 2 the WENO-Z weight
 * 0.3*(1 + (a/b)^p) with p = mode.  q_fast and q_ieee receive the two
 * results, for a bitwise comparison by the caller. */
int mgad_check_division(hipStream_t stream, int mode, const double *a,
                        const double *b, double *q_fast, double *q_ieee,
                        size_t n);

/* The reference's heat step fused (SingleGPU/Diffusion3d_baselineCode/
 * main.cpp:108-121): the same u as nsteps times {uo <- u; for rk = 1..3:
 * Call_Lu3d, Call_RK3d(rk)} with the CopyToConstantMemory coefficients, bit
 * for bit, one launch per stage.  Interior nodes 2..n-3 are updated, the two
 * ghost layers kept; w1, w2 are scratch fields (each call copies u into both).
 * Requires n >= 5 per axis, nx*ny*nz < 2^31 and nx*ny < 2^29. */
int mgad_heat13_rk3_steps(hipStream_t stream, unsigned int nx, unsigned int ny,
                          unsigned int nz, double dt, double *u, double *w1,
                          double *w2, int nsteps, int use_graph);

/* ------------------------------------------------------------------------ */
/* Host helpers (CPU only)                                                   */
/* ------------------------------------------------------------------------ */

/* u = 0.5 + sum_{m<nmodes} a_m sin(pi (kx_m x + ky_m y + kz_m z) + phi_m)
 * on the periodic grid x_i = lo + i*(hi-lo)/n; integer wavenumbers uniform in
 * 1..kmax, a_m uniform in [0, amp), phi_m uniform in [0, 2pi), drawn from
 * SplitMix64(seed).  Host array. */
int mgad_ic_fourier(double *u, unsigned int nx, unsigned int ny,
                    unsigned int nz, double lo, double hi, uint64_t seed,
                    int nmodes, int kmax, double amp, double offset);

/* u = exp(-|x-c|^2/(2 sigma^2)) + noise*U(-1,1) on x_i = i/(n-1) in [0,1]^3,
 * boundary nodes set to 0.  Host array. */
int mgad_ic_gauss_noise(double *u, unsigned int nx, unsigned int ny,
                        unsigned int nz, double sigma, double noise,
                        uint64_t seed);

/* SaveBinary3D, SingleGPU/Burgers3d_WENO5/tools.cpp:59-87: writes the host
 * field as 32-bit floats, x fastest, no header; path NULL writes result.bin
 * as the reference does. */
int mgad_save_binary3d(const double *u, unsigned int nx, unsigned int ny,
                       unsigned int nz, const char *path);
/* Reads a SaveBinary3D file of exactly nx*ny*nz floats into doubles. */
int mgad_load_binary3d(double *u, unsigned int nx, unsigned int ny,
                       unsigned int nz, const char *path);
/* Call_Init3d, SingleGPU/Burgers3d_WENO5/tools.cpp:92-140 (ic 1 square
 * jump, 2 zero, 3 Gaussian exp(-|x-1|^2/0.1) at x = i*dx).  Host array. */
int mgad_init3d_burgers(int ic, double *u0, double dx, double dy, double dz,
                        unsigned int nx, unsigned int ny, unsigned int nz);
/* Call_Init3D, SingleGPU/Diffusion3d_baselineCode/tools.cpp:92-143 (ic 1
 * top/bottom planes at 1, 2 sine mode, 3 heat kernel at t0 = 0.1 centred at
 * (5,5,5); ic 2, 3 with zero boundary nodes).  Host array. */
int mgad_init3d_heat(int ic, double *u0, double dx, double dy, double dz,
                     unsigned int nx, unsigned int ny, unsigned int nz);
/* CalcError, SingleGPU/Diffusion3d_baselineCode/tools.cpp:156-180, with the
 * node index corrected: norms of u minus the heat kernel (0.1/t)^(3/2)
 * exp(-r^2/(4t)) centred at (5,5,5); out = {L1, L2, Linf} with L1, L2
 * weighted by dx*dy*dz as the reference prints them. */
int mgad_heat_error_norms(const double *u, double t, double dx, double dy,
                          double dz, unsigned int nx, unsigned int ny,
                          unsigned int nz, double *out);
/* Init_subdomain, MultiGPU/Burgers3d_Baseline/Tools.c:181-199 (also
 * Diffusion3d_Baseline/Tools.c:180-198): copies planes n*_Nz .. n*_Nz+_Nz+5
 * of the global Nx x Ny x (P*_Nz+6) field h_q into the slab h_s_q
 * (Nx x Ny x (_Nz+6): 3 ghost planes each side).  Host arrays. */
int mgad_init_subdomain(const double *h_q, double *h_s_q, unsigned int n,
                        unsigned int Nx, unsigned int Ny, unsigned int _Nz);
/* Merge_domains, MultiGPU/Burgers3d_Baseline/Tools.c:204-223 (also
 * Diffusion3d_Baseline/Tools.c:203-222): copies the slab's own planes
 * 3 .. _Nz+2 into planes n*_Nz+3 .. n*_Nz+_Nz+2 of h_q.  Host arrays. */
int mgad_merge_domains(const double *h_s_q, double *h_q, unsigned int n,
                       unsigned int Nx, unsigned int Ny, unsigned int _Nz);

#ifdef __cplusplus
}
#endif
#endif /* MGAD_H */
