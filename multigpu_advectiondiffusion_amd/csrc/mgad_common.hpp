// mgad_common.hpp -- shared device arithmetic and host plumbing.
//
// Every translation unit is compiled with -ffp-contract=off: an expression
// written a*b + c is a multiply then an add, each correctly rounded, the same
// binary64 evaluation a host compiler gives the reference's C source.  Fused
// multiply-adds appear only where __builtin_fma is spelled out.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/mgad.h"

namespace mgad {

// ---------------------------------------------------------------------------
// Host: sticky error code
// ---------------------------------------------------------------------------
void set_error(int code);
int check_hip(hipError_t e);           // records MGAD_ERR_HIP, returns status
int check_launch();                    // hipGetLastError after a launch

// One lock for the library's host-side tables and caches (launch shapes,
// captured step graphs, per-stream slots), so entry points may be called
// from several host threads.  Recursive: an entry point that holds it reads
// the launch table, which takes it again.
std::recursive_mutex &host_mutex();

// The scheme of the Call_Adv_* sweeps (mgad_set_adv_scheme).
struct AdvParams {
    int weno, p;
    double eps;
    int lf;
    double alpha;
};
AdvParams adv_params();

// Destroys every step graph of a full cache.  It first waits for the device,
// so no exec is destroyed while a replay of it, queued by any host thread on
// any stream, is pending or running.
template <class Map>
void clear_graph_cache(Map &cache) {
    (void)hipDeviceSynchronize();
    for (auto &kv : cache) (void)hipGraphExecDestroy(kv.second);
    cache.clear();
}

// Diffusion coefficients of CopyToConstantMemory, passed to kernels by value.
struct DiffCoef { double kx, ky, kz; };
DiffCoef diff_coef();
// Call_Lu3d's operator on the heat stage kernel (heat.hip).
int heat13_lu_launch(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, const double *u,
                     double *Lu);
int diff13_b5_launch(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, const double *u, double *Lu,
                     double kx, double ky, double kz);
int lu13_b5_launch(hipStream_t s, unsigned nx, unsigned ny, unsigned nz, unsigned px, const double *u, double *Lu);

// ---------------------------------------------------------------------------
// Device: WENO5 arithmetic.
// Constants follow SingleGPU/Burgers3d_WENO5/Burgers.h:20-28 after macro
// expansion: D0N = 1./10. etc. are folded to the correctly rounded doubles.
// ---------------------------------------------------------------------------
constexpr double kD0N = 1. / 10., kD1N = 6. / 10., kD2N = 3. / 10.;
constexpr double kD0P = 3. / 10., kD1P = 6. / 10., kD2P = 1. / 10.;
constexpr double kC1312 = 13. / 12., kC14 = 1. / 4.;
constexpr double kC1o6 = 1. / 6.;  // SingleGPU/Burgers3d_WENO5_SharedMem/Burgers.h:30

// kWenoZB5S: WENO-Z with the shared-memory code's rounding of the final
// weight, *(1/6 * (1/sum)) (SingleGPU/Burgers3d_WENO5_SharedMem/
// kernels.cu:175-178, :198-201) instead of /(6*sum).
enum : int { kWenoZ = 0, kWenoJS = 1, kWenoZB5S = 2 };

// Correctly rounded a/b for a divisor fixed on the host, given rb = RN(1/b)
// computed there: q = RN(a*rb), rem = a - b*q (exact with an FMA), and
// RN(q + rem*rb) = RN(a/b) (Markstein's theorem; no over/underflow in the
// range taken).  Operands below 2^-900 in magnitude, zeros included, take the
// IEEE division so signs of zero and subnormal quotients are unchanged.
__device__ __forceinline__ double div_by(double a, double b, double rb) {
    if (__builtin_expect(fabs(a) < 0x1p-900, 0)) return a / b;
    const double q = a * rb;
    const double rem = __builtin_fma(-b, q, a);
    return __builtin_fma(rem, rb, q);
}

// k*x + y for k = +-2 or +-4, one instruction.  k*x is exact in binary64
// (a power-of-two scaling; the only exception is overflow, |x| >= 2^1021), so
// the single rounding of the fused form is the same double as the two-step
// RN(RN(k*x) + y) the source expressions spell out.  A subtraction a - k*b is
// written p2fma(-k, b, a): IEEE defines a - y as a + (-y).
__device__ __forceinline__ double p2fma(double k, double x, double y) { return __builtin_fma(k, x, y); }

// First term of a smoothness indicator, 13/12 (a-2b+c)^2 (kernels.cu:157-159):
// it depends only on the sub-stencil's centre, so callers that march along a
// line compute it once per node and pass it to weno_side_s.
__device__ __forceinline__ double sterm(double a, double b, double c) {
    const double s = p2fma(-2.0, b, a) + c;  // (a - 2b) + c
    return (kC1312 * s) * s;
}

// Jiang-Shu smoothness indicators of the three 3-point sub-stencils
// (kernels.cu:157-159): 13/12 (a-2b+c)^2 + 1/4 (a-4b+3c)^2, etc.
__device__ __forceinline__ double smooth0(double a, double b, double c) {
    const double s = p2fma(-2.0, b, a) + c;         // (a - 2b) + c
    const double q = p2fma(-4.0, b, a) + 3.0 * c;   // (a - 4b) + 3c
    return (kC1312 * s) * s + (kC14 * q) * q;
}
__device__ __forceinline__ double smooth1(double b, double c, double d) {
    const double s = p2fma(-2.0, c, b) + d;         // (b - 2c) + d
    const double q = b - d;
    return (kC1312 * s) * s + (kC14 * q) * q;
}
__device__ __forceinline__ double smooth2(double c, double d, double e) {
    const double s = p2fma(-2.0, d, c) + e;         // (c - 2d) + e
    const double q = p2fma(-4.0, d, 3.0 * c) + e;   // (3c - 4d) + e
    return (kC1312 * s) * s + (kC14 * q) * q;
}

// a/b by the reciprocal iteration that IEEE division compiles to on gfx950
// (v_rcp_f64, two Newton steps, q = a*r, one residual correction) without its
// operand scaling (v_div_scale), scaled last FMA (v_div_fmas) and special-case
// fix-up (v_div_fixup).  Those three return their inputs unchanged when
//   0 < b, 2^-1000 <= b <= 2^1000,
//   a = +0, or 2^-960 <= |a| <= 2^1000,
//   |exponent(a) - exponent(b)| <= 760,
// (no huge exponent gap, no subnormal reciprocal or quotient, no tiny
// numerator), so there the result is the same double as a / b
// (mgad_check_division compares the two on the device).
__device__ __forceinline__ double div_inrange(double a, double b) {
    double r = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-b, r, 1.0);
    r = __builtin_fma(r, e, r);
    const double q = a * r;
    const double rem = __builtin_fma(-b, q, a);
    return __builtin_fma(rem, r, q);
}

// When the divisions of weno_side_s may use div_inrange (its FD variant)
// with results identical to IEEE division.  With max|u| <= 2^40 over the
// stage input (and |alpha| <= 2^40 for a global LF split, checked here;
// callers check |u|) every split flux has |f| < 2^80, so each 13/12 s^2 < 2^165
// and 1/4 q^2 < 2^164: 0 <= beta_k < 2^166.
//   WENO-Z, 2^-150 <= eps <= 1:
//   - weights: b = beta+eps is in [2^-150, 2^167] and a = tau5 < 2^167.  For
//     tau5 = 0 or tau5 >= 2^-380 the ranges of div_inrange hold.  For
//     0 < tau5 < 2^-380 both quotients are below 2^-229, so 1 + q (p=1) and
//     1 + q^2 (p=2) round to exactly 1 either way and the weight is d_k in both.
//   - final division num/(6 sum): q = tau5/(beta+eps) < 2^317, so each
//     alpha_k = d_k (1 + q^p) < 2^634 and alpha_1 >= d_1 = 0.6: 6 sum is in
//     [3.6, 2^637].  The candidates are below 2^85 in magnitude, so
//     |num| < 2^721.  For |num| >= 2^-121 (checked per node) the exponent gap
//     is at most 757 and div_inrange applies; smaller numerators, zeros of
//     either sign and NaN take the IEEE division.
//   - the shared-memory rounding's 1/sum: a = 1, b = sum in [0.6, 2^635].
//   WENO-JS, 2^-190 <= eps <= 1: (eps+beta)^2 is in [2^-380, 2^334], so
//     alpha_k = d_k/(..)^2 is in [2^-338, 2^380] and alpha_k/sum in
//     [2^-720, 1].
// A NaN or infinite alpha fails the test.
template <int WENO>
__host__ __device__ inline bool fast_division_ok(double alpha, double eps) {
    const double eps_lo = WENO == kWenoJS ? 0x1p-190 : 0x1p-150;
    return fabs(alpha) <= 0x1p40 && eps >= eps_lo && eps <= 1.0;
}

// num / den of the final WENO-Z division under FD (fast_division_ok holds):
// div_inrange where |num| >= 2^-121, IEEE otherwise (a branch every lane of a
// wave skips in the common case).
__device__ __forceinline__ double div_final(double num, double den) {
    double r = div_inrange(num, den);
    if (__builtin_expect(!(fabs(num) >= 0x1p-121), 0)) r = num / den;
    return r;
}

// WENO-Z weight d*(1 + (tau5/(beta+eps))^P), kernels.cu:163-165 (P = 1);
// P = 2 is the Borges et al. (2008) choice.
template <int P, bool FD = false>
__device__ __forceinline__ double zweight(double d, double tau5, double beta, double eps) {
    const double q = FD ? div_inrange(tau5, beta + eps) : tau5 / (beta + eps);
    return d * (1.0 + (P == 2 ? q * q : q));
}

// One side of the reconstruction with the 13/12 terms S0..S2 supplied:
// PLUS: left-biased f+ side, v = f+ at i-2..i+2, weights (1,6,3)/10;
// else the right-biased f- side, v = f- at i-1..i+3, weights (3,6,1)/10.
// B_k = S_k + 1/4 q_k^2 is the reference's sum in its own order.
// FD: fast_division_ok holds for the stage, and the weight divisions use
// div_inrange.
template <int WENO, int P, bool PLUS, bool FD = false>
__device__ __forceinline__ double weno_side_s(double v0, double v1, double v2, double v3, double v4,
                                              double S0, double S1, double S2, double eps) {
    // the source expressions are in the comments; p2fma folds the exact
    // products by 2 and 4 into the adjacent addition
    const double q0 = p2fma(-4.0, v1, v0) + 3.0 * v2;  // (v0 - 4v1) + 3v2
    const double q1 = v1 - v3;
    const double q2 = p2fma(-4.0, v3, 3.0 * v2) + v4;  // (3v2 - 4v3) + v4
    const double B0 = S0 + (kC14 * q0) * q0;
    const double B1 = S1 + (kC14 * q1) * q1;
    const double B2 = S2 + (kC14 * q2) * q2;
    double c0, c1, c2, d0, d1, d2;
    if (PLUS) {
        c0 = p2fma(2.0, v0, -(7.0 * v1)) + 11.0 * v2;  // (2v0 - 7v1) + 11v2
        c1 = p2fma(2.0, v3, -v1 + 5.0 * v2);           // (-v1 + 5v2) + 2v3
        c2 = p2fma(2.0, v2, 5.0 * v3) - v4;            // (2v2 + 5v3) - v4
        d0 = kD0N; d1 = kD1N; d2 = kD2N;
    } else {
        c0 = p2fma(2.0, v2, -v0 + 5.0 * v1);           // (-v0 + 5v1) + 2v2
        c1 = p2fma(2.0, v1, 5.0 * v2) - v3;            // (2v1 + 5v2) - v3
        c2 = p2fma(2.0, v4, 11.0 * v2 - 7.0 * v3);     // (11v2 - 7v3) + 2v4
        d0 = kD0P; d1 = kD1P; d2 = kD2P;
    }
    if (WENO == kWenoZ || WENO == kWenoZB5S) {
        const double tau5 = fabs(B0 - B2);
        const double a0 = zweight<P, FD>(d0, tau5, B0, eps);
        const double a1 = zweight<P, FD>(d1, tau5, B1, eps);
        const double a2 = zweight<P, FD>(d2, tau5, B2, eps);
        const double sum = (a0 + a1) + a2;
        if (WENO == kWenoZB5S)
            return ((a0 * c0 + a1 * c1) + a2 * c2) * (kC1o6 * (FD ? div_inrange(1.0, sum) : 1.0 / sum));
        const double num = (a0 * c0 + a1 * c1) + a2 * c2;
        return FD ? div_final(num, 6.0 * sum) : num / (6.0 * sum);
    } else {
        const double e0 = (eps + B0) * (eps + B0);
        const double e1 = (eps + B1) * (eps + B1);
        const double e2 = (eps + B2) * (eps + B2);
        const double a0 = FD ? div_inrange(d0, e0) : d0 / e0;
        const double a1 = FD ? div_inrange(d1, e1) : d1 / e1;
        const double a2 = FD ? div_inrange(d2, e2) : d2 / e2;
        const double sum = (a0 + a1) + a2;
        const double o0 = FD ? div_inrange(a0, sum) : a0 / sum;
        const double o1 = FD ? div_inrange(a1, sum) : a1 / sum;
        const double o2 = FD ? div_inrange(a2, sum) : a2 / sum;
        return div_by((o0 * c0 + o1 * c1) + o2 * c2, 6.0, 1.0 / 6.0);
    }
}

// Face flux from split fluxes with the 13/12 terms precomputed:
// Sp_k at the centres of the f+ sub-stencils (i-1, i, i+1), Sm_k at the
// centres of the f- sub-stencils (i, i+1, i+2).
template <int WENO, int P, bool FD = false>
__device__ __forceinline__ double weno_face_s(double v0, double v1, double v2, double v3, double v4,
                                              double w0, double w1, double w2, double w3, double w4,
                                              double Sp0, double Sp1, double Sp2, double Sm0,
                                              double Sm1, double Sm2, double eps) {
    const double hp = weno_side_s<WENO, P, true, FD>(v0, v1, v2, v3, v4, Sp0, Sp1, Sp2, eps);
    const double hm = weno_side_s<WENO, P, false, FD>(w0, w1, w2, w3, w4, Sm0, Sm1, Sm2, eps);
    return hp + hm;
}

// Reconstructed flux at face i+1/2:  v = f+ at i-2..i+2, w = f- at i-1..i+3.
// WENO-Z: kernels.cu:142-192 (alpha / (6*sum)).
// WENO-JS: MultiGPU/Burgers3d_Baseline/Kernels.cu:112-220 (w = alpha/sum, /6).
template <int WENO, int P, bool FD = false>
__device__ __forceinline__ double weno_face2(double v0, double v1, double v2, double v3, double v4,
                                             double w0, double w1, double w2, double w3, double w4,
                                             double eps) {
    return weno_face_s<WENO, P, FD>(v0, v1, v2, v3, v4, w0, w1, w2, w3, w4, sterm(v0, v1, v2),
                                sterm(v1, v2, v3), sterm(v2, v3, v4), sterm(w0, w1, w2),
                                sterm(w1, w2, w3), sterm(w2, w3, w4), eps);
}

// Lax-Friedrichs split of f(u) = u^2/2 (kernels.cu:37-40, 224-238):
// f+- = 0.5*(f(u) +- a*u), a = |u| (local) or the global max|u|.
__device__ __forceinline__ void lf_split(double u, double a, double &fp, double &fm) {
    const double f = (0.5 * u) * u;
    const double au = a * u;
    fp = 0.5 * (f + au);
    fm = 0.5 * (f - au);
}
__device__ __forceinline__ void lf_split_local(double u, double &fp, double &fm) {
    const double f = (0.5 * u) * u;
    const double au = fabs(u) * u;
    fp = 0.5 * (f + au);
    fm = 0.5 * (f - au);
}

// Fourth-order second difference of the 13-point Laplacians,
// (-a + 16b - 30c + 16d - e) in the source's left-to-right order
// (kernels.cu:474-476, Diffusion3d_baselineCode/kernels.cu:94-97); 16x is
// exact, so it folds into an FMA.
__device__ __forceinline__ double d4(double a, double b, double c, double d, double e) {
    return p2fma(16.0, d, p2fma(16.0, b, -a) - 30.0 * c) - e;
}

// SSP-RK3 stage at one node, kernels.cu:520-527.
__device__ __forceinline__ double rk_stage(int step, double dt, double u, double uo, double Lu) {
    if (step == 1) return uo + dt * Lu;
    if (step == 2) return 0.75 * uo + 0.25 * (u + dt * Lu);
    return p2fma(2.0, u + dt * Lu, uo) / 3.0;  // (uo + 2(u + dt Lu)) / 3
}

// Maps the launch's linear block id to a tile so that the blocks one XCD
// runs (ids congruent mod 8) are contiguous tiles sharing halos in L2.
__device__ __forceinline__ unsigned xcd_tile(unsigned bid, unsigned nblocks) {
    if ((nblocks & 7u) != 0u) return bid;
    return (bid & 7u) * (nblocks >> 3) + (bid >> 3);
}

// Non-negative doubles order like their bit patterns as unsigned integers, so
// a max over |u| is an integer atomicMax; NaN bit patterns sort above +Inf.
__device__ __forceinline__ unsigned long long abs_bits(double x) {
    return static_cast<unsigned long long>(__double_as_longlong(fabs(x)));
}
__device__ __forceinline__ double bits_double(unsigned long long b) {
    return __longlong_as_double(static_cast<long long>(b));
}

// Wave-wide (64 lanes) maximum of an unsigned 64-bit value.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(v, off, 64);
        v = o > v ? o : v;
    }
    return v;
}

// Draws the next work-share number for the whole block from *counter (one
// added per draw) and returns it, wave-uniform, to every thread.  Wave 0 draws
// with all its lanes active (lane 0 adds 1, the others 0, and lane 0's old
// value is the draw) inside a branch on a wave-uniform value, so no lane runs
// apart from its wave around the two barriers: with the draw under
// `threadIdx.x == 0` the compiler may merge it with another lane-0 region of
// the loop and move it out of the barrier loop, which then never ends.
// slot: a __shared__ word of the caller.
template <class C>
__device__ __forceinline__ C block_draw(C *counter, C &slot) {
    const unsigned wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wave == 0u) {
        const C v = atomicAdd(counter, (threadIdx.x & 63u) == 0u ? C(1) : C(0));
        if (sizeof(C) == 8) {
            const unsigned long long b = static_cast<unsigned long long>(v);
            const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b));
            const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b >> 32));
            slot = static_cast<C>((static_cast<unsigned long long>(hi) << 32) | lo);
        } else {
            slot = static_cast<C>(__builtin_amdgcn_readfirstlane(static_cast<unsigned>(v)));
        }
    }
    __syncthreads();
    C s;
    if (sizeof(C) == 8) {
        const unsigned long long b = static_cast<unsigned long long>(slot);
        const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b));
        const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(b >> 32));
        s = static_cast<C>((static_cast<unsigned long long>(hi) << 32) | lo);
    } else {
        s = static_cast<C>(__builtin_amdgcn_readfirstlane(static_cast<unsigned>(slot)));
    }
    __syncthreads();  // every thread has read the slot before the next draw
    return s;
}

}  // namespace mgad
