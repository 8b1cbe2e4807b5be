// mgad_common.hpp -- shared device arithmetic and host plumbing.
//
// Every translation unit is compiled with -ffp-contract=off: an expression
// written a*b + c is a multiply then an add, each correctly rounded, the same
// binary64 evaluation a host compiler gives the reference's C source.  Fused
// multiply-adds appear only where __builtin_fma is spelled out.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mgad.h"

namespace mgad {

// ---------------------------------------------------------------------------
// Host: sticky error code
// ---------------------------------------------------------------------------
void set_error(int code);
int check_hip(hipError_t e);           // records MGAD_ERR_HIP, returns status
int check_launch();                    // hipGetLastError after a launch

// Diffusion coefficients of CopyToConstantMemory, passed to kernels by value.
struct DiffCoef { double kx, ky, kz; };
DiffCoef diff_coef();

// ---------------------------------------------------------------------------
// Device: WENO5 arithmetic.
// Constants follow SingleGPU/Burgers3d_WENO5/Burgers.h:20-28 after macro
// expansion: D0N = 1./10. etc. are folded to the correctly rounded doubles.
// ---------------------------------------------------------------------------
constexpr double kD0N = 1. / 10., kD1N = 6. / 10., kD2N = 3. / 10.;
constexpr double kD0P = 3. / 10., kD1P = 6. / 10., kD2P = 1. / 10.;
constexpr double kC1312 = 13. / 12., kC14 = 1. / 4.;

enum : int { kWenoZ = 0, kWenoJS = 1 };

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

// First term of a smoothness indicator, 13/12 (a-2b+c)^2 (kernels.cu:157-159):
// it depends only on the sub-stencil's centre, so callers that march along a
// line compute it once per node and pass it to weno_side_s.
__device__ __forceinline__ double sterm(double a, double b, double c) {
    const double s = (a - 2.0 * b) + c;
    return (kC1312 * s) * s;
}

// Jiang-Shu smoothness indicators of the three 3-point sub-stencils
// (kernels.cu:157-159): 13/12 (a-2b+c)^2 + 1/4 (a-4b+3c)^2, etc.
__device__ __forceinline__ double smooth0(double a, double b, double c) {
    const double s = (a - 2.0 * b) + c;
    const double q = (a - 4.0 * b) + 3.0 * c;
    return (kC1312 * s) * s + (kC14 * q) * q;
}
__device__ __forceinline__ double smooth1(double b, double c, double d) {
    const double s = (b - 2.0 * c) + d;
    const double q = b - d;
    return (kC1312 * s) * s + (kC14 * q) * q;
}
__device__ __forceinline__ double smooth2(double c, double d, double e) {
    const double s = (c - 2.0 * d) + e;
    const double q = (3.0 * c - 4.0 * d) + e;
    return (kC1312 * s) * s + (kC14 * q) * q;
}

// WENO-Z weight d*(1 + (tau5/(beta+eps))^P), kernels.cu:163-165 (P = 1);
// P = 2 is the Borges et al. (2008) choice.
template <int P>
__device__ __forceinline__ double zweight(double d, double tau5, double beta, double eps) {
    const double q = tau5 / (beta + eps);
    return d * (1.0 + (P == 2 ? q * q : q));
}

// One side of the reconstruction with the 13/12 terms S0..S2 supplied:
// PLUS: left-biased f+ side, v = f+ at i-2..i+2, weights (1,6,3)/10;
// else the right-biased f- side, v = f- at i-1..i+3, weights (3,6,1)/10.
// B_k = S_k + 1/4 q_k^2 is the reference's sum in its own order.
template <int WENO, int P, bool PLUS>
__device__ __forceinline__ double weno_side_s(double v0, double v1, double v2, double v3, double v4,
                                              double S0, double S1, double S2, double eps) {
    const double q0 = (v0 - 4.0 * v1) + 3.0 * v2;
    const double q1 = v1 - v3;
    const double q2 = (3.0 * v2 - 4.0 * v3) + v4;
    const double B0 = S0 + (kC14 * q0) * q0;
    const double B1 = S1 + (kC14 * q1) * q1;
    const double B2 = S2 + (kC14 * q2) * q2;
    double c0, c1, c2, d0, d1, d2;
    if (PLUS) {
        c0 = (2.0 * v0 - 7.0 * v1) + 11.0 * v2;
        c1 = (-v1 + 5.0 * v2) + 2.0 * v3;
        c2 = (2.0 * v2 + 5.0 * v3) - v4;
        d0 = kD0N; d1 = kD1N; d2 = kD2N;
    } else {
        c0 = (-v0 + 5.0 * v1) + 2.0 * v2;
        c1 = (2.0 * v1 + 5.0 * v2) - v3;
        c2 = (11.0 * v2 - 7.0 * v3) + 2.0 * v4;
        d0 = kD0P; d1 = kD1P; d2 = kD2P;
    }
    if (WENO == kWenoZ) {
        const double tau5 = fabs(B0 - B2);
        const double a0 = zweight<P>(d0, tau5, B0, eps);
        const double a1 = zweight<P>(d1, tau5, B1, eps);
        const double a2 = zweight<P>(d2, tau5, B2, eps);
        const double sum = (a0 + a1) + a2;
        return ((a0 * c0 + a1 * c1) + a2 * c2) / (6.0 * sum);
    } else {
        const double a0 = d0 / ((eps + B0) * (eps + B0));
        const double a1 = d1 / ((eps + B1) * (eps + B1));
        const double a2 = d2 / ((eps + B2) * (eps + B2));
        const double sum = (a0 + a1) + a2;
        const double o0 = a0 / sum, o1 = a1 / sum, o2 = a2 / sum;
        return ((o0 * c0 + o1 * c1) + o2 * c2) / 6.0;
    }
}

// Face flux from split fluxes with the 13/12 terms precomputed:
// Sp_k at the centres of the f+ sub-stencils (i-1, i, i+1), Sm_k at the
// centres of the f- sub-stencils (i, i+1, i+2).
template <int WENO, int P>
__device__ __forceinline__ double weno_face_s(double v0, double v1, double v2, double v3, double v4,
                                              double w0, double w1, double w2, double w3, double w4,
                                              double Sp0, double Sp1, double Sp2, double Sm0,
                                              double Sm1, double Sm2, double eps) {
    const double hp = weno_side_s<WENO, P, true>(v0, v1, v2, v3, v4, Sp0, Sp1, Sp2, eps);
    const double hm = weno_side_s<WENO, P, false>(w0, w1, w2, w3, w4, Sm0, Sm1, Sm2, eps);
    return hp + hm;
}

// Reconstructed flux at face i+1/2:  v = f+ at i-2..i+2, w = f- at i-1..i+3.
// WENO-Z: kernels.cu:142-192 (alpha / (6*sum)).
// WENO-JS: MultiGPU/Burgers3d_Baseline/Kernels.cu:112-220 (w = alpha/sum, /6).
template <int WENO, int P>
__device__ __forceinline__ double weno_face2(double v0, double v1, double v2, double v3, double v4,
                                             double w0, double w1, double w2, double w3, double w4,
                                             double eps) {
    return weno_face_s<WENO, P>(v0, v1, v2, v3, v4, w0, w1, w2, w3, w4, sterm(v0, v1, v2),
                                sterm(v1, v2, v3), sterm(v2, v3, v4), sterm(w0, w1, w2),
                                sterm(w1, w2, w3), sterm(w2, w3, w4), eps);
}

// Straight restatement of the same face (kept for the drop-in sweeps).
template <int WENO, int P>
__device__ __forceinline__ double weno_face(double v0, double v1, double v2, double v3, double v4,
                                            double w0, double w1, double w2, double w3, double w4,
                                            double eps) {
    double hp, hm;
    {
        const double B0 = smooth0(v0, v1, v2), B1 = smooth1(v1, v2, v3), B2 = smooth2(v2, v3, v4);
        const double p0 = (2.0 * v0 - 7.0 * v1) + 11.0 * v2;
        const double p1 = (-v1 + 5.0 * v2) + 2.0 * v3;
        const double p2 = (2.0 * v2 + 5.0 * v3) - v4;
        if (WENO == kWenoZ) {
            const double tau5 = fabs(B0 - B2);
            const double a0 = zweight<P>(kD0N, tau5, B0, eps);
            const double a1 = zweight<P>(kD1N, tau5, B1, eps);
            const double a2 = zweight<P>(kD2N, tau5, B2, eps);
            const double sum = (a0 + a1) + a2;
            hp = ((a0 * p0 + a1 * p1) + a2 * p2) / (6.0 * sum);
        } else {
            const double a0 = kD0N / ((eps + B0) * (eps + B0));
            const double a1 = kD1N / ((eps + B1) * (eps + B1));
            const double a2 = kD2N / ((eps + B2) * (eps + B2));
            const double sum = (a0 + a1) + a2;
            const double o0 = a0 / sum, o1 = a1 / sum, o2 = a2 / sum;
            hp = ((o0 * p0 + o1 * p1) + o2 * p2) / 6.0;
        }
    }
    {
        const double B0 = smooth0(w0, w1, w2), B1 = smooth1(w1, w2, w3), B2 = smooth2(w2, w3, w4);
        const double m0 = (-w0 + 5.0 * w1) + 2.0 * w2;
        const double m1 = (2.0 * w1 + 5.0 * w2) - w3;
        const double m2 = (11.0 * w2 - 7.0 * w3) + 2.0 * w4;
        if (WENO == kWenoZ) {
            const double tau5 = fabs(B0 - B2);
            const double a0 = zweight<P>(kD0P, tau5, B0, eps);
            const double a1 = zweight<P>(kD1P, tau5, B1, eps);
            const double a2 = zweight<P>(kD2P, tau5, B2, eps);
            const double sum = (a0 + a1) + a2;
            hm = ((a0 * m0 + a1 * m1) + a2 * m2) / (6.0 * sum);
        } else {
            const double a0 = kD0P / ((eps + B0) * (eps + B0));
            const double a1 = kD1P / ((eps + B1) * (eps + B1));
            const double a2 = kD2P / ((eps + B2) * (eps + B2));
            const double sum = (a0 + a1) + a2;
            const double o0 = a0 / sum, o1 = a1 / sum, o2 = a2 / sum;
            hm = ((o0 * m0 + o1 * m1) + o2 * m2) / 6.0;
        }
    }
    return hp + hm;
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

// SSP-RK3 stage at one node, kernels.cu:520-527.
__device__ __forceinline__ double rk_stage(int step, double dt, double u, double uo, double Lu) {
    if (step == 1) return uo + dt * Lu;
    if (step == 2) return 0.75 * uo + 0.25 * (u + dt * Lu);
    return (uo + 2.0 * (u + dt * Lu)) / 3.0;
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

}  // namespace mgad
