// FP64 sine and cosine of a phase in revolutions, shared by the Debye kernels
// (seeds, non-uniform grids) and the Direct kernel.  FP64 has no hardware
// sine; these replace the library sin / sincos / sincospi on the hot paths.
#pragma once

#include <hip/hip_runtime.h>

namespace pripps {

// Taylor polynomials of sin(2 pi t) and cos(2 pi t) - 1, coefficients (-1)^n (2 pi)^m / m! rounded to double (sincos_coeffs.py in
// this folder prints them); truncation bounds: quarter-turn sine to t^17
// 8.3e-20, quarter-turn cosine - 1 to t^16 2.0e-18, half-turn sine to t^23
// 5.2e-21 -- all far below the double rounding of the evaluation itself.  The
// reduction t = x - k/4 (or k/2) is exact for |x| < 2^29 revolutions.
constexpr double kDSinQ[9] = {6.283185307179586, -41.34170224039976, 81.60524927607506,
                              -76.70585975306139, 42.058693944897655, -15.09464257682299,
                              3.819952584848282, -0.7181223017785006, 0.10422916220813984};
constexpr double kDCosQ[8] = {-19.739208802178716, 64.9393940226683, -85.45681720669373,
                              60.24464137187666, -26.4262567833744, 7.903536371318469,
                              -1.714390711088672, 0.28200596845579123};
constexpr double kDSinH[12] = {6.283185307179586, -41.34170224039976, 81.60524927607506,
                               -76.70585975306139, 42.058693944897655, -15.09464257682299,
                               3.819952584848282, -0.7181223017785006, 0.10422916220813984,
                               -0.012031585942120627, 0.0011309237482517963,
                               -8.823533599243006e-05};

// The quarter-turn coefficients as values.  A kernel may hold them in
// registers the compiler cannot see through (an empty asm with a "+v"
// operand): with constants it evaluates each Horner step as a two-address
// FMA after a 64-bit copy of the coefficient, with register values as one
// three-address FMA.
struct SinCosCoeffs {
    double s[9];
    double c[8];
};

__device__ __forceinline__ SinCosCoeffs sincos_coeffs_f64() {
    SinCosCoeffs k;
#pragma unroll
    for (int i = 0; i < 9; ++i) k.s[i] = kDSinQ[i];
#pragma unroll
    for (int i = 0; i < 8; ++i) k.c[i] = kDCosQ[i];
    return k;
}

// sin(2 pi t) and cos(2 pi t) - 1 for |t| <= 1/8 revolution.
__device__ __forceinline__ void sincos_poly_f64(double t, double &s, double &cm1,
                                                const SinCosCoeffs &k) {
    const double t2 = t * t;
    double ps = k.s[8];
#pragma unroll
    for (int i = 7; i >= 0; --i) ps = fma(ps, t2, k.s[i]);
    s = t * ps;
    double pc = k.c[7];
#pragma unroll
    for (int i = 6; i >= 0; --i) pc = fma(pc, t2, k.c[i]);
    cm1 = t2 * pc;
}

__device__ __forceinline__ void sincos_poly_f64(double t, double &s, double &cm1) {
    sincos_poly_f64(t, s, cm1, sincos_coeffs_f64());
}

// sin(2 pi x) and cos(2 pi x) for x in revolutions: exact reduction to the
// nearest quarter turn, odd (sine) and even (cosine) in x bit for bit.
__device__ __forceinline__ void sincos_rev_f64(double x, double &s, double &c,
                                               const SinCosCoeffs &kc) {
    const double k = rint(4.0 * x);
    const double t = fma(k, -0.25, x);
    double st, cm1;
    sincos_poly_f64(t, st, cm1, kc);
    const double ct = 1.0 + cm1;
    const int qd = (int)k;
    const double a = (qd & 1) ? ct : st;
    const double b = (qd & 1) ? st : ct;
    s = (qd & 2) ? -a : a;
    c = ((qd + 1) & 2) ? -b : b;
}

__device__ __forceinline__ void sincos_rev_f64(double x, double &s, double &c) {
    sincos_rev_f64(x, s, c, sincos_coeffs_f64());
}

// sin(2 pi x) for x in revolutions: exact reduction to the nearest half turn,
// one polynomial on |t| <= 1/4.
__device__ __forceinline__ double sin_rev_f64(double x) {
    const double k = rint(2.0 * x);
    const double t = fma(k, -0.5, x);
    const double t2 = t * t;
    double p = kDSinH[11];
#pragma unroll
    for (int i = 10; i >= 0; --i) p = fma(p, t2, kDSinH[i]);
    const double v = t * p;
    return ((int)k & 1) ? -v : v;
}

}  // namespace pripps
