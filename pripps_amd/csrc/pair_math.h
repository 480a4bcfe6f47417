// Device helpers shared by the pair kernels: constant-address-space loads,
// the VGPR copy of wave-uniform values, and the FP32 sine/cosine of a phase in
// revolutions and square root / reciprocal square root that the pair seeds use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace pripps {

constexpr double kInv2Pi = 0.15915494309189533577;  // 1 / (2 pi)

// Read-only data shared by a whole wave (atoms, items, structure records) is
// read through the constant address space so that uniform indices become
// scalar loads into SGPRs instead of 64 identical vector loads.
#define PRIPPS_CONST __attribute__((address_space(4)))

template <typename T>
__device__ __forceinline__ const PRIPPS_CONST T *as_const(const T *p) {
    return (const PRIPPS_CONST T *)(p);
}

// A copy of x in a vector register.  On gfx950 a VALU instruction that
// reads a scalar register issues at half the rate of one whose operands are
// all vector registers (measured: 0.25 vs 0.47 wave instructions per
// SIMD-cycle for v_add_f32 / v_fmac_f32), so wave-uniform values that feed
// an inner loop are moved once into a VGPR.  The register move is opaque to
// the compiler, which would otherwise fold the scalar operand back in.
__device__ __forceinline__ float to_vgpr(float x) {
    float v;
    asm("v_mov_b32 %0, %1" : "=v"(v) : "v"(x));
    return v;
}

__device__ __forceinline__ double to_vgpr(double x) { return x; }

template <typename T>
__device__ __forceinline__ Atom<T> load_atom(const PRIPPS_CONST Atom<T> *p, int j) {
    Atom<T> a;
    a.x = p[j].x;
    a.y = p[j].y;
    a.z = p[j].z;
    a.w = p[j].w;
    return a;
}

__device__ __forceinline__ Item load_item(const Item *items, int64_t w) {
    const PRIPPS_CONST Item *p = as_const(items) + w;
    Item it;
    it.s = p->s;
    it.i0 = p->i0;
    it.icount = p->icount;
    it.ikind = p->ikind;
    it.j0 = p->j0;
    it.j1 = p->j1;
    it.strip = p->strip;
    it.pad = p->pad;
    return it;
}

// sin(2 pi t) and cos(2 pi t) - 1 for |t| <= 1/8 revolution.  Minimax
// polynomials whose leading coefficients (2 pi, -2 pi^2) are split into float
// hi + lo parts; the coefficients come from sincos_coeffs.py in this folder.
// Max error in exact arithmetic: sine 8.8e-11 relative, cosine 1.0e-10
// absolute, so what remains is the rounding of the last operations, which
// varies erratically from pair to pair and averages out over the pair sum.
// The hardware sine and cosine are not used for the recurrence seeds: their
// error, though below 1.3e-7, follows the phase smoothly (it includes a
// relative scale error of -3.2e-8 on both), so it is shared by all near
// neighbours of similar length, and the recurrence carries an error in
// delta_1 forward growing with the step count.  At the deep minima of I(q)
// that coherent error is amplified several hundred times.
constexpr float kS1_HI = 6.283185482e+00f;
constexpr float kS1_LO = -1.748455531e-07f;
constexpr float kS3 = -4.134170151e+01f;
constexpr float kS5 = 8.160494995e+01f;
constexpr float kS7 = -7.666957855e+01f;
constexpr float kS9 = 4.052738571e+01f;
constexpr float kC2_HI = -1.973920822e+01f;
constexpr float kC2_LO = -5.807431762e-07f;
constexpr float kC4 = 6.493936157e+01f;
constexpr float kC6 = -8.544713593e+01f;
constexpr float kC8 = 5.934757233e+01f;

__device__ __forceinline__ void sincos_poly(float t, float &s, float &cm1) {
    const float t2 = t * t;
    float ps = __builtin_fmaf(t2, kS9, kS7);
    ps = __builtin_fmaf(t2, ps, kS5);
    ps = __builtin_fmaf(t2, ps, kS3);
    ps = __builtin_fmaf(t2, ps, kS1_LO);
    s = __builtin_fmaf(t, kS1_HI, t * ps);
    float pc = __builtin_fmaf(t2, kC8, kC6);
    pc = __builtin_fmaf(t2, pc, kC4);
    pc = __builtin_fmaf(t2, pc, kC2_LO);
    cm1 = __builtin_fmaf(t2, pc, t2 * kC2_HI);
}

// sin(2 pi x) and cos(2 pi x) for a phase x in revolutions, |x| < 2^22.  The
// reduction to the nearest quarter turn is exact, and the result is odd
// (sine) and even (cosine) in x bit for bit.
__device__ __forceinline__ void sincos_rev(float x, float &s, float &c) {
    const float k = __builtin_rintf(4.0f * x);
    const float t = __builtin_fmaf(k, -0.25f, x);
    float st, cm1;
    sincos_poly(t, st, cm1);
    const float ct = 1.0f + cm1;
    const int qd = (int)k;
    const bool swap = (qd & 1) != 0;
    const float a = swap ? ct : st;
    const float b = swap ? st : ct;
    s = __int_as_float(__float_as_int(a) ^ ((qd & 2) << 30));
    c = __int_as_float(__float_as_int(b) ^ (((qd + 1) & 2) << 30));
}

// r = sqrt(r2) and rinv = 1/r, each within about half an ulp.  The hardware
// reciprocal-square-root estimate has an error that is a fixed function of the
// mantissa of r2: it is shared by every pair of about the same length, so it
// does not average out over pairs, and at a deep minimum of I(q) -- a small
// difference of the large same-sign sum over near neighbours and the sum over
// distant pairs -- it is amplified by the ratio of the two.  A plain Newton
// step y (1.5 - r2 y^2 / 2) does not remove it: the factor in parentheses is
// rounded at the coarse ulp of 1.  Here the corrections are formed from exact
// FMA residuals instead (r2 - r0^2, 1 - r y), which are small numbers whose
// rounding is negligible.
__device__ __forceinline__ void sqrt_and_rsqrt(float r2, float &r, float &rinv) {
    const float y = __builtin_amdgcn_rsqf(r2);
    const float r0 = r2 * y;
    const float d = __builtin_fmaf(-r0, r0, r2);
    r = __builtin_fmaf(d, 0.5f * y, r0);
    const float e = __builtin_fmaf(-r, y, 1.0f);
    rinv = __builtin_fmaf(y, e, y);
}

// The i = j terms of the six profiles at q-point m (sinc(0) = 1), and the
// profiles at q = 0 in closed form, (sum f_X)(sum f_Y).
__device__ inline void diagonal_and_zero(const PRIPPS_CONST DevStruct &sd, int m, double (&diag)[kProfiles],
                                  double (&zero)[kProfiles]) {
    double aa = 0, ee = 0, ae = 0, sa = 0, se = 0, aw = 0, ew = 0, sw = 0;
    for (int k = 0; k < sd.K; ++k) {
        const double n = sd.kind_n[k];
        const double a = sd.ffa[k * sd.nq + m];
        const double e = sd.ffe ? sd.ffe[k * sd.nq + m] : 0.0;
        const double wk = sd.kind_w[k];
        aa += n * a * a;
        ee += n * e * e;
        ae += n * a * e;
        sa += n * a;
        se += n * e;
        aw += a * wk;
        ew += e * wk;
        sw += wk;
    }
    const double W = sd.ffw ? sd.ffw[m] : 0.0;
    diag[0] = aa;
    diag[1] = ee;
    diag[2] = W * W * sd.sum_w2;
    diag[3] = ae;
    diag[4] = W * aw;
    diag[5] = W * ew;
    const double sh = W * sw;
    zero[0] = sa * sa;
    zero[1] = se * se;
    zero[2] = sh * sh;
    zero[3] = sa * se;
    zero[4] = sa * sh;
    zero[5] = se * sh;
}

}  // namespace pripps
