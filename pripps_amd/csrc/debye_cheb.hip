// FP32 pair kernel with a Chebyshev expansion in q per pair (gfx950).
//
// A strip is a run of up to kChebQ consecutive q-points of one structure, on
// any grid, with centre q_c and half-width H (x = (q - q_c) / H in [-1, 1]).
// For a pair at distance r, sin(q r) / r over the strip is, by the
// Jacobi-Anger expansion (a = q_c r, beta = H r),
//   sin(a + beta x) / r = (1/r) [sin a (J_0(beta) + 2 sum_m (-1)^m J_2m(beta) T_2m(x))
//                               + cos a (2 sum_m (-1)^m J_2m+1(beta) T_2m+1(x))],
// a Chebyshev series whose terms beyond order P are below 1e-9 of the pair's
// amplitude (the host picks P from beta_max = H r_max with the bound
// |J_p(beta)| <= (beta/2)^p / p!).  Each lane (one i atom) adds the P
// coefficients of its pairs; once per chunk of up to `block` j atoms of one
// class the wave sums them (DPP and lane-swap adds, bitwise the same in every
// lane), and each lane evaluates the series at its own q-points in FP64
// (Clenshaw) and carries the result into six FP64 profile sums, as the
// recurrence kernels do.  Per pair and strip this costs about 60 + 4.5 P
// instructions instead of three per q-point.
//
// Bessel weights per pair: Miller's backward recurrence on the scaled values
// u_p = J_p(beta) (2 / beta)^p, u_{p-1} = p u_p - z u_{p+1} with z = (beta/2)^2,
// started at N = P + 2 from (u_{N+1}, u_N) = (0, 2^-100).  The coefficients are
// exact integers, so no rounded constant enters the weights: their rounding
// errors vary erratically from pair to pair instead of following r.  The scale
// comes from Neumann's sum 1 = J_0 + 2 sum_m J_2m, which in u is
// u_0 + 2 sum_m u_2m z^m (Horner along the recurrence).  The unwanted
// solution that the arbitrary start admits shrinks against the wanted one by
// about z / p^2 per order on the way down; at the largest beta an order
// admits (J_P below 5e-10 of the amplitude) z / P^2 is at most 0.04, so two
// orders leave it under 2e-3 of J_P at order P, below 1e-12 of the pair's
// amplitude, and smaller still at every lower order.
//
// The phases use the same split constants as the recurrence kernels: q_c / 2 pi
// and H / 2 as float hi + lo parts, r and 1/r from residual-corrected square
// roots, sin a and cos a from the polynomial sincos of a phase in revolutions.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "internal.h"
#include "pair_math.h"

namespace pripps {

constexpr int kChebM = kChebQ / 64;  // q-points per lane

// v + v of the lane DPP_CTRL names, fused into one v_add_f32_dpp.
template <int DPP_CTRL>
__device__ __forceinline__ float dpp_add(float v) {
    const int o = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), DPP_CTRL, 0xF, 0xF,
                                              false);
    return v + __builtin_bit_cast(float, o);
}

// Sum over the 64 lanes, the same bits in every lane: quads (xor 1, xor 2),
// half rows (mirror of 8), rows (mirror of 16), then row pairs and halves by
// lane swaps.  Each step adds two operands that every participating lane sees
// in the same order, so all lanes end with one identical value.
__device__ __forceinline__ float wave_allsum(float v) {
    v = dpp_add<0xB1>(v);   // quad_perm [1,0,3,2]
    v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1]
    v = dpp_add<0x141>(v);  // row_half_mirror
    v = dpp_add<0x140>(v);  // row_mirror
    const int b = __builtin_bit_cast(int, v);
    auto r16 = __builtin_amdgcn_permlane16_swap(b, b, false, false);
    v = __builtin_bit_cast(float, (int)r16[0]) + __builtin_bit_cast(float, (int)r16[1]);
    const int c = __builtin_bit_cast(int, v);
    auto r32 = __builtin_amdgcn_permlane32_swap(c, c, false, false);
    return __builtin_bit_cast(float, (int)r32[0]) + __builtin_bit_cast(float, (int)r32[1]);
}

// Chebyshev coefficients of NP pairs, pair k added to its accumulator set
// acc[k][0..P): see the file comment.  amp = take / r (0 for padding lanes
// and pairs on or below the diagonal).
template <int P, int NP>
__device__ __forceinline__ void cheb_pairs(const float (&r)[NP], const float (&amp)[NP], float qc_hi,
                                           float qc_lo, float hh_hi, float hh_lo,
                                           float (&acc)[NP][P]) {
    constexpr int N = P + 2;
    float sa[NP], ca[NP], b[NP], z[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float arev = __builtin_fmaf(qc_hi, r[k], qc_lo * r[k]);
        sincos_rev(arev, sa[k], ca[k]);
        b[k] = __builtin_fmaf(hh_hi, r[k], hh_lo * r[k]);  // beta / 2
        z[k] = b[k] * b[k];
    }
    // pass 1: u_{p-1} = p u_p - z u_{p+1}, downward from u_N
    float u[NP][P];
    float un[NP], uc[NP], h[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        un[k] = 0.0f;
        uc[k] = 7.888609052e-31f;  // 2^-100
        h[k] = (N % 2 == 0) ? uc[k] : 0.0f;
    }
#pragma unroll
    for (int p = N; p >= 1; --p) {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
            // -z u_{p+1} is known a step early: one FMA per order on the
            // chain instead of a product and an FMA
            const float up = __builtin_fmaf(float(p), uc[k], -z[k] * un[k]);
            un[k] = uc[k];
            uc[k] = up;
            if (p - 1 < P) u[k][p - 1] = up;
            if ((p - 1) % 2 == 0) h[k] = __builtin_fmaf(h[k], z[k], up);
        }
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        // Neumann normalisation: S = u_0 + 2 sum_m u_2m z^m = 2 h - u_0
        const float S = __builtin_fmaf(2.0f, h[k], -u[k][0]);
        float inv = __builtin_amdgcn_rcpf(S);
        inv = __builtin_fmaf(inv, __builtin_fmaf(-S, inv, 1.0f), inv);
        // pass 2: c_p = u_p W_p with W_0 = amp sin a / S, W_1 = 2 amp cos a b / S,
        // W_2 = -2 z W_0 and W_{p+2} = -z W_p
        const float a0 = amp[k] * inv;
        float we = a0 * sa[k];
        float wo = 2.0f * ((a0 * ca[k]) * b[k]);
        acc[k][0] = __builtin_fmaf(u[k][0], we, acc[k][0]);
        if constexpr (P > 1) acc[k][1] = __builtin_fmaf(u[k][1], wo, acc[k][1]);
        we = -2.0f * (z[k] * we);
#pragma unroll
        for (int p = 2; p < P; ++p) {
            if (p % 2 == 0) {
                acc[k][p] = __builtin_fmaf(u[k][p], we, acc[k][p]);
                we = -z[k] * we;
            } else {
                wo = -z[k] * wo;
                acc[k][p] = __builtin_fmaf(u[k][p], wo, acc[k][p]);
            }
        }
    }
}

// One chunk's pairs: j in [jb, je), all of one class.  BYW: amplitudes are
// multiplied by w_j (the hydrated pass: sum w_j s).  Even and odd j go to two
// accumulator sets, added at the end: each FP32 partial sum then holds half
// the chunk.  The Chebyshev form evaluates about P rounded coefficient sums
// at every q-point, so their FP32 accumulation error weighs about sqrt(P)
// times more than in a sum per q-point; halving the summation length
// without halving the chunk keeps the carry's cost per pair.
template <int P, bool MASK, bool BYW>
__device__ __forceinline__ void cheb_block(const Atom<float> *__restrict__ atoms, int jb, int je,
                                           const Atom<float> &ai, int iidx, float lane_take,
                                           float qc_hi, float qc_lo, float hh_hi, float hh_lo,
                                           float (&C)[P]) {
    const PRIPPS_CONST Atom<float> *ca = as_const(atoms);
    float acc[2][P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        acc[0][p] = C[p];
        acc[1][p] = 0.0f;
    }
    auto geom = [&](const Atom<float> &aj, int j, float &r, float &amp) {
        const float dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
        const float r2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-12f);  // r -> 0: the limit s = q
        float rinv;
        sqrt_and_rsqrt(r2, r, rinv);
        float take = MASK ? (j > iidx ? 1.0f : 0.0f) : lane_take;
        if constexpr (BYW) take *= to_vgpr(aj.w);
        amp = take * rinv;
    };
    Atom<float> next = load_atom(ca, jb);
    int j = jb;
    for (; j + 1 < je; j += 2) {
        const Atom<float> a0 = next;
        const Atom<float> a1 = load_atom(ca, j + 1);
        next = load_atom(ca, min(j + 2, je - 1));  // prefetch the next j atom
        // one pair after the other: side by side their stored weights would
        // take 2 P more VGPRs, beyond three waves per SIMD
        float r[1], amp[1];
        geom(a0, j, r[0], amp[0]);
        cheb_pairs<P, 1>(r, amp, qc_hi, qc_lo, hh_hi, hh_lo, *reinterpret_cast<float(*)[1][P]>(&acc[0]));
        geom(a1, j + 1, r[0], amp[0]);
        cheb_pairs<P, 1>(r, amp, qc_hi, qc_lo, hh_hi, hh_lo, *reinterpret_cast<float(*)[1][P]>(&acc[1]));
    }
    if (j < je) {
        float r[1], amp[1];
        geom(next, j, r[0], amp[0]);
        cheb_pairs<P, 1>(r, amp, qc_hi, qc_lo, hh_hi, hh_lo, *reinterpret_cast<float(*)[1][P]>(&acc[0]));
    }
#pragma unroll
    for (int p = 0; p < P; ++p) C[p] = acc[0][p] + acc[1][p];
}

// One Clenshaw step in FP64 on an FP32 coefficient: b0 = 2x b1 - b2 + c.
// The conversion of c is bound to its use in one asm block; left to itself
// the compiler converts all P coefficients ahead of the recursion and holds
// their FP64 copies (2 P VGPRs) through the carry.
__device__ __forceinline__ double clenshaw_step(float c, double x2, double b1, double b2) {
    double b0, t;
    asm("v_cvt_f64_f32 %[t], %[c]\n\t"
        "v_add_f64 %[t], %[t], -%[b2]\n\t"
        "v_fma_f64 %[b0], %[x2], %[b1], %[t]"
        : [b0] "=&v"(b0), [t] "=&v"(t)
        : [c] "v"(c), [x2] "v"(x2), [b1] "v"(b1), [b2] "v"(b2));
    return b0;
}

// sum_p c_p T_p(x) (Clenshaw in FP64 on FP32 coefficients)
template <int P>
__device__ __forceinline__ double cheb_eval(const float (&c)[P], double x) {
    double b1 = 0.0, b2 = 0.0;
    const double x2 = 2.0 * x;
#pragma unroll
    for (int p = P - 1; p >= 1; --p) {
        const double b0 = clenshaw_step(c[p], x2, b1, b2);
        b2 = b1;
        b1 = b0;
    }
    // c_0 + x b1 - b2
    return clenshaw_step(c[0], x, b1, b2);
}

// A wave-uniform pointer the compiler cannot see through: the carry's
// per-lane addresses are then formed where they are used.  Hoisted out of
// the chunk loop (they are invariant per item), the 64-bit addresses of the
// q and form-factor loads of every q-point of the lane stay live through the
// pair loop and cost some 60 VGPRs there.
template <typename T>
__device__ __forceinline__ const T *fresh(const T *p) {
    asm volatile("" : "+s"(p));
    return p;
}

template <int P>
__global__ void __launch_bounds__(256) debye_cheb_kernel(TiledArgs args) {
    constexpr int M = kChebM;
    const int lane = __lane_id();
    // the wave's FP64 profile sums [profile][q-local]; each lane touches only
    // its own M columns (lane + 64 t), so no barrier is needed.  In VGPRs they
    // would take 48 registers from the pair loop's occupancy.
    __shared__ double sums[4][kProfiles * kChebQ];
    double *const pr = sums[threadIdx.x / 64];
    for (;;) {
        int w = 0;
        if (lane == 0) w = atomicAdd(args.counter, 1);
        w = __builtin_amdgcn_readfirstlane(__shfl(w, 0, 64));
        if (w >= args.n_items) break;

        const Item it = load_item(args.items, w);
        const PRIPPS_CONST DevStruct &sd = as_const(args.structs)[it.s];
        const Atom<float> *__restrict__ atoms = static_cast<const Atom<float> *>(sd.atoms);
        const int nq = sd.nq;

        const bool ivalid = lane < it.icount;
        const int iidx = ivalid ? it.i0 + lane : 0x7fffffff;
        const Atom<float> ai = atoms[ivalid ? it.i0 + lane : it.i0];
        const float wi = ivalid ? ai.w : 0.0f;
        const float lane_take = ivalid ? 1.0f : 0.0f;
        const bool wet_i = __ballot(wi != 0.0f) != 0;

        // strip geometry
        const int qf = sd.strip_first[it.strip];
        const int qend = sd.strip_first[it.strip + 1];
        const double qc = sd.strip_geo[2 * it.strip];
        const double H = sd.strip_geo[2 * it.strip + 1];
        const double qcr = qc * kInv2Pi, hh = 0.5 * H;
        const float qc_hi = to_vgpr(float(qcr));
        const float qc_lo = to_vgpr(float(qcr - double(float(qcr))));
        const float hh_hi = to_vgpr(float(hh));
        const float hh_lo = to_vgpr(float(hh - double(float(hh))));
        // the lane's q-points m = qf + 64 t + lane, at x = (q_m - q_c) / H
        const double inv_h = H > 0.0 ? 1.0 / H : 0.0;
#pragma unroll
        for (int k = 0; k < kProfiles; ++k)
#pragma unroll
            for (int t = 0; t < M; ++t) pr[k * kChebQ + 64 * t + lane] = 0.0;

        int j = it.j0;
        int c = 0;
        while (j < it.j1) {
            while (sd.cls_start[c + 1] <= j) ++c;
            const int jend = min(min(j + sd.block, it.j1), sd.cls_start[c + 1]);
            const int lk = sd.cls_kind[c];
            const bool wet = sd.cls_wet[c] != 0;
            const bool mask = j <= it.i0 + it.icount - 1;  // chunk meets the diagonal

            // pass 0: S = sum s and Si = sum w_i s; pass 1 (hydrated j only):
            // Sw = sum w_j s and Siw = sum w_i w_j s -- at the lane's q-points,
            // each carried into the FP64 profile sums as soon as it is known
#pragma unroll
            for (int pass = 0; pass < 2; ++pass) {
                if (pass == 1 && !wet) break;
                float C[P];
#pragma unroll
                for (int p = 0; p < P; ++p) C[p] = 0.0f;
                if (pass == 0) {
                    if (mask) cheb_block<P, true, false>(atoms, j, jend, ai, iidx, lane_take, qc_hi, qc_lo, hh_hi, hh_lo, C);
                    else cheb_block<P, false, false>(atoms, j, jend, ai, iidx, lane_take, qc_hi, qc_lo, hh_hi, hh_lo, C);
                } else {
                    if (mask) cheb_block<P, true, true>(atoms, j, jend, ai, iidx, lane_take, qc_hi, qc_lo, hh_hi, hh_lo, C);
                    else cheb_block<P, false, true>(atoms, j, jend, ai, iidx, lane_take, qc_hi, qc_lo, hh_hi, hh_lo, C);
                }
                // wave sums (the w_i-weighted ones first, then C in place),
                // then one q-point at a time: the series at x, the form
                // factors, the carry into the LDS profile sums
                float ci[P];
                if (wet_i) {
#pragma unroll
                    for (int p = 0; p < P; ++p) ci[p] = wave_allsum(C[p] * wi);
                }
#pragma unroll
                for (int p = 0; p < P; ++p) C[p] = wave_allsum(C[p]);
#pragma unroll 1
                for (int t = 0; t < M; ++t) {
                    const int m = qf + 64 * t + lane;
                    if (m >= qend) continue;
                    const double *fq = fresh(sd.q);
                    const double *fa = fresh(sd.ffa);
                    const double *fe = sd.ffe ? fresh(sd.ffe) : nullptr;
                    const double *fw = sd.ffw ? fresh(sd.ffw) : nullptr;
                    const double xt = (fq[m] - qc) * inv_h;
                    const double s0 = cheb_eval<P>(C, xt);
                    const double s1 = wet_i ? cheb_eval<P>(ci, xt) : 0.0;
                    double *o = pr + 64 * t + lane;
                    const double ak = fa[it.ikind * nq + m];
                    const double ek = fe ? fe[it.ikind * nq + m] : 0.0;
                    const double wq = fw ? fw[m] : 0.0;
                    const double al = fa[lk * nq + m];
                    const double el = fe ? fe[lk * nq + m] : 0.0;
                    if (pass == 0) {  // s0 = S, s1 = Si
                        o[0 * kChebQ] += 2.0 * ak * al * s0;
                        o[1 * kChebQ] += 2.0 * ek * el * s0;
                        o[3 * kChebQ] += (ak * el + al * ek) * s0;
                        o[4 * kChebQ] += wq * (al * s1);
                        o[5 * kChebQ] += wq * (el * s1);
                    } else {          // s0 = Sw, s1 = Siw
                        o[2 * kChebQ] += 2.0 * wq * wq * s1;
                        o[4 * kChebQ] += wq * (ak * s0);
                        o[5 * kChebQ] += wq * (ek * s0);
                    }
                }
            }
            j = jend;
        }
        double *out = args.partials + (int64_t)w * (kProfiles * kChebQ);
#pragma unroll
        for (int k = 0; k < kProfiles; ++k)
#pragma unroll
            for (int t = 0; t < M; ++t) out[k * kChebQ + 64 * t + lane] = pr[k * kChebQ + 64 * t + lane];
    }
}

// Orders the kernel is built for; the host rounds a structure's need up.
#define PRIPPS_CHEB_ORDERS(X) X(12) X(16) X(20) X(24) X(28) X(32) X(40)

hipError_t launch_cheb(int P, const TiledArgs &a, int blocks, hipStream_t st) {
    switch (P) {
#define PRIPPS_CASE(PP)                                                                     \
    case PP:                                                                                \
        hipLaunchKernelGGL(debye_cheb_kernel<PP>, dim3(blocks), dim3(256), 0, st, a);       \
        return hipGetLastError();
        PRIPPS_CHEB_ORDERS(PRIPPS_CASE)
#undef PRIPPS_CASE
    }
    return hipErrorInvalidValue;
}

hipError_t cheb_occupancy(int P, int *blocks_per_cu) {
    switch (P) {
#define PRIPPS_CASE(PP)                                                                     \
    case PP:                                                                                \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,                 \
                                                            debye_cheb_kernel<PP>, 256, 0);
        PRIPPS_CHEB_ORDERS(PRIPPS_CASE)
#undef PRIPPS_CASE
    }
    return hipErrorInvalidValue;
}

static bool cheb_order_ok(int P, double beta_max) {
    const double b = 0.5 * beta_max;
    double term = 1.0;  // b^P / P!
    for (int k = 1; k <= P; ++k) term *= b / k;
    double tail = 0.0, t = term;
    for (int k = P; k < P + 200 && t > 1e-30; ++k) {
        tail += t;
        t *= b / (k + 1);
    }
    return 2.0 * tail < 1e-9;
}

double cheb_beta_limit(int P) {
    double lo = 0.0, hi = 64.0;  // cheb_order_ok is monotone in beta
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (cheb_order_ok(P, mid)) lo = mid;
        else hi = mid;
    }
    return lo;
}

int cheb_order_for(double beta_max) {
    // smallest built order P with 2 sum_{p >= P} (beta/2)^p / p! < 1e-9
    // (|J_p(beta)| <= (beta/2)^p / p! for all p >= 0)
    const int orders[] = {12, 16, 20, 24, 28, 32, 40};
    for (int P : orders)
        if (cheb_order_ok(P, beta_max)) return P;
    return 0;  // beta too large for any built order
}

}  // namespace pripps
