// HIP kernels of the exact Debye engine (gfx950).
//
// Work decomposition
//   Atoms are sorted into class segments, hydrated classes first, then by
//   kind (a hydrated j atom therefore only meets the hydrated i atoms before
//   it, so few pairs need the w_j sums).  A work item is
//   one wave: a 64-atom i-tile of a single kind (one atom per lane) against a
//   span of j atoms, for one strip of Q q-points held by every lane.  Only
//   pairs with j > i are taken, so each unordered pair of a structure is
//   visited once (the upper triangle of the N x N pair matrix).
//
// Per pair and strip (uniform q grid, q_k = q_s + k dq):
//   a = q_s r, h = dq r / 2,  s_0 = sin(a)/r,  delta_1 = 2 cos(a + h) sin(h)/r,
//   lambda = 4 sin^2(h), then for k = 0..Q-1
//       acc0[k] += s;  accw[k] += w_j s;  s += delta;  delta -= lambda s
//   which is the angle-addition recurrence sin(a + (k+1)d) = 2 cos(d)
//   sin(a + kd) - sin(a + (k-1)d) written in Reinsch's difference form: the
//   constant lambda keeps its full relative accuracy for small d, where the
//   plain 2 cos(d) coefficient would lose it, so the FP32 phase does not drift.
//   Seeds use a polynomial sincos and residual-corrected r and 1/r (see
//   seed_f32).  The tiled kernel seeds every strip afresh; the dual kernel
//   (FP32, 128 q-points per item) seeds a dry pair once and lets the
//   recurrence run on through both 64-point halves.  On non-uniform grids
//   each term takes a direct hardware sin instead.
//
// Accumulation
//   Lanes sum at most `block` pairs in FP32 (acc0/accw), then a cross-lane
//   reduce-scatter leaves each lane the wave sums of Q/64 q-points, which are
//   multiplied by the kind form factors and added to six FP64 profile
//   accumulators.  Each item writes its FP64 partial profiles to its own slot;
//   a finalize kernel sums the slots in a fixed order (deterministic), divides
//   by q, and adds the diagonal.  Profiles never need the reference's 4.5 N^2
//   sinc count: the six profiles come from the colocated-set identity
//     sum_ij fa_i fb_j s_ij = sum_i fa_i fb_i + sum_{i<j} (fa_i fb_j + fa_j fb_i) s_ij.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "internal.h"
#include "pair_math.h"
#include "trig_f64.h"

namespace pripps {

// ---------------------------------------------------------------------------
// cross-lane helpers
// ---------------------------------------------------------------------------
template <int Q>
struct Split {
    // values per lane after the reduce-scatter
    static constexpr int M = Q >= 64 ? Q / 64 : 1;
    // lanes sharing one q-point when Q < 64
    static constexpr int kShare = Q >= 64 ? 1 : 64 / Q;
};

// q-local index held in slot t by `lane` after reduce_scatter.
template <int Q>
__device__ __forceinline__ int split_index(int lane, int t) {
    if constexpr (Q >= 64) return (Q / 64) * lane + t;
    else return lane / (64 / Q);
}

template <int Q>
__device__ __forceinline__ bool split_owner(int lane) {
    if constexpr (Q >= 64) return true;
    else return (lane % (64 / Q)) == 0;
}

// Sum in[0..Q) (optionally times a per-lane scale) over the 64 lanes; lane L
// receives the totals of the q-local indices split_index<Q>(L, t).
template <typename T, int Q, bool SCALED>
__device__ __forceinline__ void reduce_scatter(const T (&in)[Q], T scale, T (&out)[Split<Q>::M]) {
    const int lane = __lane_id();
    constexpr int H = Q / 2;
    T buf[H];
    {
        const bool up = lane & 32;
#pragma unroll
        for (int t = 0; t < H; ++t) {
            T lo = in[t], hi = in[t + H];
            if constexpr (SCALED) {
                lo *= scale;
                hi *= scale;
            }
            T keep = up ? hi : lo;
            T send = up ? lo : hi;
            buf[t] = keep + __shfl_xor(send, 32, 64);
        }
    }
    int len = H;
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) {
        if (len > 1) {
            const bool up = lane & o;
            const int half = len / 2;
#pragma unroll
            for (int t = 0; t < half; ++t) {
                T lo = buf[t], hi = buf[t + half];
                T keep = up ? hi : lo;
                T send = up ? lo : hi;
                buf[t] = keep + __shfl_xor(send, o, 64);
            }
            len = half;
        } else {
            buf[0] += __shfl_xor(buf[0], o, 64);
        }
    }
#pragma unroll
    for (int t = 0; t < Split<Q>::M; ++t) out[t] = buf[t];
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------
// per-pair seeds
// ---------------------------------------------------------------------------
struct Seed32 {
    float s, d, lam;
};

// FP32 seed: q0rev = q_s / 2pi split as q0rev + q0lo and hrev = dq / 4pi split
// as hrev + hlo (revolutions per A).  The lo parts remove the rounding of the
// constants, which would otherwise shift q coherently for every pair -- an
// error that deep minima of I(q), where |dlnI/dlnq| reaches ~100, amplify.
// small_h: the host guarantees hrev * r <= 1/8 for every pair.
// The strip-independent part of a pair's seed: r, 1/r (times take), and the
// half step h = dq r / 2.
struct PairGeom32 {
    float r, amp, sh, chm1, lam;
};

__device__ __forceinline__ PairGeom32 pair_geom_f32(float r2, float hrev, float hlo, float take,
                                                    bool small_h) {
    r2 = fmaxf(r2, 1e-12f);  // coincident atoms: the r -> 0 limit s_k -> q_k
    float r, rinv;
    sqrt_and_rsqrt(r2, r, rinv);
    const float hr = __builtin_fmaf(hrev, r, hlo * r);
    PairGeom32 g;
    if (small_h) {
        sincos_poly(hr, g.sh, g.chm1);
    } else {
        float ch;
        sincos_rev(hr, g.sh, ch);
        g.chm1 = ch - 1.0f;
    }
    g.r = r;
    g.amp = take * rinv;
    g.lam = 4.0f * (g.sh * g.sh);
    return g;
}

// The strip part: s_0 = sin(a)/r and delta_1 = 2 cos(a + h) sin(h)/r.
__device__ __forceinline__ Seed32 seed_from_geom(const PairGeom32 &g, float q0rev, float q0lo) {
    const float arev = __builtin_fmaf(q0rev, g.r, q0lo * g.r);
    float sa, ca;
    sincos_rev(arev, sa, ca);
    // cos(a + h) = cos a + (cos a (cos h - 1) - sin a sin h)
    const float cah = ca + __builtin_fmaf(ca, g.chm1, -sa * g.sh);
    Seed32 z;
    z.s = sa * g.amp;
    z.d = (2.0f * g.amp) * (cah * g.sh);
    z.lam = g.lam;
    return z;
}

__device__ __forceinline__ Seed32 seed_f32(float r2, float q0rev, float q0lo, float hrev,
                                           float hlo, float take, bool small_h) {
    return seed_from_geom(pair_geom_f32(r2, hrev, hlo, take, small_h), q0rev, q0lo);
}

struct Seed64 {
    double s, d, lam;
};

// FP64 seed: q0rev = q_s / 2pi and hrev = dq / 4pi in revolutions per A.
// small_h: the host guarantees hrev * r <= 1/8 for every pair.
__device__ __forceinline__ Seed64 seed_f64(double r2, double q0rev, double hrev, double take,
                                           bool small_h) {
    r2 = fmax(r2, 1e-24);
    const double r = sqrt(r2);
    const double rinv = 1.0 / r;
    double sa, ca, sh, chm1;
    sincos_rev_f64(q0rev * r, sa, ca);
    if (small_h) {
        sincos_poly_f64(hrev * r, sh, chm1);
    } else {
        double ch;
        sincos_rev_f64(hrev * r, sh, ch);
        chm1 = ch - 1.0;
    }
    // cos(a + h) = cos a + (cos a (cos h - 1) - sin a sin h)
    const double cah = ca + fma(ca, chm1, -sa * sh);
    const double amp = take * rinv;
    Seed64 z;
    z.s = sa * amp;
    z.d = (2.0 * amp) * (cah * sh);
    z.lam = 4.0 * (sh * sh);
    return z;
}

// ---------------------------------------------------------------------------
// the tiled pair kernel
// ---------------------------------------------------------------------------
// Two consecutive j atoms, fetched with one 32-byte scalar load.
struct alignas(16) AtomPair32 {
    Atom<float> a, b;
};

__device__ __forceinline__ AtomPair32 load_pair(const PRIPPS_CONST Atom<float> *ca, int j) {
    const PRIPPS_CONST float *f = (const PRIPPS_CONST float *)(ca + j);
    AtomPair32 p;
    p.a.x = f[0];
    p.a.y = f[1];
    p.a.z = f[2];
    p.a.w = f[3];
    p.b.x = f[4];
    p.b.y = f[5];
    p.b.z = f[6];
    p.b.w = f[7];
    return p;
}

// FP32, uniform q grid: two j atoms per iteration (one paired scalar load).
// INTERLEAVE runs their two recurrence chains step by step together; otherwise
// one after the other (which measured faster on gfx950).  Dry
// chunks cost 3 VALU ops per pair-q term, hydrated chunks 4.  The diagonal
// test j > i is always applied; padding lanes carry i = INT_MAX and so never
// contribute.  Odd chunks end with one single-atom step.
template <int Q, bool WET, bool SMALL_H, bool INTERLEAVE>
__device__ __forceinline__ void pair_block_f32x2(const Atom<float> *__restrict__ atoms, int jb,
                                                 int je, const Atom<float> &ai, int iidx,
                                                 float q0p, float q0lo, float hp, float hlo,
                                                 float (&acc0)[Q], float (&accw)[Q]) {
    const PRIPPS_CONST Atom<float> *ca = as_const(atoms);
    int j = jb;
    AtomPair32 nxt{};
    if (je - jb >= 2) nxt = load_pair(ca, jb);  // never reads past je - 1
    for (; j + 1 < je; j += 2) {
        const AtomPair32 cur = nxt;
        nxt = load_pair(ca, min(j + 2, je - 2));  // prefetch the next two atoms
        const Atom<float> &a = cur.a, &b = cur.b;
        const float ax = ai.x - a.x, ay = ai.y - a.y, az = ai.z - a.z;
        const float bx = ai.x - b.x, by = ai.y - b.y, bz = ai.z - b.z;
        const float ta = j > iidx ? 1.0f : 0.0f;
        const float tb = j + 1 > iidx ? 1.0f : 0.0f;
        const Seed32 za = seed_f32(ax * ax + ay * ay + az * az, q0p, q0lo, hp, hlo, ta, SMALL_H);
        const Seed32 zb = seed_f32(bx * bx + by * by + bz * bz, q0p, q0lo, hp, hlo, tb, SMALL_H);
        float s1 = za.s, d1 = za.d, s2 = zb.s, d2 = zb.d;
        const float l1 = za.lam, l2 = zb.lam;
        const float wa = WET ? to_vgpr(a.w) : 0.0f, wb = WET ? to_vgpr(b.w) : 0.0f;
        if constexpr (INTERLEAVE) {
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                acc0[k] += s1 + s2;
                if constexpr (WET) {
                    accw[k] = __builtin_fmaf(wa, s1, accw[k]);
                    accw[k] = __builtin_fmaf(wb, s2, accw[k]);
                }
                s1 += d1;
                s2 += d2;
                d1 = __builtin_fmaf(-l1, s1, d1);
                d2 = __builtin_fmaf(-l2, s2, d2);
            }
        } else {
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                acc0[k] += s1;
                if constexpr (WET) accw[k] = __builtin_fmaf(wa, s1, accw[k]);
                s1 += d1;
                d1 = __builtin_fmaf(-l1, s1, d1);
            }
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                acc0[k] += s2;
                if constexpr (WET) accw[k] = __builtin_fmaf(wb, s2, accw[k]);
                s2 += d2;
                d2 = __builtin_fmaf(-l2, s2, d2);
            }
        }
    }
    if (j < je) {
        const Atom<float> a = load_atom(ca, j);
        const float ax = ai.x - a.x, ay = ai.y - a.y, az = ai.z - a.z;
        const float ta = j > iidx ? 1.0f : 0.0f;
        const Seed32 za = seed_f32(ax * ax + ay * ay + az * az, q0p, q0lo, hp, hlo, ta, SMALL_H);
        float s1 = za.s, d1 = za.d;
        const float l1 = za.lam;
        const float wa = WET ? to_vgpr(a.w) : 0.0f;
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            acc0[k] += s1;
            if constexpr (WET) accw[k] = __builtin_fmaf(wa, s1, accw[k]);
            s1 += d1;
            d1 = __builtin_fmaf(-l1, s1, d1);
        }
    }
}

// BYW: the pass of a hydrated chunk in the one-accumulator kernel that sums
// w_j s into acc0 (acc0 += w_j s, the same arithmetic as the two-array
// kernel's accw).  Scaling the pair amplitude by w_j instead would round
// w_j / r once per pair with an error that depends on r alone; on lattice
// points, where many pairs share r, that error adds up coherently.
// SIN_DIRECT (FP32, non-uniform grids): every phase is below 255 revolutions,
// so the hardware sine takes it without a fract (see the term loop).
template <typename T, int Q, bool UNIFORM, bool WET, bool MASK, int SMALL_H, bool BYW = false,
          bool SIN_DIRECT = false>
__device__ __forceinline__ void pair_block(const Atom<T> *__restrict__ atoms, int jb, int je,
                                           const Atom<T> &ai, int iidx, T lane_take,
                                           const T (&qv)[Q], const T (&ql)[Q], T q0p, T q0lo,
                                           T hp, T hlo, bool small_h_rt, T (&acc0)[Q],
                                           T (&accw)[Q]) {
    // SMALL_H: 1 / 0 fixed at compile time, -1 decided per pair from small_h_rt
    const bool small_h = SMALL_H < 0 ? small_h_rt : SMALL_H == 1;
    const PRIPPS_CONST Atom<T> *ca = as_const(atoms);
    Atom<T> next = load_atom(ca, jb);
    for (int j = jb; j < je; ++j) {
        const Atom<T> aj = next;
        next = load_atom(ca, min(j + 1, je - 1));  // prefetch the next j atom
        const T dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T wj = (WET || BYW) ? to_vgpr(aj.w) : T(0);
        // padding lanes (iidx = INT_MAX, lane_take = 0) never contribute
        const T take = MASK ? (j > iidx ? T(1) : T(0)) : lane_take;
        if constexpr (UNIFORM) {
            T s, d, lam;
            if constexpr (sizeof(T) == 4) {
                Seed32 z = seed_f32(r2, q0p, q0lo, hp, hlo, take, small_h);
                s = z.s;
                d = z.d;
                lam = z.lam;
            } else {
                Seed64 z = seed_f64(r2, q0p, hp, take, small_h);
                s = z.s;
                d = z.d;
                lam = z.lam;
            }
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                if constexpr (BYW) acc0[k] += wj * s;
                else acc0[k] += s;
                if constexpr (WET) accw[k] += wj * s;
                s += d;
                d -= lam * s;
            }
        } else {
            if constexpr (sizeof(T) == 4) {
                const float rr2 = fmaxf(r2, 1e-12f);
                float r, rinv;
                sqrt_and_rsqrt(rr2, r, rinv);
                const float amp = take * rinv;
#pragma unroll
                for (int k = 0; k < Q; ++k) {
                    // phase (q / 2pi) r with q / 2pi split as qv + ql: see seed_f32.
                    // v_sin_f32 reduces |x| <= 256 revolutions itself: with every
                    // phase proven below 255 (SIN_DIRECT) the fract is left out.
                    // For x >= 0 the two agree but for a few ulp-level cases (0.03 %
                    // of inputs in [1, 256), measured), with the same error bound.
                    const float x = __builtin_fmaf(qv[k], r, ql[k] * r);
                    const float xr = SIN_DIRECT ? x : __builtin_amdgcn_fractf(x);
                    const float s = __builtin_amdgcn_sinf(xr) * amp;
                    if constexpr (BYW) acc0[k] += wj * s;
                    else acc0[k] += s;
                    if constexpr (WET) accw[k] += wj * s;
                }
            } else {
                const double rr2 = fmax(r2, 1e-24);
                const double r = sqrt(rr2);
                const double amp = take / r;
#pragma unroll
                for (int k = 0; k < Q; ++k) {
                    const double s = sin_rev_f64(qv[k] * r) * amp;
                    if constexpr (BYW) acc0[k] += wj * s;
                    else acc0[k] += s;
                    if constexpr (WET) accw[k] += wj * s;
                }
            }
        }
    }
}

// VARIANT (FP32 uniform grids only): 1 paired loads, chains one after the other;
// 2 paired loads, chains interleaved; 3 one atom per iteration, half-step
// sine path chosen per pair; 4 one atom per iteration, path fixed per chunk.
// ONEACC (FP32 on uniform grids, variants 3 and 4): one accumulator array.
// A hydrated j chunk runs twice, first for sum s, then for sum w_j s.  Atoms are sorted hydrated-first, so
// a hydrated j only meets the hydrated i before it and the second pass covers
// a few percent of the pairs; in exchange the kernel needs about 60 fewer
// VGPRs and runs three waves per SIMD instead of two, which lets the VALU
// dual-issue far more often.
template <typename T, int Q, bool UNIFORM, int VARIANT, bool ONEACC = false>
__device__ __forceinline__ void tiled_body(const TiledArgs &args) {
    constexpr bool INTERLEAVE = VARIANT == 2;
    constexpr int M = Split<Q>::M;
    const int lane = __lane_id();
    const bool owner = split_owner<Q>(lane);

    for (;;) {
        int w = 0;
        if (lane == 0) w = atomicAdd(args.counter, 1);
        w = __builtin_amdgcn_readfirstlane(__shfl(w, 0, 64));
        if (w >= args.n_items) break;

        const Item it = load_item(args.items, w);
        const PRIPPS_CONST DevStruct &sd = as_const(args.structs)[it.s];
        const Atom<T> *__restrict__ atoms = static_cast<const Atom<T> *>(sd.atoms);
        const int nq = sd.nq;

        const bool ivalid = lane < it.icount;
        const int iidx = ivalid ? it.i0 + lane : 0x7fffffff;
        Atom<T> ai = atoms[ivalid ? it.i0 + lane : it.i0];
        const T wi = ivalid ? ai.w : T(0);
        const T lane_take = ivalid ? T(1) : T(0);
        const bool wet_i = __ballot(wi != T(0)) != 0;

        const int qbase = it.strip * Q;
        const bool small_h = sd.small_h != 0;
        const bool sin_direct = sd.sin_direct != 0;
        // strip constants
        T qv[Q], ql[Q];
        T q0p = T(0), q0lo = T(0), hp = T(0), hlo = T(0);
        if constexpr (UNIFORM) {
            const double qs = sd.q0 + double(qbase) * sd.dq;
            if constexpr (sizeof(T) == 4) {
                q0p = float(qs * kInv2Pi);
                q0lo = float(qs * kInv2Pi - double(q0p));
                const double hd = sd.dq * (0.5 * kInv2Pi);
                hp = float(hd);
                hlo = float(hd - double(hp));
                // item-uniform: used by every pair's seed
                q0p = to_vgpr(q0p);
                q0lo = to_vgpr(q0lo);
                hp = to_vgpr(hp);
                hlo = to_vgpr(hlo);
            } else {
                q0p = qs * kInv2Pi;
                hp = sd.dq * (0.5 * kInv2Pi);
            }
#pragma unroll
            for (int k = 0; k < Q; ++k) qv[k] = ql[k] = T(0);
        } else {
#pragma unroll
            for (int k = 0; k < Q; ++k) {
                const int m = min(qbase + k, nq - 1);
                if constexpr (sizeof(T) == 4) {
                    const double qd = sd.q[m] * kInv2Pi;
                    qv[k] = float(qd);
                    ql[k] = float(qd - double(qv[k]));
                } else {
                    qv[k] = sd.q[m] * kInv2Pi;
                    ql[k] = T(0);
                }
            }
        }

        // per-lane form factors of the tile kind at the lane's q-points
        double ak[M], ek[M], wq[M];
        double pAA[M], pEE[M], pHH[M], pAE[M], pAH[M], pEH[M];
#pragma unroll
        for (int t = 0; t < M; ++t) {
            const int m = min(qbase + split_index<Q>(lane, t), nq - 1);
            ak[t] = sd.ffa[it.ikind * nq + m];
            ek[t] = sd.ffe ? sd.ffe[it.ikind * nq + m] : 0.0;
            wq[t] = sd.ffw ? sd.ffw[m] : 0.0;
            pAA[t] = pEE[t] = pHH[t] = pAE[t] = pAH[t] = pEH[t] = 0.0;
        }

        int j = it.j0;
        int c = 0;
        while (j < it.j1) {
            while (sd.cls_start[c + 1] <= j) ++c;
            const int jend = min(min(j + sd.block, it.j1), sd.cls_start[c + 1]);
            const int lk = sd.cls_kind[c];
            const bool wet = sd.cls_wet[c] != 0;
            const bool mask = j <= it.i0 + it.icount - 1;  // chunk meets the diagonal

            // FP64 carry: wave sums of S = sum s, Si = sum w_i s, Sw = sum w_j s,
            // Siw = sum w_i w_j s for the lane's q-points.
            T r0[M], ri[M], rw[M], riw[M];
#pragma unroll
            for (int t = 0; t < M; ++t) ri[t] = rw[t] = riw[t] = T(0);
            if constexpr (ONEACC) {
                T acc0[Q], unused[Q];
#pragma unroll
                for (int k = 0; k < Q; ++k) acc0[k] = T(0);
#define PRIPPS_PB1(M, S, B) pair_block<T, Q, UNIFORM, false, M, S, B>(atoms, j, jend, ai, iidx, lane_take, qv, ql, q0p, q0lo, hp, hlo, small_h, acc0, unused)
#define PRIPPS_PASS(B)                                                                   \
                if constexpr (UNIFORM && sizeof(T) == 4 && VARIANT == 4) {              \
                    if (mask) { if (small_h) PRIPPS_PB1(true, 1, B); else PRIPPS_PB1(true, 0, B); } \
                    else { if (small_h) PRIPPS_PB1(false, 1, B); else PRIPPS_PB1(false, 0, B); }    \
                } else {                                                                 \
                    if (mask) PRIPPS_PB1(true, -1, B);                                   \
                    else PRIPPS_PB1(false, -1, B);                                       \
                }
                PRIPPS_PASS(false)
                reduce_scatter<T, Q, false>(acc0, T(1), r0);
                if (wet_i) reduce_scatter<T, Q, true>(acc0, wi, ri);
                if (wet) {
#pragma unroll
                    for (int k = 0; k < Q; ++k) acc0[k] = T(0);
                    PRIPPS_PASS(true)
                    reduce_scatter<T, Q, false>(acc0, T(1), rw);
                    if (wet_i) reduce_scatter<T, Q, true>(acc0, wi, riw);
                }
#undef PRIPPS_PASS
#undef PRIPPS_PB1
            } else {
            T acc0[Q], accw[Q];
#pragma unroll
            for (int k = 0; k < Q; ++k) acc0[k] = accw[k] = T(0);

            if constexpr (UNIFORM && sizeof(T) == 4 && (VARIANT == 1 || VARIANT == 2)) {
                if (wet) {
                    if (small_h) pair_block_f32x2<Q, true, true, INTERLEAVE>(atoms, j, jend, ai, iidx, q0p, q0lo, hp, hlo, acc0, accw);
                    else pair_block_f32x2<Q, true, false, INTERLEAVE>(atoms, j, jend, ai, iidx, q0p, q0lo, hp, hlo, acc0, accw);
                } else {
                    if (small_h) pair_block_f32x2<Q, false, true, INTERLEAVE>(atoms, j, jend, ai, iidx, q0p, q0lo, hp, hlo, acc0, accw);
                    else pair_block_f32x2<Q, false, false, INTERLEAVE>(atoms, j, jend, ai, iidx, q0p, q0lo, hp, hlo, acc0, accw);
                }
            } else if constexpr (UNIFORM && sizeof(T) == 4 && VARIANT == 4) {
#define PRIPPS_PB(W, M, S) pair_block<T, Q, UNIFORM, W, M, S>(atoms, j, jend, ai, iidx, lane_take, qv, ql, q0p, q0lo, hp, hlo, small_h, acc0, accw)
                if (wet) {
                    if (mask) { if (small_h) PRIPPS_PB(true, true, 1); else PRIPPS_PB(true, true, 0); }
                    else { if (small_h) PRIPPS_PB(true, false, 1); else PRIPPS_PB(true, false, 0); }
                } else {
                    if (mask) { if (small_h) PRIPPS_PB(false, true, 1); else PRIPPS_PB(false, true, 0); }
                    else { if (small_h) PRIPPS_PB(false, false, 1); else PRIPPS_PB(false, false, 0); }
                }
            } else if constexpr (!UNIFORM && sizeof(T) == 4) {
#define PRIPPS_PBN(W, M, D) pair_block<T, Q, UNIFORM, W, M, -1, false, D>(atoms, j, jend, ai, iidx, lane_take, qv, ql, q0p, q0lo, hp, hlo, small_h, acc0, accw)
                if (sin_direct) {
                    if (wet) { if (mask) PRIPPS_PBN(true, true, true); else PRIPPS_PBN(true, false, true); }
                    else { if (mask) PRIPPS_PBN(false, true, true); else PRIPPS_PBN(false, false, true); }
                } else {
                    if (wet) { if (mask) PRIPPS_PBN(true, true, false); else PRIPPS_PBN(true, false, false); }
                    else { if (mask) PRIPPS_PBN(false, true, false); else PRIPPS_PBN(false, false, false); }
                }
#undef PRIPPS_PBN
            } else {
                if (wet) {
                    if (mask) PRIPPS_PB(true, true, -1);
                    else PRIPPS_PB(true, false, -1);
                } else {
                    if (mask) PRIPPS_PB(false, true, -1);
                    else PRIPPS_PB(false, false, -1);
                }
#undef PRIPPS_PB
            }

            reduce_scatter<T, Q, false>(acc0, T(1), r0);
            if (wet_i) reduce_scatter<T, Q, true>(acc0, wi, ri);
            if (wet) {
                reduce_scatter<T, Q, false>(accw, T(1), rw);
                if (wet_i) reduce_scatter<T, Q, true>(accw, wi, riw);
            }
            }
#pragma unroll
            for (int t = 0; t < M; ++t) {
                const int m = min(qbase + split_index<Q>(lane, t), nq - 1);
                const double al = sd.ffa[lk * nq + m];
                const double el = sd.ffe ? sd.ffe[lk * nq + m] : 0.0;
                const double S = double(r0[t]), Si = double(ri[t]);
                const double Sw = double(rw[t]), Siw = double(riw[t]);
                pAA[t] += 2.0 * ak[t] * al * S;
                pEE[t] += 2.0 * ek[t] * el * S;
                pAE[t] += (ak[t] * el + al * ek[t]) * S;
                pHH[t] += 2.0 * wq[t] * wq[t] * Siw;
                pAH[t] += wq[t] * (ak[t] * Sw + al * Si);
                pEH[t] += wq[t] * (ek[t] * Sw + el * Si);
            }
            j = jend;
        }

        double *out = args.partials + (int64_t)w * (kProfiles * Q);
        if (owner) {
#pragma unroll
            for (int t = 0; t < M; ++t) {
                const int ql = split_index<Q>(lane, t);
                out[0 * Q + ql] = pAA[t];
                out[1 * Q + ql] = pEE[t];
                out[2 * Q + ql] = pHH[t];
                out[3 * Q + ql] = pAE[t];
                out[4 * Q + ql] = pAH[t];
                out[5 * Q + ql] = pEH[t];
            }
        }
    }
}

template <typename T, int Q, bool UNIFORM, int VARIANT, bool ONEACC = false>
__global__ void __launch_bounds__(256) debye_tiled_kernel(TiledArgs args) {
    tiled_body<T, Q, UNIFORM, VARIANT, ONEACC>(args);
}

// FP64 kernels, capped at two waves per SIMD.  Left to itself the compiler
// gives them 258-288 registers, one wave per SIMD; with the cap the pair loops
// stay spill-free (the few spilled values live in the per-item set-up and the
// carry).  Shapes: Q = 64 on uniform grids with one accumulator array (64 FP64
// accumulators, one seed per pair per 64 q-points), and Q = 32 with two
// arrays on uniform and non-uniform grids.
template <int Q, bool UNIFORM, bool ONEACC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
debye_tiled_f64_kernel(TiledArgs args) {
    tiled_body<double, Q, UNIFORM, 3, ONEACC>(args);
}

// ---------------------------------------------------------------------------
// FP32, uniform grid, 128 q points per item as two 64-point halves.
//   dry j class: per pair one seed starts a 128-step recurrence whose first
//     and second 64 steps go to the two halves' accumulators;
//   hydrated j class: the halves run one after the other (acc0 = a, accw = b),
//     with the 64-point pair block.
// The first half of a dry pair and both halves of a hydrated pair see the
// same seeds, recurrence and carry as the 64-point kernel; the second half of
// a dry pair continues the recurrence for 64 more steps instead of starting
// from a fresh seed.  The FP64 profile sums live in LDS (6 KB per wave) and
// go to the item's partial slot once at the end: in registers they would push
// the kernel past 256 VGPRs, to one wave per SIMD, and in the partial slot
// their read-modify-write per chunk cost HBM traffic (288 against 85 MB per
// launch on a 10,000-atom structure).
// ---------------------------------------------------------------------------
// TWO = false: the strip's second half lies beyond the grid's last q-point
// (the last strip of a grid with nq mod 128 in (0, 64]); only the first 64
// steps are run.
template <bool MASK, bool TWO>
__device__ __forceinline__ void dual_block_dry(const Atom<float> *__restrict__ atoms, int jb, int je,
                                               const Atom<float> &ai, int iidx, float lane_take,
                                               const float (&q0p)[2], const float (&q0lo)[2],
                                               float hp, float hlo, bool small_h,
                                               float (&acc_a)[64], float (&acc_b)[64]) {
    const PRIPPS_CONST Atom<float> *ca = as_const(atoms);
    Atom<float> next = load_atom(ca, jb);
    for (int j = jb; j < je; ++j) {
        const Atom<float> aj = next;
        next = load_atom(ca, min(j + 1, je - 1));  // prefetch the next j atom
        const float dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
        const float r2 = dx * dx + dy * dy + dz * dz;
        const float take = MASK ? (j > iidx ? 1.0f : 0.0f) : lane_take;
        const PairGeom32 g = pair_geom_f32(r2, hp, hlo, take, small_h);
        // one seed per pair: the second half continues the first half's
        // recurrence (its state after 64 steps is the second half's seed)
        const Seed32 z = seed_from_geom(g, q0p[0], q0lo[0]);
        float sv = z.s, d = z.d;
#pragma unroll
        for (int k = 0; k < 64; ++k) {
            acc_a[k] += sv;
            sv += d;
            d -= z.lam * sv;
        }
        if constexpr (TWO) {
#pragma unroll
            for (int k = 0; k < 64; ++k) {
                acc_b[k] += sv;
                sv += d;
                d -= z.lam * sv;
            }
        }
    }
}

__global__ void __launch_bounds__(256) debye_dual_kernel(TiledArgs args) {
    constexpr int Q = 128;
    const int lane = __lane_id();
    // the wave's FP64 profile sums: [profile][q-local], each lane only ever
    // touches its own two columns (lane, 64 + lane), so no barrier is needed
    __shared__ double sums[4][kProfiles * Q];
    double *const lsum = sums[threadIdx.x / 64];
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
        Atom<float> ai = atoms[ivalid ? it.i0 + lane : it.i0];
        const float wi = ivalid ? ai.w : 0.0f;
        const float lane_take = ivalid ? 1.0f : 0.0f;
        const bool wet_i = __ballot(wi != 0.0f) != 0;
        const bool small_h = sd.small_h != 0;

        const int qbase = it.strip * Q;
        // the second 64-point half holds grid points (else it is left at zero)
        const bool half2 = qbase + 64 < nq;
        float q0p[2], q0lo[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const double qs = sd.q0 + double(qbase + 64 * h) * sd.dq;
            q0p[h] = float(qs * kInv2Pi);
            q0lo[h] = float(qs * kInv2Pi - double(q0p[h]));
        }
        const double hd = sd.dq * (0.5 * kInv2Pi);
        const float hp = float(hd);
        const float hlo = float(hd - double(hp));
        float qv_unused[64], ql_unused[64];
#pragma unroll
        for (int k = 0; k < 64; ++k) qv_unused[k] = ql_unused[k] = 0.0f;

        // Lane owns q = qbase + 64 h + lane of each half h.  The FP64 profile
        // accumulators live in LDS (6 KB per wave): a carry adds to them once
        // per chunk, which keeps them and the form factors out of the register
        // file, and the item's partial slot is written once at the end.
        double *out = args.partials + (int64_t)w * (kProfiles * Q);
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int p = 0; p < kProfiles; ++p) lsum[p * Q + 64 * h + lane] = 0.0;

        auto carry = [&](int h, int lk, const float (&a0)[64], const float (&aw)[64], bool wet) {
            float r0[1], ri[1], rw[1], riw[1];
            reduce_scatter<float, 64, false>(a0, 1.0f, r0);
            ri[0] = rw[0] = riw[0] = 0.0f;
            if (wet_i) reduce_scatter<float, 64, true>(a0, wi, ri);
            if (wet) {
                reduce_scatter<float, 64, false>(aw, 1.0f, rw);
                if (wet_i) reduce_scatter<float, 64, true>(aw, wi, riw);
            }
            const int m = min(qbase + 64 * h + lane, nq - 1);
            const double ak = sd.ffa[it.ikind * nq + m];
            const double ek = sd.ffe ? sd.ffe[it.ikind * nq + m] : 0.0;
            const double wq = sd.ffw ? sd.ffw[m] : 0.0;
            const double al = sd.ffa[lk * nq + m];
            const double el = sd.ffe ? sd.ffe[lk * nq + m] : 0.0;
            const double S = double(r0[0]), Si = double(ri[0]);
            const double Sw = double(rw[0]), Siw = double(riw[0]);
            double *o = lsum + 64 * h + lane;
            o[0 * Q] += 2.0 * ak * al * S;
            o[1 * Q] += 2.0 * ek * el * S;
            o[2 * Q] += 2.0 * wq * wq * Siw;
            o[3 * Q] += (ak * el + al * ek) * S;
            o[4 * Q] += wq * (ak * Sw + al * Si);
            o[5 * Q] += wq * (ek * Sw + el * Si);
        };

        int j = it.j0;
        int c = 0;
        while (j < it.j1) {
            while (sd.cls_start[c + 1] <= j) ++c;
            const int jend = min(min(j + sd.block, it.j1), sd.cls_start[c + 1]);
            const int lk = sd.cls_kind[c];
            const bool wet = sd.cls_wet[c] != 0;
            const bool mask = j <= it.i0 + it.icount - 1;  // chunk meets the diagonal

            float acc_a[64], acc_b[64];
            if (!wet) {
#pragma unroll
                for (int k = 0; k < 64; ++k) acc_a[k] = acc_b[k] = 0.0f;
                if (half2) {
                    if (mask) dual_block_dry<true, true>(atoms, j, jend, ai, iidx, lane_take, q0p, q0lo, hp, hlo, small_h, acc_a, acc_b);
                    else dual_block_dry<false, true>(atoms, j, jend, ai, iidx, lane_take, q0p, q0lo, hp, hlo, small_h, acc_a, acc_b);
                } else {
                    if (mask) dual_block_dry<true, false>(atoms, j, jend, ai, iidx, lane_take, q0p, q0lo, hp, hlo, small_h, acc_a, acc_b);
                    else dual_block_dry<false, false>(atoms, j, jend, ai, iidx, lane_take, q0p, q0lo, hp, hlo, small_h, acc_a, acc_b);
                }
                carry(0, lk, acc_a, acc_b, false);
                if (half2) carry(1, lk, acc_b, acc_a, false);
            } else {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (h == 1 && !half2) break;
#pragma unroll
                    for (int k = 0; k < 64; ++k) acc_a[k] = acc_b[k] = 0.0f;
                    if (mask)
                        pair_block<float, 64, true, true, true, -1>(atoms, j, jend, ai, iidx, lane_take, qv_unused, ql_unused, q0p[h], q0lo[h], hp, hlo, small_h, acc_a, acc_b);
                    else
                        pair_block<float, 64, true, true, false, -1>(atoms, j, jend, ai, iidx, lane_take, qv_unused, ql_unused, q0p[h], q0lo[h], hp, hlo, small_h, acc_a, acc_b);
                    carry(h, lk, acc_a, acc_b, true);
                }
            }
            j = jend;
        }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int p = 0; p < kProfiles; ++p) out[p * Q + 64 * h + lane] = lsum[p * Q + 64 * h + lane];
    }
}

// ---------------------------------------------------------------------------
// finalize: sum item partials per (structure, strip) in a fixed order, divide
// by q, add the diagonal; q = 0 takes the closed form (sum f_X)(sum f_Y).
// ---------------------------------------------------------------------------

// Stage 1: block (strip g, chunk c) sums the partials of items
// csr[g][c*kFinChunk .. (c+1)*kFinChunk) into stage[g][c][6][Q].  Thread t
// takes q-point t % Q and every G-th item of the chunk (G = 256 / Q); the G
// group sums are combined in LDS in group order, so the result is fixed by
// the plan alone.
__global__ void __launch_bounds__(256) debye_reduce_kernel(FinalizeArgs args) {
    __shared__ double red[256][kProfiles];
    // One block per chunk slot of `stage`: the strip is the last one whose
    // first slot is <= blockIdx.x (strips without items own no slot).
    const int slot = blockIdx.x;
    int lo = 0, hi = args.n_strips;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (args.stage_ptr[mid] <= slot) lo = mid;
        else hi = mid;
    }
    const int g_strip = lo;
    const int chunk = slot - args.stage_ptr[g_strip];
    const int Q = args.Q;
    const int G = 256 / Q;
    const int ql = threadIdx.x % Q;
    const int g = threadIdx.x / Q;
    const int b = args.csr_ptr[g_strip] + chunk * kFinChunk;
    const int e = min(args.csr_ptr[g_strip + 1], b + kFinChunk);
    double acc[kProfiles] = {0, 0, 0, 0, 0, 0};
    if (g < G) {
#pragma unroll 4
        for (int idx = b + g; idx < e; idx += G) {
            const double *p = args.partials + (int64_t)args.csr_items[idx] * (kProfiles * Q);
#pragma unroll
            for (int k = 0; k < kProfiles; ++k) acc[k] += p[k * Q + ql];
        }
    }
#pragma unroll
    for (int k = 0; k < kProfiles; ++k) red[threadIdx.x][k] = acc[k];
    __syncthreads();
    if (g == 0) {
        for (int gg = 1; gg < G; ++gg)
#pragma unroll
            for (int k = 0; k < kProfiles; ++k) acc[k] += red[gg * Q + ql][k];
        double *o = args.stage + ((int64_t)args.stage_ptr[g_strip] + chunk) * (kProfiles * Q);
#pragma unroll
        for (int k = 0; k < kProfiles; ++k) o[k * Q + ql] = acc[k];
    }
}

// Stage 2: one block per strip.  The block's kFinThreads threads form
// kFinThreads / Q groups of Q; group g sums the strip's chunk totals
// g, g + G, ... in chunk order, and the group sums are added in group order
// (the result is a function of the plan alone).  Then the q-point's thread
// divides by q and adds the diagonal (q = 0: closed form).
constexpr int kFinThreads = 1024;

__global__ void __launch_bounds__(kFinThreads) debye_finalize_kernel(FinalizeArgs args) {
    __shared__ double red[kFinThreads][kProfiles];
    const int g_strip = blockIdx.x;
    const int Q = args.Q;
    const int G = kFinThreads / Q;
    const int ql = threadIdx.x % Q, g = threadIdx.x / Q;
    const int s = args.strip_struct[g_strip];
    const int strip = args.strip_index[g_strip];
    const PRIPPS_CONST DevStruct &sd = as_const(args.structs)[s];
    // fixed-width strips, or the strip table of a Chebyshev plan
    const int m = (sd.strip_first ? sd.strip_first[strip] : strip * Q) + ql;
    const int mend = sd.strip_first ? sd.strip_first[strip + 1] : sd.nq;
    const int nchunks = args.stage_ptr[g_strip + 1] - args.stage_ptr[g_strip];
    double acc[kProfiles] = {0, 0, 0, 0, 0, 0};
    if (m < mend) {
#pragma unroll 4
        for (int c = g; c < nchunks; c += G) {
            const double *p = args.stage + ((int64_t)args.stage_ptr[g_strip] + c) * (kProfiles * Q);
#pragma unroll
            for (int k = 0; k < kProfiles; ++k) acc[k] += p[k * Q + ql];
        }
    }
#pragma unroll
    for (int k = 0; k < kProfiles; ++k) red[threadIdx.x][k] = acc[k];
    __syncthreads();
    if (g != 0 || m >= mend) return;
    for (int gg = 1; gg < G; ++gg)
#pragma unroll
        for (int k = 0; k < kProfiles; ++k) acc[k] += red[gg * Q + ql][k];
    double diag[kProfiles], zero[kProfiles];
    diagonal_and_zero(sd, m, diag, zero);
    const double q = sd.q[m];
#pragma unroll
    for (int k = 0; k < kProfiles; ++k) sd.profiles[k * sd.nq + m] = q > 0.0 ? acc[k] / q + diag[k] : zero[k];
}

// ---------------------------------------------------------------------------
// plain FP64 path: one sin per pair and q, one wave per (i-tile, q)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) debye_plain_kernel(PlainArgs args) {
    const PRIPPS_CONST DevStruct &sd = *as_const(args.sd);
    const int lane = __lane_id();
    const int m = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int tile = blockIdx.x;
    if (m >= sd.nq) return;  // whole wave leaves; no block-level barrier below
    const int4 tl = args.tiles[tile];
    const Atom<double> *__restrict__ atoms = static_cast<const Atom<double> *>(sd.atoms);
    const bool ivalid = lane < tl.y;
    const int iidx = ivalid ? tl.x + lane : 0x7fffffff;
    const Atom<double> ai = atoms[ivalid ? tl.x + lane : tl.x];
    const double wi = ivalid ? ai.w : 0.0;
    const int nq = sd.nq;
    const double q = sd.q[m];
    const double ak = sd.ffa[tl.z * nq + m];
    const double ek = sd.ffe ? sd.ffe[tl.z * nq + m] : 0.0;
    const double W = sd.ffw ? sd.ffw[m] : 0.0;
    double pAA = 0, pEE = 0, pHH = 0, pAE = 0, pAH = 0, pEH = 0;
    int c = 0;
    for (int j = tl.x; j < sd.n;) {
        while (sd.cls_start[c + 1] <= j) ++c;
        const int jend = sd.cls_start[c + 1];
        const int lk = sd.cls_kind[c];
        const double al = sd.ffa[lk * nq + m];
        const double el = sd.ffe ? sd.ffe[lk * nq + m] : 0.0;
        double T0 = 0.0, Tw = 0.0;
        for (int jj = j; jj < jend; ++jj) {
            const Atom<double> aj = load_atom(as_const(atoms), jj);
            const double dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double x = q * r;
            const double sc = fabs(x) < 1e-8 ? 1.0 : sin(x) / x;
            const double s = jj > iidx ? sc : 0.0;
            T0 += s;
            Tw += aj.w * s;
        }
        pAA += 2.0 * ak * al * T0;
        pEE += 2.0 * ek * el * T0;
        pAE += (ak * el + al * ek) * T0;
        pHH += 2.0 * W * W * wi * Tw;
        pAH += W * (ak * Tw + al * wi * T0);
        pEH += W * (ek * Tw + el * wi * T0);
        j = jend;
    }
    pAA = wave_sum(pAA);
    pEE = wave_sum(pEE);
    pHH = wave_sum(pHH);
    pAE = wave_sum(pAE);
    pAH = wave_sum(pAH);
    pEH = wave_sum(pEH);
    if (lane == 0) {
        double *o = args.partials + ((int64_t)tile * nq + m) * kProfiles;
        o[0] = pAA;
        o[1] = pEE;
        o[2] = pHH;
        o[3] = pAE;
        o[4] = pAH;
        o[5] = pEH;
    }
}

__global__ void __launch_bounds__(256) debye_plain_finalize_kernel(PlainArgs args) {
    const PRIPPS_CONST DevStruct &sd = *as_const(args.sd);
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= sd.nq) return;
    double acc[kProfiles] = {0, 0, 0, 0, 0, 0};
    for (int t = 0; t < args.n_tiles; ++t) {
        const double *p = args.partials + ((int64_t)t * sd.nq + m) * kProfiles;
#pragma unroll
        for (int k = 0; k < kProfiles; ++k) acc[k] += p[k];
    }
    double diag[kProfiles], zero[kProfiles];
    diagonal_and_zero(sd, m, diag, zero);
#pragma unroll
    for (int k = 0; k < kProfiles; ++k) sd.profiles[k * sd.nq + m] = acc[k] + diag[k];
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------
__global__ void fill_int_kernel(int *p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void fill_double_kernel(double *p, int64_t n, double v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Weighted sum of the profiles of a chunk of trajectory frames into the
// running FP64 accumulator, slots in a fixed order (deterministic).
__global__ void frames_accumulate_kernel(const double *__restrict__ prof, int n_slots,
                                         const double *__restrict__ w, int64_t n,
                                         double *__restrict__ acc) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = acc[i];
    for (int c = 0; c < n_slots; ++c) a += w[c] * prof[(int64_t)c * n + i];
    acc[i] = a;
}

__global__ void frames_normalize_kernel(const double *__restrict__ acc, int64_t n, double wsum,
                                        double *__restrict__ out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = acc[i] / wsum;
}

// I(q) = AA + vs^2 EE + cd^2 HH - 2 vs AE + 2 cd AH - 2 vs cd EH, evaluated in
// the reference's operation order without FMA contraction
// (src/debye.rs:186-195).  out: [4][nq] total, atoms, excluded, hydration.
__global__ void combine_kernel(const double *__restrict__ prof, int nq, double vs, double cd,
                               double *__restrict__ out) {
#pragma clang fp contract(off)
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= nq) return;
    const double aa = prof[0 * nq + m], ee = prof[1 * nq + m], hh = prof[2 * nq + m];
    const double ae = prof[3 * nq + m], ah = prof[4 * nq + m], eh = prof[5 * nq + m];
    const double t1 = aa + vs * vs * ee;
    const double t2 = t1 + cd * cd * hh;
    const double t3 = t2 - 2.0 * vs * ae;
    const double t4 = t3 + 2.0 * cd * ah;
    const double total = t4 - 2.0 * vs * cd * eh;
    out[0 * nq + m] = total;
    out[1 * nq + m] = aa;
    out[2 * nq + m] = ee;
    out[3 * nq + m] = hh;
}

// ---------------------------------------------------------------------------
// launch wrappers (host side of this translation unit)
// ---------------------------------------------------------------------------
// The one-accumulator kernel where it saves a wave's worth of registers: FP32
// on uniform grids, variants 3 and 4 (the non-uniform FP32 kernel and FP64 at
// Q = 32 stay at two waves per SIMD either way, so they keep two arrays).
// FP64 has its own entry points, debye_tiled_f64_kernel.
template <typename T, int Q, bool U, int V>
constexpr bool kOneAcc = sizeof(T) == 4 && U && V != 1 && V != 2;

template <typename T, int Q, bool U, int V = 3>
static hipError_t launch_tiled_t(const TiledArgs &a, int blocks, hipStream_t st) {
    constexpr bool ONE = kOneAcc<T, Q, U, V>;
    hipLaunchKernelGGL((debye_tiled_kernel<T, Q, U, V, ONE>), dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_tiled(bool fp64, int Q, bool uniform, const TiledArgs &a, int blocks,
                        hipStream_t st) {
    // Instantiated shapes: those whose accumulators fit the 256 VGPRs of a wave.
    if (!fp64) {
        if (uniform && a.variant == 1 && Q == 64) return launch_tiled_t<float, 64, true, 1>(a, blocks, st);
        if (uniform && a.variant == 2 && Q == 64) return launch_tiled_t<float, 64, true, 2>(a, blocks, st);
        if (uniform && a.variant == 4 && Q == 64) return launch_tiled_t<float, 64, true, 4>(a, blocks, st);
        if (uniform && a.variant == 4 && Q == 32) return launch_tiled_t<float, 32, true, 4>(a, blocks, st);
        if (Q == 32) return uniform ? launch_tiled_t<float, 32, true>(a, blocks, st)
                                    : launch_tiled_t<float, 32, false>(a, blocks, st);
        if (Q == 64) return uniform ? launch_tiled_t<float, 64, true>(a, blocks, st)
                                    : launch_tiled_t<float, 64, false>(a, blocks, st);
        if (Q == 128 && uniform) {
            hipLaunchKernelGGL(debye_dual_kernel, dim3(blocks), dim3(256), 0, st, a);
            return hipGetLastError();
        }
    } else {
        if (Q == 32 && uniform)
            hipLaunchKernelGGL((debye_tiled_f64_kernel<32, true, false>), dim3(blocks), dim3(256), 0, st, a);
        else if (Q == 32)
            hipLaunchKernelGGL((debye_tiled_f64_kernel<32, false, false>), dim3(blocks), dim3(256), 0, st, a);
        else if (Q == 64 && uniform)
            hipLaunchKernelGGL((debye_tiled_f64_kernel<64, true, true>), dim3(blocks), dim3(256), 0, st, a);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    }
    return hipErrorInvalidValue;
}

template <typename T, int Q, bool U, int V = 3>
static hipError_t occupancy_t(int *blocks_per_cu) {
    constexpr bool ONE = kOneAcc<T, Q, U, V>;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, reinterpret_cast<const void *>(&debye_tiled_kernel<T, Q, U, V, ONE>), 256, 0);
}

hipError_t tiled_occupancy(bool fp64, int Q, bool uniform, int variant, int *blocks_per_cu) {
    if (!fp64) {
        if (uniform && Q == 64 && (variant == 1 || variant == 2))
            return variant == 1 ? occupancy_t<float, 64, true, 1>(blocks_per_cu)
                                : occupancy_t<float, 64, true, 2>(blocks_per_cu);
        if (Q == 32) return uniform ? occupancy_t<float, 32, true>(blocks_per_cu) : occupancy_t<float, 32, false>(blocks_per_cu);
        if (Q == 64) return uniform ? occupancy_t<float, 64, true>(blocks_per_cu) : occupancy_t<float, 64, false>(blocks_per_cu);
        if (Q == 128 && uniform)
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(
                blocks_per_cu, reinterpret_cast<const void *>(&debye_dual_kernel), 256, 0);
    } else {
        const void *f = nullptr;
        if (Q == 32 && uniform) f = reinterpret_cast<const void *>(&debye_tiled_f64_kernel<32, true, false>);
        else if (Q == 32) f = reinterpret_cast<const void *>(&debye_tiled_f64_kernel<32, false, false>);
        else if (Q == 64 && uniform) f = reinterpret_cast<const void *>(&debye_tiled_f64_kernel<64, true, true>);
        if (f) return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, f, 256, 0);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_finalize(const FinalizeArgs &a, int n_strips_total, hipStream_t st) {
    if (a.total_chunks > 0) {
        hipLaunchKernelGGL(debye_reduce_kernel, dim3(a.total_chunks), dim3(256), 0, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(debye_finalize_kernel, dim3(n_strips_total), dim3(kFinThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_plain(const PlainArgs &a, int nq, hipStream_t st) {
    if (a.n_tiles > 0) {
        hipLaunchKernelGGL(debye_plain_kernel, dim3(a.n_tiles, (nq + 3) / 4), dim3(256), 0, st, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(debye_plain_finalize_kernel, dim3((nq + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_fill_int(int *p, int n, int v, hipStream_t st) {
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_combine(const double *prof, int nq, double vs, double cd, double *out,
                          hipStream_t st) {
    hipLaunchKernelGGL(combine_kernel, dim3((nq + 255) / 256), dim3(256), 0, st, prof, nq, vs, cd,
                       out);
    return hipGetLastError();
}

hipError_t launch_fill_double(double *p, int64_t n, double v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(fill_double_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_frames_accumulate(const double *prof, int n_slots, const double *w, int nq,
                                    double *acc, hipStream_t st) {
    const int64_t n = (int64_t)kProfiles * nq;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(frames_accumulate_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       prof, n_slots, w, n, acc);
    return hipGetLastError();
}

hipError_t launch_frames_normalize(const double *acc, int nq, double wsum, double *out,
                                   hipStream_t st) {
    const int64_t n = (int64_t)kProfiles * nq;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(frames_normalize_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       acc, n, wsum, out);
    return hipGetLastError();
}

}  // namespace pripps
