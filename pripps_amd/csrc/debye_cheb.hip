// FP32 pair kernel with a Chebyshev expansion in q per pair (gfx950).
//
// A strip is a run of up to kChebQ consecutive q-points of one structure, on
// any grid, with centre q_c and half-width H (x = (q - q_c) / H in [-1, 1]).
// For a pair at distance r, sin(q r) / r over the strip is, by the
// Jacobi-Anger expansion (a = q_c r, beta = H r),
//   sin(a + beta x) / r = (1/r) [sin a (J_0(beta) + 2 sum_m (-1)^m J_2m(beta) T_2m(x))
//                               + cos a (2 sum_m (-1)^m J_2m+1(beta) T_2m+1(x))],
// a Chebyshev series whose terms beyond order P are below 1e-8 of the pair's
// amplitude (the host picks P from beta_max = H r_max with the bound
// |J_p(beta)| <= (beta/2)^p / p!).
//
// Work item: a 64-atom i-tile (one lane per i atom) against a j-span, for a
// group of one or two strips that share H; a wave takes a unit (a run of
// items of one group) and sums it into the unit's coefficient records.
// Sharing H shares beta, so the pair's Bessel weights -- most of the per-pair
// work -- serve both strips; only sin a, cos a and the P accumulations are
// per strip.  Each lane adds the coefficients of its pairs in FP32 over a
// chunk of up to `block` j atoms of one class; then the wave reduce-scatters
// them (lane f receives coefficient f summed over the 64 lanes) and adds them
// to FP64 sums in LDS, which go to a record when the j class changes and at
// the end of the item.  The reduction kernels below sum the records per
// (i kind, j kind), evaluate the series at the q-points and multiply in the
// form factors.
//
// Bessel weights per pair: Miller's backward recurrence.  With z = (beta/2)^2
// and u_p = J_p(beta) (2 / beta)^p the recurrence reads
// u_{p-1} = p u_p - z u_{p+1}; the kernel runs it on w_p = 2 (-z)^floor(p/2) u_p,
// which are the series weights themselves (2 (-1)^m J_2m for even p = 2m,
// 2 (-1)^m J_2m+1 (2 / beta) for odd p = 2m + 1, times one scale per pair):
//   w_{p-1} = (2m + 1) w_p + w_{p+1}   (p = 2m + 1 odd),
//   w_{p-1} = (-p / z) w_p + w_{p+1}   (p even),
// one FMA per order and one multiply per two.  Started at N = P + 2 from
// (w_{N+1}, w_N) = (0, s), it has the scale of s; Neumann's sum
// 1 = J_0 + 2 sum_m J_2m, in w: S = w_0 / 2 + sum_{m >= 1} (-1)^m w_2m, removes
// it.  The unwanted solution that the arbitrary start admits shrinks against
// the wanted one by about z / p^2 per order on the way down; at the largest
// beta an order admits (J_P below 5e-9 of the amplitude) z / P^2 is at most
// 0.06, so two orders leave it under 4e-3 of J_P at order P, below 2e-11 of
// the pair's amplitude, and smaller still at every lower order.
//
// Range: w_0 / w_N grows like N! (2 / beta)^N as beta falls, beyond FP32 for
// small beta.  Every |w_p| is kept within [2^-126, 2^50] -- so that the
// weights amp sin a / S below stay normal numbers -- by three starts that
// every lane runs through (ChebLimits): s = 2^-113 at order N for beta above a
// threshold t1 (the largest growth from N is 2^163 there) and 0 below it; a
// restart of 2^-126 added at order 14 (for N > 16) for beta above t2 (growth
// from order 14 below 2^176), 0 below; and a restart of 2^-126 at order 4 for
// every lane.  Below t1 the sequence starts at order 14, where J_14(t1) is
// below 1e-10 of the amplitude, and below t2 at order 4 (J_4(t2) < 1e-10).
// Where a start is on, a restart added below it changes the sequence by less
// than 2^-13 J_N / J_L, far below the FP32 rounding.  z is clamped below at a
// floor that keeps -p / z finite and changes no weight.
//
// The series coefficients are w_p times a per-strip factor: amp sin a / S for
// even p, amp b cos a / S for odd p (b = beta / 2), with w_p formed once for
// both strips; coefficient 0 is added in full and the reduction halves it.
//
// The phases use the same split constants as the recurrence kernels: q_c / 2 pi
// and H / 2 as float hi + lo parts, r and 1/r from residual-corrected square
// roots, sin a and cos a from the polynomial sincos of a phase in revolutions.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "internal.h"
#include "pair_math.h"
#include "trig_f64.h"

namespace pripps {

#define PRIPPS_LDS __attribute__((address_space(3)))
#define PRIPPS_GLOBAL __attribute__((address_space(1)))

__device__ __forceinline__ int fbits(float v) { return __builtin_bit_cast(int, v); }
__device__ __forceinline__ float bitsf(int v) { return __builtin_bit_cast(float, v); }

// v as seen in the lane that DPP_CTRL names
template <int DPP_CTRL>
__device__ __forceinline__ float dpp_from(float v) {
    return bitsf(__builtin_amdgcn_update_dpp(0, fbits(v), DPP_CTRL, 0xF, 0xF, false));
}

// One in-row level of reduce_scatter: lanes with `hi` set keep value t + H of
// each pair (t, t + H), the others value t, and each adds its DPP partner's
// copy of the value it keeps.
template <int DPP_CTRL, int H>
__device__ __forceinline__ void rs_row_level(const float *in, float *out, bool hi) {
#pragma unroll
    for (int t = 0; t < H; ++t) {
        const float keep = hi ? in[t + H] : in[t];
        const float send = hi ? in[t] : in[t + H];
        out[t] = keep + dpp_from<DPP_CTRL>(send);
    }
}

// Reduce-scatter of NV <= 64 per-lane values in(0..NV): lane L ends with the
// sum over the 64 lanes of value L (NV <= 32: lanes L and L + 32 both hold
// value L mod 32; lanes at or beyond NV hold sums of zeros).  Levels: wave
// halves (permlane32 swap), row pairs (permlane16 swap), then inside each
// 16-lane row the mirror (L ^ 15), the half-row mirror (L ^ 7), L ^ 2 and
// L ^ 1 by DPP.  At each level a lane keeps the half of its values that the
// level's lane bit selects and adds its partner's copy; the six lane masks
// are independent, so each total covers every lane once, always in the same
// order.
template <int NV, typename F>
__device__ __forceinline__ float reduce_scatter(F in) {
    static_assert(NV >= 1 && NV <= 64, "one round holds up to 64 values");
    const int lane = __lane_id();
    float x[32];
#pragma unroll
    for (int t = 0; t < 32; ++t) {
        if (NV > 32) {
            const float a = in(t);
            const float c = t + 32 < NV ? in(t + 32) : 0.0f;
            const auto s = __builtin_amdgcn_permlane32_swap(fbits(a), fbits(c), false, false);
            x[t] = bitsf(s[0]) + bitsf(s[1]);
        } else if (t < NV) {
            const float a = in(t);
            const auto s = __builtin_amdgcn_permlane32_swap(fbits(a), fbits(a), false, false);
            x[t] = bitsf(s[0]) + bitsf(s[1]);
        } else {
            x[t] = 0.0f;
        }
    }
    float y[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) {
        const auto s = __builtin_amdgcn_permlane16_swap(fbits(x[t]), fbits(x[t + 16]), false, false);
        y[t] = bitsf(s[0]) + bitsf(s[1]);
    }
    float z[8], w[4], v[2];
    rs_row_level<0x140, 8>(y, z, (lane & 8) != 0);  // row_mirror
    rs_row_level<0x141, 4>(z, w, (lane & 4) != 0);  // row_half_mirror
    rs_row_level<0x4E, 2>(w, v, (lane & 2) != 0);   // quad_perm [2,3,0,1]
    float o[1];
    rs_row_level<0xB1, 1>(v, o, (lane & 1) != 0);   // quad_perm [1,0,3,2]
    return o[0];
}

// Starts of the recurrence (see the file comment): the start at order
// N = P + 2 (2^-113), a restart at order kChebMid (2^-126, for N > 16) and a
// fixed restart 2^-126 at order kChebLow that every lane receives.
constexpr float kChebStartTop = 0x1p-113f;
constexpr float kChebStartLow = 0x1p-126f;
constexpr int kChebMid = 14;
constexpr int kChebLow = 4;
template <int P>
constexpr bool kChebThree = P + 2 > 16;

// Coefficients of one pair for K strips of one group, added to acc[k][0..P):
// see the file comment.  amp = take / r (0 for padding lanes and pairs on or
// below the diagonal).  Coefficient 0 is added in full (w_0 = 2 u_0); the
// reduction halves it.
template <int P, int K>
__device__ __forceinline__ void cheb_pair(float r, float amp, const float (&qc_hi)[K],
                                          const float (&qc_lo)[K], float hh_hi, float hh_lo,
                                          const ChebLimits &lim, float (&acc)[K][P]) {
    constexpr int N = P + 2;
    constexpr int LOW = kChebLow;
    float sa[K], ca[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const float arev = __builtin_fmaf(qc_hi[k], r, qc_lo[k] * r);
        sincos_rev(arev, sa[k], ca[k]);
    }
    const float b = __builtin_fmaf(hh_hi, r, hh_lo * r);  // beta / 2
    const float z = fmaxf(b * b, lim.z_floor);
    float rz = __builtin_amdgcn_rcpf(z);
    rz = __builtin_fmaf(rz, __builtin_fmaf(-z, rz, 1.0f), rz);
    const float s_top = __builtin_amdgcn_fmed3f(__builtin_fmaf(z, lim.k1, lim.b1), 0.0f, kChebStartTop);
    float s_mid = 0.0f;
    if constexpr (kChebThree<P>)
        s_mid = __builtin_amdgcn_fmed3f(__builtin_fmaf(z, lim.k2, lim.b2), 0.0f, kChebStartLow);
    // w_{p-1} = (-p / z) w_p + w_{p+1} (p even), p w_p + w_{p+1} (p odd), down
    // from (w_{N+1}, w_N) = (0, s_top); the restarts enter the addend.  The
    // Neumann sum runs along: E = sum_{m >= 1} (-1)^m w_2m.
    float w[P];
    float wn = 0.0f, wc = s_top;
    float E = ((N / 2) % 2 == 0) ? s_top : -s_top;
#pragma unroll
    for (int p = N; p >= 1; --p) {
        float add = wn;
        if (p - 1 == LOW) add += kChebStartLow;
        if constexpr (kChebThree<P>)
            if (p - 1 == kChebMid) add += s_mid;
        const float wm = (p % 2 == 0) ? __builtin_fmaf(float(-p) * rz, wc, add)
                                      : __builtin_fmaf(float(p), wc, add);
        if (p - 1 < P) w[p - 1] = wm;
        if ((p - 1) % 2 == 0 && p - 1 >= 2) E = ((p - 1) / 2 % 2 == 0) ? E + wm : E - wm;
        wn = wc;
        wc = wm;
    }
    // Neumann normalisation: S = u_0 + 2 sum_m u_2m z^m = w_0 / 2 + E
    const float S = __builtin_fmaf(0.5f, w[0], E);
    float inv = __builtin_amdgcn_rcpf(S);
    inv = __builtin_fmaf(inv, __builtin_fmaf(-S, inv, 1.0f), inv);
    const float a0 = amp * inv;
    const float a0b = a0 * b;
    float e[K], o[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        e[k] = a0 * sa[k];   // even orders: amp sin a / S
        o[k] = a0b * ca[k];  // odd orders: amp (beta/2) cos a / S
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int k = 0; k < K; ++k)
            acc[k][p] = __builtin_fmaf(w[p], (p % 2 == 0) ? e[k] : o[k], acc[k][p]);
}

// One chunk's pairs: j in [jb, je), all of one class, summed per lane into
// acc (cleared first).  BYW: amplitudes are multiplied by w_j (the hydrated
// pass: sum w_j s).
template <int P, int K, bool MASK, bool BYW>
__device__ __forceinline__ void cheb_chunk(const Atom<float> *__restrict__ atoms, int jb, int je,
                                           const Atom<float> &ai, int iidx, float lane_take,
                                           const float (&qc_hi)[K], const float (&qc_lo)[K],
                                           float hh_hi, float hh_lo, const ChebLimits &lim,
                                           float (&acc)[K][P]) {
    const PRIPPS_CONST Atom<float> *ca = as_const(atoms);
#pragma unroll
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int p = 0; p < P; ++p) acc[k][p] = 0.0f;
    auto pair = [&](const Atom<float> &aj, int j) {
        const float dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
        const float r2 = fmaxf(dx * dx + dy * dy + dz * dz, 1e-12f);  // r -> 0: the limit s = q
        float r, rinv;
        sqrt_and_rsqrt(r2, r, rinv);
        float take = MASK ? (j > iidx ? 1.0f : 0.0f) : lane_take;
        if constexpr (BYW) take *= to_vgpr(aj.w);
        cheb_pair<P, K>(r, take * rinv, qc_hi, qc_lo, hh_hi, hh_lo, lim, acc);
    };
    // Two pairs per iteration, each atom's scalar load issued one pair ahead
    // into its own registers.  The empty asm that reads an atom at the end of
    // the pair before its use keeps its load there: left to itself the
    // compiler issues the load at the top of the next iteration (or copies
    // the loaded registers early) and waits a full cache latency per pair.
    // The pairs are not interleaved: at three waves per SIMD two interleaved
    // pairs issue slower than one pair after the other (measured).
    Atom<float> a0 = load_atom(ca, jb);
    asm volatile("" ::"s"(a0.x), "s"(a0.y), "s"(a0.z), "s"(a0.w));  // waited for here, once
    int j = jb;
    for (; j + 1 < je; j += 2) {
        // (scheduling barriers keep each load ahead of the pair and each
        // empty asm after it, and the two pairs apart)
        const Atom<float> a1 = load_atom(ca, j + 1);
        __builtin_amdgcn_sched_barrier(0);
        pair(a0, j);
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("" ::"s"(a1.x), "s"(a1.y), "s"(a1.z), "s"(a1.w));
        a0 = load_atom(ca, min(j + 2, je - 1));
        __builtin_amdgcn_sched_barrier(0);
        pair(a1, j + 1);
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("" ::"s"(a0.x), "s"(a0.y), "s"(a0.z), "s"(a0.w));
    }
    if (j < je) pair(a0, j);
}

// acc summed over the wave into the FP64 sums of set `set` (and w_i acc into
// set + 1 when wi_on), lane-distributed: coefficient f = k P + p of round
// f / 64 lives at d[set * RS + f].  The w_i-weighted sums go first, on
// products; the plain sums last, so that their lane swaps can work on the
// accumulators in place once nothing else reads them.
template <int P, int K, int RS>
__device__ __forceinline__ void cheb_carry(const float (&acc)[K][P], PRIPPS_LDS double *d, int set,
                                           bool wi_on, float wi) {
    constexpr int NV = K * P;
    constexpr int n0 = NV < 64 ? NV : 64;
    constexpr int n1 = NV > 64 ? NV - 64 : 1;
    static_assert(NV <= 128, "two rounds at most");
    const int lane = __lane_id();
    auto flat = [&](int f) { return acc[f / P][f % P]; };
    if (wi_on) {
        float si = reduce_scatter<n0>([&](int t) { return flat(t) * wi; });
        d[(set + 1) * RS + lane] += (double)si;
        if constexpr (NV > 64) {
            si = reduce_scatter<n1>([&](int t) { return flat(64 + t) * wi; });
            d[(set + 1) * RS + 64 + lane] += (double)si;
        }
    }
    float s = reduce_scatter<n0>([&](int t) { return flat(t); });
    d[set * RS + lane] += (double)s;
    if constexpr (NV > 64) {
        s = reduce_scatter<n1>([&](int t) { return flat(64 + t); });
        d[set * RS + 64 + lane] += (double)s;
    }
}

// End of a class run: the wave's FP64 coefficient sums (every set, zero
// where unused) go to record `rec` as they lie in LDS, one coefficient per
// lane, and the LDS sums restart at zero.
template <int RS>
__device__ __forceinline__ void cheb_record(PRIPPS_LDS double *d, double *records, int rec) {
    const int lane = __lane_id();
    double *o = records + (int64_t)rec * (kChebSets * RS) + lane;
#pragma unroll
    for (int s = 0; s < kChebSets; ++s)
#pragma unroll
        for (int r = 0; r < RS; r += 64) {
            o[s * RS + r] = d[s * RS + r + lane];
            d[s * RS + r + lane] = 0.0;
        }
}

// cls_wet[c] through the 32-bit word that holds it: the scalar memory path
// has no byte loads, and a vector byte load here would make the chunk's carry
// wait for every vector memory operation still in flight (record stores, the
// queue fetch).
static_assert(offsetof(DevStruct, cls_wet) % 4 == 0, "cls_wet is read by words");
__device__ __forceinline__ bool class_wet(const PRIPPS_CONST DevStruct &sd, int c) {
    const PRIPPS_CONST int32_t *w = (const PRIPPS_CONST int32_t *)sd.cls_wet;
    return ((w[c >> 2] >> (8 * (c & 3))) & 0xff) != 0;
}

template <int P, int K, int RS>
__device__ __forceinline__ void cheb_item(const TiledArgs &args, const Item &it, PRIPPS_LDS double *d,
                                          int &rec) {
    const int lane = __lane_id();
    const PRIPPS_CONST DevStruct &sd = as_const(args.structs)[it.s];
    const Atom<float> *__restrict__ atoms = static_cast<const Atom<float> *>(sd.atoms);

    const bool ivalid = lane < it.icount;
    const int iidx = ivalid ? it.i0 + lane : 0x7fffffff;
    const Atom<float> ai = atoms[ivalid ? it.i0 + lane : it.i0];
    const float wi = ivalid ? ai.w : 0.0f;
    const float lane_take = ivalid ? 1.0f : 0.0f;
    const bool wet_i = __ballot(wi != 0.0f) != 0;

    // phases of the strips' centres and the group's half-width (shared)
    float qc_hi[K], qc_lo[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double qcr = as_const(sd.strip_geo)[2 * (it.strip + k)] * kInv2Pi;
        qc_hi[k] = to_vgpr(float(qcr));
        qc_lo[k] = to_vgpr(float(qcr - double(float(qcr))));
    }
    const double hh = 0.5 * as_const(sd.strip_geo)[2 * it.strip + 1];
    const float hh_hi = to_vgpr(float(hh));
    const float hh_lo = to_vgpr(float(hh - double(float(hh))));
    const ChebLimits lim = args.cheb;

    int j = it.j0;
    int c = 0;
    int cur = -1;
    while (j < it.j1) {
        while (sd.cls_start[c + 1] <= j) ++c;
        if (cur >= 0 && c != cur) cheb_record<RS>(d, args.partials, rec++);
        cur = c;
        const int jend = min(min(j + sd.block, it.j1), sd.cls_start[c + 1]);
        const bool wet = class_wet(sd, c);
        const bool mask = j <= it.i0 + it.icount - 1;  // chunk meets the diagonal
        // pass 0: sum s (and w_i s); pass 1 (hydrated j only): sum w_j s (and w_i w_j s)
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1 && !wet) break;
            float acc[K][P];
            if (pass == 0) {
                if (mask) cheb_chunk<P, K, true, false>(atoms, j, jend, ai, iidx, lane_take, qc_hi, qc_lo, hh_hi, hh_lo, lim, acc);
                else cheb_chunk<P, K, false, false>(atoms, j, jend, ai, iidx, lane_take, qc_hi, qc_lo, hh_hi, hh_lo, lim, acc);
            } else {
                if (mask) cheb_chunk<P, K, true, true>(atoms, j, jend, ai, iidx, lane_take, qc_hi, qc_lo, hh_hi, hh_lo, lim, acc);
                else cheb_chunk<P, K, false, true>(atoms, j, jend, ai, iidx, lane_take, qc_hi, qc_lo, hh_hi, hh_lo, lim, acc);
            }
            cheb_carry<P, K, RS>(acc, d, 2 * pass, wet_i, wi);
        }
        j = jend;
    }
    cheb_record<RS>(d, args.partials, rec++);
}

template <int P>
__global__ void __launch_bounds__(256) debye_cheb_kernel(TiledArgs args) {
    // per wave: the FP64 coefficient sums of the current class run, four sets
    // of up to two rounds of 64 (the two strips' K P coefficients)
    constexpr int RS = (2 * P + 63) / 64 * 64;
    __shared__ double coef[4][kChebSets * RS];
    PRIPPS_LDS double *const d =
        (PRIPPS_LDS double *)coef[__builtin_amdgcn_readfirstlane(threadIdx.x / 64)];
    const int lane = __lane_id();
#pragma unroll
    for (int s = 0; s < kChebSets; ++s)
#pragma unroll
        for (int r = 0; r < RS; r += 64) d[s * RS + r + lane] = 0.0;
    // Unit w of the queue: wave g of the grid takes unit g first, without
    // touching the queue (at the start every wave would otherwise wait on one
    // contended atomic), and then units n_waves + counter++ as it finishes.
    // Waves of one SIMD do not progress alike (the oldest wave issues first),
    // so the later units go to whichever wave is free.
    const int n_waves = gridDim.x * 4;
    int w = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    while (w < args.n_items) {
        const PRIPPS_CONST Unit &un = as_const(args.units)[w];
        const int i0 = un.item0, i1 = un.item1, nk = un.nk;
        int rec = un.rec0;
        for (int i = i0; i < i1; ++i) {
            const Item it = load_item(args.items, i);
            if (nk == 2) cheb_item<P, 2, RS>(args, it, d, rec);
            else cheb_item<P, 1, RS>(args, it, d, rec);
        }
        // fetched here, not ahead: at the start every wave would contend for
        // the counter at once, and every later vector-memory wait (record
        // stores, the next i tile) would wait for the atomic too
        int nx = 0;
        if (lane == 0) nx = atomicAdd(args.counter, 1);
        w = n_waves + __builtin_amdgcn_readfirstlane(__shfl(nx, 0, 64));
    }
    // The last wave out leaves the queue at zero for the next launch: every
    // other wave has taken its last unit index before it counts itself out.
    if (lane == 0 && atomicAdd(args.counter + 1, 1) == n_waves - 1) {
        atomicExch(args.counter, 0);
        atomicExch(args.counter + 1, 0);
    }
}

// ---------------------------------------------------------------------------
// FP64 form (FP64-throughout builds)
// ---------------------------------------------------------------------------
// The same expansion, strips, groups, units and coefficient records as the
// FP32 kernel, computed in FP64: orders whose truncation stays below 1e-15 of
// a pair's amplitude, the weights by the w-recurrence in FP64, and each
// lane's coefficient sums in FP64 over the whole class run (no FP32 chunks,
// no carry), summed across the wave by a butterfly when the run ends.
//
// Range: FP64 holds the growth N! (2 / beta)^N down to beta ~ 1e-6, so two
// starts suffice: 2^-600 at order N for beta above t1 (growth from N below
// 2^900, ChebLimits64), 0 below it, and a restart of 2^-1000 at order 4 for
// every lane (growth below 2^1000 down to z_floor); J_4(t1) is below 1e-24.
constexpr double kCheb64StartTop = 0x1p-600;
constexpr double kCheb64StartLow = 0x1p-1000;
constexpr int kCheb64Low = 4;

template <int P, int K>
__device__ __forceinline__ void cheb64_pair(double r, double amp, const double (&qc)[K], double hh,
                                            const ChebLimits64 &lim, const double (&kodd)[P / 2 + 1],
                                            const SinCosCoeffs &kc, double (&acc)[K][P]) {
    constexpr int N = P + 2;
    double sa[K], ca[K];
#pragma unroll
    for (int k = 0; k < K; ++k) sincos_rev_f64(qc[k] * r, sa[k], ca[k], kc);
    const double b = hh * r;  // beta / 2
    const double z = fmax(b * b, lim.z_floor);
    double rz = __builtin_amdgcn_rcp(z);
    rz = fma(rz, fma(-z, rz, 1.0), rz);
    rz = fma(rz, fma(-z, rz, 1.0), rz);
    const double s_top = fmin(fmax(fma(z, lim.k1, lim.b1), 0.0), kCheb64StartTop);
    double w[P];
    double wn = 0.0, wc = s_top;
    double E = ((N / 2) % 2 == 0) ? s_top : -s_top;
#pragma unroll
    for (int p = N; p >= 1; --p) {
        double add = wn;
        if (p - 1 == kCheb64Low) add += kCheb64StartLow;
        const double wm = (p % 2 == 0) ? fma(double(-p) * rz, wc, add)
                          : p == 1     ? wc + add
                                       : fma(kodd[(p - 1) / 2], wc, add);
        if (p - 1 < P) w[p - 1] = wm;
        if ((p - 1) % 2 == 0 && p - 1 >= 2) E = ((p - 1) / 2 % 2 == 0) ? E + wm : E - wm;
        wn = wc;
        wc = wm;
    }
    const double S = fma(0.5, w[0], E);
    double inv = __builtin_amdgcn_rcp(S);
    inv = fma(inv, fma(-S, inv, 1.0), inv);
    inv = fma(inv, fma(-S, inv, 1.0), inv);
    const double a0 = amp * inv;
    const double a0b = a0 * b;
    double e[K], o[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        e[k] = a0 * sa[k];
        o[k] = a0b * ca[k];
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k][p] = fma(w[p], (p % 2 == 0) ? e[k] : o[k], acc[k][p]);
}

// r = sqrt(r2) and 1 / r in FP64 (reciprocal square root estimate, two
// Newton steps, and the residual-corrected product)
__device__ __forceinline__ void sqrt_and_rsqrt_f64(double r2, double &r, double &rinv) {
    double y = __builtin_amdgcn_rsq(r2);
    y = fma(0.5 * y, fma(-r2 * y, y, 1.0), y);
    y = fma(0.5 * y, fma(-r2 * y, y, 1.0), y);
    const double r0 = r2 * y;
    r = fma(fma(-r0, r0, r2), 0.5 * y, r0);
    rinv = fma(y, fma(-r, y, 1.0), y);
}

// The pairs of j in [jb, je) with the lane's i atom, added to acc.  MASK:
// only j > i counts (the run meets the diagonal).  BYW: amplitudes times w_j.
template <int P, int K, bool MASK, bool BYW>
__device__ __forceinline__ void cheb64_span(const Atom<double> *__restrict__ atoms, int jb, int je,
                                            const Atom<double> &ai, int iidx, double lane_take,
                                            const double (&qc)[K], double hh, const ChebLimits64 &lim,
                                            const double (&kodd)[P / 2 + 1], const SinCosCoeffs &kc,
                                            double (&acc)[K][P]) {
    const PRIPPS_CONST Atom<double> *ca = as_const(atoms);
    for (int j = jb; j < je; ++j) {
        const Atom<double> aj = load_atom(ca, j);
        const double dx = ai.x - aj.x, dy = ai.y - aj.y, dz = ai.z - aj.z;
        const double r2 = fmax(fma(dx, dx, fma(dy, dy, dz * dz)), 1e-24);  // r -> 0: the limit s = q
        double r, rinv;
        sqrt_and_rsqrt_f64(r2, r, rinv);
        double take = MASK ? (j > iidx ? 1.0 : 0.0) : lane_take;
        if constexpr (BYW) take *= aj.w;
        cheb64_pair<P, K>(r, take * rinv, qc, hh, lim, kodd, kc, acc);
    }
}

// acc (times scale) summed over the wave into record set `set`: coefficient
// f = k P + p of the group at o[set * RS + f] (butterfly over the lanes, in a
// fixed order; lane f mod 64 writes it).
template <int P, int K, int RS>
__device__ __forceinline__ void cheb64_store(const double (&acc)[K][P], double scale, double *o,
                                             int set) {
    const int lane = __lane_id();
    double tot[RS / 64];
#pragma unroll
    for (int r = 0; r < RS / 64; ++r) tot[r] = 0.0;
#pragma unroll
    for (int f = 0; f < K * P; ++f) {
        double v = acc[f / P][f % P] * scale;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
        if (lane == f % 64) tot[f / 64] = v;
    }
#pragma unroll
    for (int r = 0; r < RS / 64; ++r) o[set * RS + r * 64 + lane] = tot[r];
}

template <int RS>
__device__ __forceinline__ void cheb64_zero(double *o, int set) {
#pragma unroll
    for (int r = 0; r < RS / 64; ++r) o[set * RS + r * 64 + __lane_id()] = 0.0;
}

template <int P, int K, int RS>
__device__ __forceinline__ void cheb64_item(const TiledArgs &args, const Item &it, int &rec,
                                            const double (&kodd)[P / 2 + 1], const SinCosCoeffs &kc) {
    const int lane = __lane_id();
    const PRIPPS_CONST DevStruct &sd = as_const(args.structs)[it.s];
    const Atom<double> *__restrict__ atoms = static_cast<const Atom<double> *>(sd.atoms);
    const bool ivalid = lane < it.icount;
    const int iidx = ivalid ? it.i0 + lane : 0x7fffffff;
    const Atom<double> ai = atoms[ivalid ? it.i0 + lane : it.i0];
    const double wi = ivalid ? ai.w : 0.0;
    const double lane_take = ivalid ? 1.0 : 0.0;
    const bool wet_i = __ballot(wi != 0.0) != 0;
    double qc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) qc[k] = as_const(sd.strip_geo)[2 * (it.strip + k)] * kInv2Pi;
    const double hh = 0.5 * as_const(sd.strip_geo)[2 * it.strip + 1];
    const ChebLimits64 lim = args.cheb64;
    const int jdiag = it.i0 + it.icount;  // runs below this meet the diagonal
    int j = it.j0;
    int c = 0;
    while (j < it.j1) {
        while (sd.cls_start[c + 1] <= j) ++c;
        const int jr = min(it.j1, sd.cls_start[c + 1]);
        const int jm = max(j, min(jr, jdiag));  // [j, jm) meets the diagonal
        const bool wet = class_wet(sd, c);
        double *o = args.partials + (int64_t)rec * (kChebSets * RS);
        // pass 0: sum s (and w_i s); pass 1 (hydrated j only): sum w_j s (and w_i w_j s)
#pragma unroll
        for (int pass = 0; pass < 2; ++pass) {
            double acc[K][P];
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
                for (int p = 0; p < P; ++p) acc[k][p] = 0.0;
            if (pass == 1 && !wet) {
                cheb64_zero<RS>(o, 2);
                cheb64_zero<RS>(o, 3);
                break;
            }
            if (pass == 0) {
                cheb64_span<P, K, true, false>(atoms, j, jm, ai, iidx, lane_take, qc, hh, lim, kodd, kc, acc);
                cheb64_span<P, K, false, false>(atoms, jm, jr, ai, iidx, lane_take, qc, hh, lim, kodd, kc, acc);
            } else {
                cheb64_span<P, K, true, true>(atoms, j, jm, ai, iidx, lane_take, qc, hh, lim, kodd, kc, acc);
                cheb64_span<P, K, false, true>(atoms, jm, jr, ai, iidx, lane_take, qc, hh, lim, kodd, kc, acc);
            }
            cheb64_store<P, K, RS>(acc, 1.0, o, 2 * pass);
            if (wet_i) cheb64_store<P, K, RS>(acc, wi, o, 2 * pass + 1);
            else cheb64_zero<RS>(o, 2 * pass + 1);
        }
        ++rec;
        j = jr;
    }
}

template <int P>
__global__ void __launch_bounds__(256) debye_cheb64_kernel(TiledArgs args) {
    constexpr int RS = (P + 63) / 64 * 64;  // one strip per group
    const int lane = __lane_id();
    const int n_waves = gridDim.x * 4;
    // The odd orders of the recurrence and the sine/cosine coefficients as
    // registers the compiler cannot see through: as 64-bit constants they
    // only fit two-address FMAs, each after a register copy of the addend.
    double kodd[P / 2 + 1];
#pragma unroll
    for (int m = 0; m <= P / 2; ++m) {
        kodd[m] = double(2 * m + 1);
        asm("" : "+s"(kodd[m]));
    }
    SinCosCoeffs kc = sincos_coeffs_f64();
#pragma unroll
    for (int i = 0; i < 9; ++i) asm("" : "+v"(kc.s[i]));
#pragma unroll
    for (int i = 0; i < 8; ++i) asm("" : "+v"(kc.c[i]));
    int w = blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    while (w < args.n_items) {
        const PRIPPS_CONST Unit &un = as_const(args.units)[w];
        const int i0 = un.item0, i1 = un.item1, nk = un.nk;
        int rec = un.rec0;
        (void)nk;  // FP64 plans keep one strip per group (see cheb_strips)
        for (int i = i0; i < i1; ++i) cheb64_item<P, 1, RS>(args, load_item(args.items, i), rec, kodd, kc);
        int nx = 0;
        if (lane == 0) nx = atomicAdd(args.counter, 1);
        w = n_waves + __builtin_amdgcn_readfirstlane(__shfl(nx, 0, 64));
    }
    if (lane == 0 && atomicAdd(args.counter + 1, 1) == n_waves - 1) {
        atomicExch(args.counter, 0);
        atomicExch(args.counter + 1, 0);
    }
}

// ---------------------------------------------------------------------------
// reduction of the coefficient records
// ---------------------------------------------------------------------------
// Stage 1: block c sums the records of chunk c (up to kFinChunk records of one
// combo) in list order; thread t takes coefficients t, t + 256, ...
__global__ void __launch_bounds__(256) cheb_chunk_kernel(ChebReduceArgs a) {
    const int c = blockIdx.x;
    const int b = a.chunk_first[c], e = a.chunk_first[c + 1];
    const int n = kChebSets * a.RS;
    static_assert(kFinChunk == 32, "one load per record of a chunk in flight");
    for (int t = threadIdx.x; t < n; t += 256) {
        // every load first, then the sum in record order
        double v[kFinChunk];
#pragma unroll
        for (int i = 0; i < kFinChunk; ++i)
            v[i] = b + i < e ? a.records[(int64_t)a.chunk_recs[b + i] * n + t] : 0.0;
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < kFinChunk; ++i)
            if (b + i < e) acc += v[i];
        a.chunk_tot[(int64_t)c * n + t] = acc;
    }
}

// sum_{c in [c0, c1)} x[c * stride] in order of c, eight loads in flight
__device__ __forceinline__ double sum_strided(const double *x, int c0, int c1, int64_t stride) {
    double acc = 0.0;
    int c = c0;
    for (; c + 8 <= c1; c += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = x[(int64_t)(c + k) * stride];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc += v[k];
    }
    for (; c < c1; ++c) acc += x[(int64_t)c * stride];
    return acc;
}

// Stage 2: block b sums its combo's chunk totals (in order) into LDS, then
// thread ql evaluates the four series at q-point m = first + ql of its strip
// (Clenshaw in FP64, coefficients k P .. k P + P - 1) and writes the six
// profile contributions of the combo's form factors.
__global__ void __launch_bounds__(kChebQ) cheb_combo_kernel(ChebReduceArgs a) {
    __shared__ double coef[kChebSets * 128];
    const int4 sa = a.sb_a[blockIdx.x], sb = a.sb_b[blockIdx.x];
    const int n = kChebSets * a.RS;
    const PRIPPS_CONST DevStruct &sd = as_const(a.structs)[sa.x];
    const int strip = sa.y, ql = threadIdx.x;
    const int first = sd.strip_first[strip], end = sd.strip_first[strip + 1];
    const int m = first + ql;
    // this q-point's grid value and form factors, loaded while the block sums
    const int nq = sd.nq, ik = sa.w, jk = sb.x;
    double qm = 0.0, ak = 0.0, al = 0.0, ek = 0.0, el = 0.0, wq = 0.0;
    if (m < end) {
        qm = sd.q[m];
        ak = sd.ffa[ik * nq + m];
        al = sd.ffa[jk * nq + m];
        if (sd.ffe) {
            ek = sd.ffe[ik * nq + m];
            el = sd.ffe[jk * nq + m];
        }
        if (sd.ffw) wq = sd.ffw[m];
    }
    for (int t = threadIdx.x; t < n; t += kChebQ)
        coef[t] = sum_strided(a.chunk_tot + t, sb.y, sb.z, n);
    __syncthreads();
    double v[kProfiles] = {0, 0, 0, 0, 0, 0};
    if (m < end) {
        const double qc = sd.strip_geo[2 * strip], H = sd.strip_geo[2 * strip + 1];
        const double x = H > 0.0 ? (qm - qc) / H : 0.0;
        // the four sets' Clenshaw recurrences side by side (independent chains)
        double s[kChebSets], b1[kChebSets], b2[kChebSets];
        const double *c = coef + sa.z * a.P;
#pragma unroll
        for (int set = 0; set < kChebSets; ++set) b1[set] = b2[set] = 0.0;
        for (int p = a.P - 1; p >= 1; --p) {
#pragma unroll
            for (int set = 0; set < kChebSets; ++set) {
                const double b0 = __builtin_fma(2.0 * x, b1[set], c[set * a.RS + p] - b2[set]);
                b2[set] = b1[set];
                b1[set] = b0;
            }
        }
#pragma unroll
        for (int set = 0; set < kChebSets; ++set)  // c_0 holds w_0 = 2 u_0
            s[set] = __builtin_fma(x, b1[set], 0.5 * c[set * a.RS] - b2[set]);
        // sets: S = sum s, Si = sum w_i s, Sw = sum w_j s, Siw = sum w_i w_j s
        v[0] = 2.0 * ak * al * s[0];
        v[1] = 2.0 * ek * el * s[0];
        v[2] = 2.0 * wq * wq * s[3];
        v[3] = (ak * el + al * ek) * s[0];
        v[4] = wq * (al * s[1]) + wq * (ak * s[2]);
        v[5] = wq * (el * s[1]) + wq * (ek * s[2]);
    }
    double *o = a.sb_out + (int64_t)blockIdx.x * (kProfiles * kChebQ) + ql;
#pragma unroll
    for (int pr = 0; pr < kProfiles; ++pr) o[pr * kChebQ] = v[pr];
}

// Stage 3: block (t, pr) adds the profile-pr contributions of strip t's
// combos in order, divides by q and adds the diagonal (q = 0: closed form).
// One block per profile: each sum is a chain of dependent loads, and six of
// them in one block took six times as long.
__global__ void __launch_bounds__(kChebQ) cheb_strip_kernel(ChebReduceArgs a) {
    const int4 st = a.st[blockIdx.x / kProfiles];
    const int pr = blockIdx.x % kProfiles;
    const PRIPPS_CONST DevStruct &sd = as_const(a.structs)[st.x];
    const int strip = st.y, ql = threadIdx.x;
    const int m = sd.strip_first[strip] + ql;
    if (m >= sd.strip_first[strip + 1]) return;
    const double acc = sum_strided(a.sb_out + pr * kChebQ + ql, st.z, st.w, kProfiles * kChebQ);
    double diag[kProfiles], zero[kProfiles];
    diagonal_and_zero(sd, m, diag, zero);
    const double q = sd.q[m];
    sd.profiles[pr * sd.nq + m] = q > 0.0 ? acc / q + diag[pr] : zero[pr];
}

hipError_t launch_cheb_reduce(const ChebReduceArgs &a, hipStream_t st) {
    if (a.n_chunks > 0) {
        hipLaunchKernelGGL(cheb_chunk_kernel, dim3(a.n_chunks), dim3(256), 0, st, a);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (a.n_sb > 0) {
        hipLaunchKernelGGL(cheb_combo_kernel, dim3(a.n_sb), dim3(kChebQ), 0, st, a);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (a.n_st > 0) {
        hipLaunchKernelGGL(cheb_strip_kernel, dim3(a.n_st * kProfiles), dim3(kChebQ), 0, st, a);
        return hipGetLastError();
    }
    return hipSuccess;
}

// Largest |w_p| over the recurrence started at order L from (w_{L+1}, w_L) =
// (0, 1) at beta (FP64; infinity when it leaves the FP64 range).
static double cheb_ratio(double beta, int L) {
    const double z = 0.25 * beta * beta;
    double w1 = 0.0, w0 = 1.0, mx = 1.0;
    for (int p = L; p >= 1; --p) {
        const double wm = (p % 2 == 0) ? (-p / z) * w0 + w1 : p * w0 + w1;
        w1 = w0;
        w0 = wm;
        mx = std::max(mx, std::fabs(wm));
        if (!std::isfinite(mx)) return INFINITY;
    }
    return mx;
}

// The smallest beta whose recurrence from order L stays within 2^bits of its
// start (the ratio falls as beta grows).
static double cheb_beta_min(int L, int bits) {
    double lo = 1e-30, hi = 64.0;
    for (int it = 0; it < 200; ++it) {
        const double mid = std::sqrt(lo * hi);
        if (cheb_ratio(mid, L) <= std::ldexp(1.0, bits)) hi = mid;
        else lo = mid;
    }
    return hi;
}

// Start values of the recurrence for order P (see ChebLimits).  Every |w_p|
// stays within [2^-126, 2^50], so that the normalised weights amp sin a / S
// stay normal FP32 numbers: the top start 2^-113 may grow by 2^163 and takes
// over from z >= z1 (its threshold, reached within 2^-20 of z1); the
// restarts 2^-126 may grow by 2^176.  A restart below a start must give the
// pair's weights for every beta under that start's threshold: the truncation
// J_L(beta) <= (beta/2)^L / L! stays below 1e-10 there.  z_floor keeps the
// coefficients -p / z finite and the lowest restart within range; it changes
// no weight (J_p(beta) = delta_p0 to below 1e-24 there).
bool cheb_limits(int P, ChebLimits *out) {
    static std::mutex mu;
    static ChebLimits cache[64];
    static bool have[64] = {};
    if (P <= 0 || P >= 62) return false;
    std::lock_guard<std::mutex> lock(mu);
    if (!have[P]) {
        const int N = P + 2;
        const bool three = N > 16;
        auto weight_bound = [](double beta, int L) {
            double t = 1.0;
            for (int k = 1; k <= L; ++k) t *= 0.5 * beta / k;
            return t;
        };
        const double t1 = cheb_beta_min(N, 163) * (1.0 + 1e-6);
        const double t2 = three ? cheb_beta_min(kChebMid, 176) * (1.0 + 1e-6) : t1;
        if (three && weight_bound(t1 * (1.0 + 1e-5), kChebMid) > 1e-10) return false;
        if (weight_bound(t2 * (1.0 + 1e-5), kChebLow) > 1e-10) return false;
        const double tf = cheb_beta_min(kChebLow, 176) * (1.0 + 1e-6);
        const double zf = std::max(0.25 * tf * tf, std::ldexp(2.0 * N, -120));
        if (0.25 * tf * tf > 1e-24) return false;
        ChebLimits L{};
        L.z_floor = (float)zf * (1.0f + 0x1p-20f);
        L.k1 = (float)(std::ldexp(1.0, -93) / (0.25 * t1 * t1));
        L.b1 = -0x1p-93f;
        if (three) {
            L.k2 = (float)(std::ldexp(1.0, -106) / (0.25 * t2 * t2));
            L.b2 = -0x1p-106f;
        }
        cache[P] = L;
        have[P] = true;
    }
    *out = cache[P];
    return true;
}

// FP64 form: start values (see ChebLimits64 and the FP64 kernel's comment).
bool cheb64_limits(int P, ChebLimits64 *out) {
    static std::mutex mu;
    static ChebLimits64 cache[64];
    static bool have[64] = {};
    if (P <= 0 || P >= 62) return false;
    std::lock_guard<std::mutex> lock(mu);
    if (!have[P]) {
        const int N = P + 2;
        const double t1 = cheb_beta_min(N, 900) * (1.0 + 1e-9);
        double wb = 1.0;  // J_4(t1) bound
        for (int k = 1; k <= kCheb64Low; ++k) wb *= 0.5 * t1 / k;
        if (wb > 1e-20) return false;
        const double tf = cheb_beta_min(kCheb64Low, 1000) * (1.0 + 1e-9);
        ChebLimits64 L{};
        L.z_floor = std::max(0.25 * tf * tf, std::ldexp(2.0 * N, -1000));
        L.k1 = std::ldexp(1.0, -580) / (0.25 * t1 * t1);
        L.b1 = -std::ldexp(1.0, -580);
        cache[P] = L;
        have[P] = true;
    }
    *out = cache[P];
    return true;
}

static bool cheb64_order_ok(int P, double beta_max) {
    const double b = 0.5 * beta_max;
    double term = 1.0;
    for (int k = 1; k <= P; ++k) term *= b / k;
    double tail = 0.0, t = term;
    for (int k = P; k < P + 300 && t > 1e-40; ++k) {
        tail += t;
        t *= b / (k + 1);
    }
    return 2.0 * tail < 1e-15;
}

double cheb64_beta_limit(int P) {
    double lo = 0.0, hi = 64.0;
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (cheb64_order_ok(P, mid)) lo = mid;
        else hi = mid;
    }
    return lo;
}

#define PRIPPS_CHEB64_ORDERS(X) X(16) X(20) X(24) X(28) X(32) X(36)

hipError_t launch_cheb64(int P, const TiledArgs &a, int blocks, hipStream_t st) {
    TiledArgs b = a;
    if (!cheb64_limits(P, &b.cheb64)) return hipErrorInvalidValue;
    switch (P) {
#define PRIPPS_CASE(PP)                                                                     \
    case PP:                                                                                \
        hipLaunchKernelGGL(debye_cheb64_kernel<PP>, dim3(blocks), dim3(256), 0, st, b);     \
        return hipGetLastError();
        PRIPPS_CHEB64_ORDERS(PRIPPS_CASE)
#undef PRIPPS_CASE
    }
    return hipErrorInvalidValue;
}

hipError_t cheb64_occupancy(int P, int *blocks_per_cu) {
    switch (P) {
#define PRIPPS_CASE(PP)                                                                     \
    case PP:                                                                                \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,                 \
                                                            debye_cheb64_kernel<PP>, 256, 0);
        PRIPPS_CHEB64_ORDERS(PRIPPS_CASE)
#undef PRIPPS_CASE
    }
    return hipErrorInvalidValue;
}

// Orders the kernel is built for; the host rounds a structure's need up.
#define PRIPPS_CHEB_ORDERS(X) \
    X(12) X(16) X(20) X(22) X(24) X(26) X(28) X(30) X(32) X(36) X(40)

hipError_t launch_cheb(int P, const TiledArgs &a, int blocks, hipStream_t st) {
    TiledArgs b = a;
    if (!cheb_limits(P, &b.cheb)) return hipErrorInvalidValue;
    switch (P) {
#define PRIPPS_CASE(PP)                                                                     \
    case PP:                                                                                \
        hipLaunchKernelGGL(debye_cheb_kernel<PP>, dim3(blocks), dim3(256), 0, st, b);       \
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
    return 2.0 * tail < 1e-8;
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
    // smallest built order P with 2 sum_{p >= P} (beta/2)^p / p! < 1e-8
    // (|J_p(beta)| <= (beta/2)^p / p! for all p >= 0)
    for (int P : kChebOrders)
        if (cheb_order_ok(P, beta_max)) return P;
    return 0;  // beta too large for any built order
}

}  // namespace pripps
