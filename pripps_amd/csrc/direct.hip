// Direct Fourier transform of a periodic box on reciprocal-lattice q-vectors
// (gfx950): the reference's DirectTransform / build_model
// (src/explicit.rs:151-326) and its wgpu structure_factor kernel
// (src/gpu.rs:48-150).
//
// Work decomposition
//   One lane owns one q-vector n (q = 2 pi n / L); a wave holds a tile of 64
//   vectors and walks a span of atoms that is the same for every lane, so each
//   atom is one scalar load (constant address space) shared by the wave.  The
//   lanes never exchange data: every vector's sums stay in its own lane.
//
// Per atom and vector
//   The phase in revolutions is n . t with t = x / L reduced to [-1/2, 1/2) on
//   the host (exact: n is an integer vector, so only t mod 1 matters), then
//   cos and sin of 2 pi fract(n . t).  FP32 mode uses the hardware sine and
//   cosine of a revolution argument; FP64 mode uses the FP64 sine and cosine
//   of a revolution argument in trig_f64.h (exact reduction to the nearest
//   quarter turn, Taylor polynomials on |t| <= 1/8 turn).
//
// Accumulation
//   Atoms are sorted by class (kind, hydrated) inside each molecule, and the
//   box is described by a flat table of segments (runs of one class inside
//   one molecule), so a run of atoms shares its form factors: the lane sums cos and sin (and w cos,
//   w sin for hydrated atoms) over at most kDirBlock atoms of one class in the
//   working precision, then carries them into FP64 amplitudes multiplied by
//   the class's form factors at the lane's |q|:
//       A += fA (C, S),  E += fE (C, S),  H += fW (Cw, Sw).
//   With molecules, the atom amplitude of molecule m is completed in FP64 and
//   added to A, and |A_m|^2 to the self term.  Items write FP64 partials to
//   their own slot; a finalize kernel sums the slots of each vector in item
//   order, so results depend on the plan only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "internal.h"
#include "trig_f64.h"
#include "pripps_direct.h"

namespace pripps {
EngineAccess engine_access(pripps_engine *e);
void engine_retain(pripps_engine *e);
void engine_unref(pripps_engine *e);
void set_last_error(const std::string &msg);
}  // namespace pripps

using namespace pripps;

#define PRIPPS_CONST __attribute__((address_space(4)))

namespace {

constexpr int kDirVals = 7;     // Re A, Im A, Re E, Im E, Re H, Im H, self
constexpr int kDirBlock = 256;  // atoms summed in the working precision per carry
constexpr int kAtomPad = 8;   // zero atoms after the last one (FP32 prefetch)

struct DirectItem {
    int32_t tile;    // vectors [64 tile, 64 tile + 64)
    int32_t a0, a1;  // atom span (whole molecules when the self term is on)
    int32_t seg0;    // the segment holding atom a0
};

// One run of atoms of one class inside one molecule, in device memory: read
// with wave-uniform indices as scalar loads.  flags bit 0: hydrated class;
// bit 1: the run ends its molecule.
struct DirectSegment {
    int32_t start;   // first atom; the next segment's start ends it
    int16_t kind;
    int16_t flags;
};

struct DirectArgs {
    const void *atoms;        // Atom<T>[n]: fractional coordinates t and weight w
    const DirectSegment *segs;  // [n_segs + 1], the last a sentinel at n
    const void *nvec;         // T[V_pad][4]: lattice indices as T
    const double *ffa;        // [K][V]
    const double *ffe;        // [K][V] or nullptr
    const double *ffw;        // [V] or nullptr
    const DirectItem *items;
    int64_t n_items;
    int *counter;             // work-queue head, zeroed before launch
    double *partials;         // [n_items][7][64]
    int32_t V;                // q-vectors
    int32_t self;             // 1: accumulate the per-molecule self term
};

struct FinalizeDirectArgs {
    const double *partials;
    const int32_t *tile_ptr;    // [n_tiles + 1] offsets into tile_items
    const int32_t *tile_items;  // item ids of each tile, in span order
    double *amp;                // [V][7]
    int32_t V;
};

struct GroupArgs {
    const double *amp;          // [V][7]
    const int32_t *group_ptr;   // [G + 1] offsets into group_vec
    const int32_t *group_vec;   // vectors of each group, ascending
    double *out;                // [8][G]: total, atoms, excluded, hydration, self,
                                //   Re A E*, Re A H*, Re E H*
    int32_t G;
    double vs, cd;
};

__device__ __forceinline__ int direct_lane() { return __lane_id(); }

template <typename T>
__device__ __forceinline__ Atom<T> load_frac(const PRIPPS_CONST Atom<T> *p, int j) {
    Atom<T> a;
    a.x = p[j].x;
    a.y = p[j].y;
    a.z = p[j].z;
    a.w = p[j].w;
    return a;
}

// cos and sin of 2 pi (n . t) for one atom.
// DIRECT: the host proved |n . t| < 255 revolutions for every vector (|t| <=
// 1/2 per axis), inside the [-256, 256] domain where v_sin_f32 / v_cos_f32
// reduce the argument themselves, so the fract is left out.  On negative
// phases that is also the more accurate form: fract(x) for x < 0 rounds
// 1 - |x| to the coarser ulp of a number near 1.
template <bool DIRECT>
__device__ __forceinline__ void cis_rev(float nx, float ny, float nz, const Atom<float> &t,
                                        float &c, float &s) {
    const float p = __builtin_fmaf(nz, t.z, __builtin_fmaf(ny, t.y, nx * t.x));
    const float f = DIRECT ? p : __builtin_amdgcn_fractf(p);
    c = __builtin_amdgcn_cosf(f);
    s = __builtin_amdgcn_sinf(f);
}

template <bool DIRECT>
__device__ __forceinline__ void cis_rev(double nx, double ny, double nz, const Atom<double> &t,
                                        double &c, double &s) {
    const double p = fma(nz, t.z, fma(ny, t.y, nx * t.x));
    sincos_rev_f64(p, s, c);
}

// Sums over atoms [jb, je) of one class (je > jb): C = sum cos, S = sum sin
// and, for hydrated classes, Cw = sum w cos, Sw = sum w sin.  Two atoms per
// step with two chains per sum keep consecutive adds independent, and the
// next two atoms' scalar loads are issued before the current two are used
// (every index is clamped to je - 1, so no load leaves the run).
// Four consecutive FP32 atoms: one 64-byte scalar load.
struct alignas(16) AtomQuad {
    Atom<float> a[4];
};

__device__ __forceinline__ AtomQuad load_quad(const PRIPPS_CONST Atom<float> *ca, int j) {
    const PRIPPS_CONST float *f = (const PRIPPS_CONST float *)(ca + j);
    AtomQuad r;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        r.a[t].x = f[4 * t];
        r.a[t].y = f[4 * t + 1];
        r.a[t].z = f[4 * t + 2];
        r.a[t].w = f[4 * t + 3];
    }
    return r;
}

// FP32 run over atoms [jb, je): quads of four atoms, each one 64-byte scalar
// load, two quads in flight.  The scalar unit is shared by the CU's four
// SIMDs; with one load, clamp and address update per atom (the FP64 form
// below) the kernel issued as many scalar as vector instructions per term.
// This form issues ~1.75 scalar instructions per atom (measured: 1.04 ->
// 1.00 ms on the box workload, same sums).  Loads past je read the following
// atoms or the zero padding after the last atom (kAtomPad >= 8), and are not
// used.
template <bool WET, bool DIRECT>
__device__ __forceinline__ void class_run_f32(const PRIPPS_CONST Atom<float> *ca, int jb, int je,
                                              float nx, float ny, float nz, float &C, float &S,
                                              float &Cw, float &Sw) {
    float c0 = 0.f, s0 = 0.f, c1 = 0.f, s1 = 0.f;
    float cw0 = 0.f, sw0 = 0.f, cw1 = 0.f, sw1 = 0.f;
    auto one = [&](const Atom<float> &a, float &c, float &s, float &cw, float &sw) {
        float ct, st;
        cis_rev<DIRECT>(nx, ny, nz, a, ct, st);
        c += ct;
        s += st;
        if constexpr (WET) {
            cw = fmaf(a.w, ct, cw);
            sw = fmaf(a.w, st, sw);
        }
    };
    // two quads in flight, alternating (no register copies between steps)
    auto quad = [&](const AtomQuad &q) {
        one(q.a[0], c0, s0, cw0, sw0);
        one(q.a[1], c1, s1, cw1, sw1);
        one(q.a[2], c0, s0, cw0, sw0);
        one(q.a[3], c1, s1, cw1, sw1);
    };
    auto use_w = [](const AtomQuad &q) {
        if constexpr (!WET) {
            // keep the weights "used" so that a quad stays one 64-byte scalar
            // load (no instruction is emitted)
#pragma unroll
            for (int t = 0; t < 4; ++t) asm volatile("" ::"s"(q.a[t].w));
        }
    };
    AtomQuad q0 = load_quad(ca, jb);
    int j = jb;
    for (; j + 7 < je; j += 8) {
        const AtomQuad q1 = load_quad(ca, j + 4);
        use_w(q1);
        quad(q0);
        q0 = load_quad(ca, j + 8);
        use_w(q0);
        quad(q1);
    }
    if (j + 3 < je) {
        const AtomQuad q1 = load_quad(ca, j + 4);
        quad(q0);
        q0 = q1;
        j += 4;
    }
    // 0-3 atoms left, in q0
    if (j < je) one(q0.a[0], c0, s0, cw0, sw0);
    if (j + 1 < je) one(q0.a[1], c1, s1, cw1, sw1);
    if (j + 2 < je) one(q0.a[2], c0, s0, cw0, sw0);
    C = c0 + c1;
    S = s0 + s1;
    Cw = cw0 + cw1;
    Sw = sw0 + sw1;
}

template <typename T, bool WET, bool DIRECT>
__device__ __forceinline__ void class_run(const PRIPPS_CONST Atom<T> *ca, int jb, int je, T nx,
                                          T ny, T nz, T &C, T &S, T &Cw, T &Sw) {
    if constexpr (sizeof(T) == 4) {
        class_run_f32<WET, DIRECT>(ca, jb, je, nx, ny, nz, C, S, Cw, Sw);
        return;
    }
    T c0 = T(0), s0 = T(0), c1 = T(0), s1 = T(0);
    T cw0 = T(0), sw0 = T(0), cw1 = T(0), sw1 = T(0);
    const int last = je - 1;
    Atom<T> a = load_frac(ca, jb);
    Atom<T> b = load_frac(ca, min(jb + 1, last));
    int j = jb;
    for (; j + 1 < je; j += 2) {
        const Atom<T> na = load_frac(ca, min(j + 2, last));
        const Atom<T> nb = load_frac(ca, min(j + 3, last));
        T ca_, sa_, cb_, sb_;
        cis_rev<DIRECT>(nx, ny, nz, a, ca_, sa_);
        cis_rev<DIRECT>(nx, ny, nz, b, cb_, sb_);
        c0 += ca_;
        s0 += sa_;
        c1 += cb_;
        s1 += sb_;
        if constexpr (WET) {
            cw0 = fma(a.w, ca_, cw0);
            sw0 = fma(a.w, sa_, sw0);
            cw1 = fma(b.w, cb_, cw1);
            sw1 = fma(b.w, sb_, sw1);
        }
        a = na;
        b = nb;
    }
    if (j < je) {  // odd count: `a` holds atom j
        T ca_, sa_;
        cis_rev<DIRECT>(nx, ny, nz, a, ca_, sa_);
        c0 += ca_;
        s0 += sa_;
        if constexpr (WET) {
            cw0 = fma(a.w, ca_, cw0);
            sw0 = fma(a.w, sa_, sw0);
        }
    }
    C = c0 + c1;
    S = s0 + s1;
    Cw = cw0 + cw1;
    Sw = sw0 + sw1;
}

template <typename T, bool DIRECT>
__global__ void __launch_bounds__(256) direct_kernel(DirectArgs args) {
    const int lane = direct_lane();
    const PRIPPS_CONST DirectSegment *segs = (const PRIPPS_CONST DirectSegment *)(args.segs);
    const PRIPPS_CONST Atom<T> *atoms = (const PRIPPS_CONST Atom<T> *)(args.atoms);
    for (;;) {
        int w = 0;
        if (lane == 0) w = atomicAdd(args.counter, 1);
        w = __builtin_amdgcn_readfirstlane(__shfl(w, 0, 64));
        if (w >= args.n_items) break;
        const PRIPPS_CONST DirectItem *ip = (const PRIPPS_CONST DirectItem *)(args.items) + w;
        const int tile = ip->tile, a0 = ip->a0, a1 = ip->a1;
        int sg = ip->seg0;

        const int V = args.V;
        const int v = min(tile * 64 + lane, V - 1);  // padding lanes repeat the last vector
        const T *nv = static_cast<const T *>(args.nvec) + 4 * (int64_t)v;
        const T nx = nv[0], ny = nv[1], nz = nv[2];
        const double fw = args.ffw ? args.ffw[v] : 0.0;

        double Ar = 0, Ai = 0, Er = 0, Ei = 0, Hr = 0, Hi = 0, self = 0;
        double Mr = 0, Mi = 0;  // atom amplitude of the current molecule (self mode)
        int j = a0;
        while (j < a1) {
            const int s_end = segs[sg + 1].start;
            const int seg_end = min(min(s_end, a1), j + kDirBlock);
            const int k = segs[sg].kind;
            const int flags = segs[sg].flags;
            T C, S, Cw = T(0), Sw = T(0);
            if (flags & 1) class_run<T, true, DIRECT>(atoms, j, seg_end, nx, ny, nz, C, S, Cw, Sw);
            else class_run<T, false, DIRECT>(atoms, j, seg_end, nx, ny, nz, C, S, Cw, Sw);
            const double dC = double(C), dS = double(S);
            const double fa = args.ffa[(int64_t)k * V + v];
            if (args.self) {
                Mr += fa * dC;
                Mi += fa * dS;
            } else {
                Ar += fa * dC;
                Ai += fa * dS;
            }
            if (args.ffe) {
                const double fe = args.ffe[(int64_t)k * V + v];
                Er += fe * dC;
                Ei += fe * dS;
            }
            Hr += fw * double(Cw);
            Hi += fw * double(Sw);
            j = seg_end;
            if (j == s_end) {
                if (args.self && (flags & 2)) {  // molecule complete
                    Ar += Mr;
                    Ai += Mi;
                    self += Mr * Mr + Mi * Mi;
                    Mr = Mi = 0.0;
                }
                ++sg;
            }
        }
        double *out = args.partials + (int64_t)w * (kDirVals * 64) + lane;
        out[0 * 64] = Ar;
        out[1 * 64] = Ai;
        out[2 * 64] = Er;
        out[3 * 64] = Ei;
        out[4 * 64] = Hr;
        out[5 * 64] = Hi;
        out[6 * 64] = self;
    }
}

// amp[v][7] = sum over the items of v's tile.  One block per 64-vector
// tile: four groups of 64 threads take the tile's items g, g + 4, ... (span
// order), and the four group sums are added in group order, so the result is
// a function of the plan alone.
__global__ void __launch_bounds__(256) direct_finalize_kernel(FinalizeDirectArgs a) {
    __shared__ double red[4][kDirVals][64];
    const int tile = blockIdx.x;
    const int lane = threadIdx.x % 64, g = threadIdx.x / 64;
    double acc[kDirVals] = {0, 0, 0, 0, 0, 0, 0};
    const int b = a.tile_ptr[tile], e = a.tile_ptr[tile + 1];
#pragma unroll 4
    for (int p = b + g; p < e; p += 4) {
        const double *src = a.partials + (int64_t)a.tile_items[p] * (kDirVals * 64) + lane;
#pragma unroll
        for (int t = 0; t < kDirVals; ++t) acc[t] += src[t * 64];
    }
#pragma unroll
    for (int t = 0; t < kDirVals; ++t) red[g][t][lane] = acc[t];
    __syncthreads();
    const int v = tile * 64 + lane;
    if (g == 0 && v < a.V) {
#pragma unroll
        for (int t = 0; t < kDirVals; ++t)
            a.amp[(int64_t)v * kDirVals + t] =
                ((red[0][t][lane] + red[1][t][lane]) + red[2][t][lane]) + red[3][t][lane];
    }
}

// Group means, in the reference's operation order per vector:
// |A - vs E + cd H|^2, |A|^2, |E|^2, |H|^2 (norm_sqr = re re + im im), summed
// over the group's vectors and divided by the count.  The cross means
// Re A E*, Re A H*, Re E H* complete the six profiles in which the total is
// the Debye combine: |A|^2 + vs^2 |E|^2 + cd^2 |H|^2 - 2 vs Re A E*
// + 2 cd Re A H* - 2 vs cd Re E H*.  One block per group: thread t takes the
// group's vectors t, t + 256, ... in order, and a fixed LDS tree combines the
// threads, so the result depends on the group size only.
constexpr int kGroupThreads = 256;

__global__ void __launch_bounds__(kGroupThreads) direct_group_kernel(GroupArgs a) {
#pragma clang fp contract(off)
    __shared__ double red[8][kGroupThreads];
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int b = a.group_ptr[g], e = a.group_ptr[g + 1];
    for (int p = b + tid; p < e; p += kGroupThreads) {
        const double *r = a.amp + (int64_t)a.group_vec[p] * kDirVals;
        const double re = (r[0] - a.vs * r[2]) + a.cd * r[4];
        const double im = (r[1] - a.vs * r[3]) + a.cd * r[5];
        acc[0] += re * re + im * im;
        acc[1] += r[0] * r[0] + r[1] * r[1];
        acc[2] += r[2] * r[2] + r[3] * r[3];
        acc[3] += r[4] * r[4] + r[5] * r[5];
        acc[4] += r[6];
        acc[5] += r[0] * r[2] + r[1] * r[3];
        acc[6] += r[0] * r[4] + r[1] * r[5];
        acc[7] += r[2] * r[4] + r[3] * r[5];
    }
#pragma unroll
    for (int t = 0; t < 8; ++t) red[t][tid] = acc[t];
    __syncthreads();
    for (int half = kGroupThreads / 2; half > 0; half /= 2) {
        if (tid < half) {
#pragma unroll
            for (int t = 0; t < 8; ++t) red[t][tid] += red[t][tid + half];
        }
        __syncthreads();
    }
    if (tid < 8) a.out[tid * a.G + g] = red[tid][0] / double(e - b);
}

// Trajectory frames: acc[v] += w (|A - vs E + cd H|^2, self) for one frame.
__global__ void __launch_bounds__(256) direct_frames_accumulate_kernel(const double *amp, int V,
                                                                       double vs, double cd,
                                                                       double w, double *acc) {
#pragma clang fp contract(off)
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const double *r = amp + (int64_t)v * kDirVals;
    const double re = (r[0] - vs * r[2]) + cd * r[4];
    const double im = (r[1] - vs * r[3]) + cd * r[5];
    acc[2 * (int64_t)v] += w * (re * re + im * im);
    acc[2 * (int64_t)v + 1] += w * r[6];
}

// out[t][g] = mean over group g of acc[v][t] / wsum (t = 0 total, 1 self);
// the same fixed-order block reduction as direct_group_kernel.
__global__ void __launch_bounds__(kGroupThreads) direct_frames_group_kernel(
    const double *acc, const int32_t *group_ptr, const int32_t *group_vec, int G, double wsum,
    double *out) {
    __shared__ double red[2][kGroupThreads];
    const int g = blockIdx.x;
    const int tid = threadIdx.x;
    double t0 = 0, t1 = 0;
    const int b = group_ptr[g], e = group_ptr[g + 1];
    for (int p = b + tid; p < e; p += kGroupThreads) {
        t0 += acc[2 * (int64_t)group_vec[p]];
        t1 += acc[2 * (int64_t)group_vec[p] + 1];
    }
    red[0][tid] = t0;
    red[1][tid] = t1;
    __syncthreads();
    for (int half = kGroupThreads / 2; half > 0; half /= 2) {
        if (tid < half) {
            red[0][tid] += red[0][tid + half];
            red[1][tid] += red[1][tid + half];
        }
        __syncthreads();
    }
    if (tid < 2) out[tid * G + g] = red[tid][0] / double(e - b) / wsum;
}

__global__ void direct_fill_int_kernel(int *p, int v) { *p = v; }

__global__ void direct_fill_double_kernel(double *p, int64_t n) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
int fail(int code, const std::string &msg) {
    set_last_error(msg);
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    return fail(e == hipErrorOutOfMemory ? PRIPPS_ERR_OUT_OF_MEMORY : PRIPPS_ERR_DEVICE,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define DCHK(expr)                                         \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);  \
    } while (0)

bool finite_all(const double *p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(p[i])) return false;
    return true;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

struct pripps_direct_model {
    pripps_engine *engine = nullptr;
    EngineAccess acc{};
    bool fp64 = false;
    int32_t V = 0, G = 0, n_items = 0, n_tiles = 0, blocks = 0;
    int64_t n_atoms = 0;
    bool has_self = false;
    bool sin_direct = false;  // every |n . t| < 255 revolutions (FP32 kernel)
    std::vector<double> group_q;
    void *dev = nullptr;  // one allocation: inputs, items, partials, amplitudes, outputs
    DirectArgs args{};
    FinalizeDirectArgs fargs{};
    GroupArgs gargs{};
    double *d_amp = nullptr;
    double *d_out = nullptr;
    pripps_build_stats stats{};
    // re-staging of positions (trajectory frames)
    std::vector<int32_t> order;  // sorted atom -> input atom
    double L[3] = {0, 0, 0};
    char *d_atoms = nullptr;
};

namespace {

hipError_t enqueue(const pripps_direct_model &m, hipStream_t st, hipEvent_t ev_a,
                   hipEvent_t ev_b) {
    hipLaunchKernelGGL(direct_fill_int_kernel, dim3(1), dim3(1), 0, st, m.args.counter, 0);
    hipError_t err = hipGetLastError();
    if (err == hipSuccess && ev_a) err = hipEventRecord(ev_a, st);
    if (err == hipSuccess && m.n_items > 0) {
        if (m.fp64)
            hipLaunchKernelGGL((direct_kernel<double, false>), dim3(m.blocks), dim3(256), 0, st, m.args);
        else if (m.sin_direct)
            hipLaunchKernelGGL((direct_kernel<float, true>), dim3(m.blocks), dim3(256), 0, st, m.args);
        else
            hipLaunchKernelGGL((direct_kernel<float, false>), dim3(m.blocks), dim3(256), 0, st, m.args);
        err = hipGetLastError();
    }
    if (err == hipSuccess && ev_b) err = hipEventRecord(ev_b, st);
    if (err == hipSuccess && m.V > 0) {
        hipLaunchKernelGGL(direct_finalize_kernel, dim3(m.n_tiles), dim3(256), 0, st,
                           m.fargs);
        err = hipGetLastError();
    }
    return err;
}

template <typename T>
void stage_atoms(const pripps_debye_input *in, const double L[3], const std::vector<int32_t> &order,
                 char *dst) {
    Atom<T> *a = reinterpret_cast<Atom<T> *>(dst);
    for (size_t s = 0; s < order.size(); ++s) {
        const int32_t i = order[s];
        const double p[3] = {in->x[i], in->y[i], in->z[i]};
        double t[3];
        for (int d = 0; d < 3; ++d) {
            const double u = p[d] / L[d];
            t[d] = u - std::nearbyint(u);
        }
        a[s].x = T(t[0]);
        a[s].y = T(t[1]);
        a[s].z = T(t[2]);
        a[s].w = in->hydration_weight ? T(in->hydration_weight[i]) : T(0);
    }
}

template <typename T>
void stage_vectors(const int32_t *idx, int V, int V_pad, char *dst) {
    T *n = reinterpret_cast<T *>(dst);
    for (int v = 0; v < V_pad; ++v) {
        const int s = std::min(v, V - 1);
        for (int d = 0; d < 3; ++d) n[4 * v + d] = T(idx[3 * s + d]);
        n[4 * v + 3] = T(0);
    }
}

}  // namespace

extern "C" int pripps_direct_model_build(pripps_engine *engine, const pripps_debye_input *in,
                                         const pripps_direct_grid *grid, int32_t precision,
                                         pripps_direct_model **out) {
    const auto t_start = std::chrono::steady_clock::now();
    if (!engine || !in || !grid || !out)
        return fail(PRIPPS_ERR_USAGE, "pripps_direct_model_build: NULL argument");
    *out = nullptr;
    if (precision != PRIPPS_FP32_COMPENSATED && precision != PRIPPS_FP64)
        return fail(PRIPPS_ERR_USAGE, "Direct transform precision must be FP32_COMPENSATED or FP64");
    const int64_t N = in->n_atoms;
    const int32_t V = grid->n_vectors, G = grid->n_groups, K = in->n_kinds;
    if (N < 0 || V < 0 || G < 0) return fail(PRIPPS_ERR_VALIDATION, "negative size");
    if (in->nq != V)
        return fail(PRIPPS_ERR_USAGE, "input nq must equal the number of q-vectors");
    if (N > INT32_MAX / 2) return fail(PRIPPS_ERR_VALIDATION, "too many atoms");
    if (K < 1 || K > INT16_MAX)
        return fail(PRIPPS_ERR_VALIDATION, "n_kinds must be in [1, 32767]");
    for (int d = 0; d < 3; ++d)
        if (!(std::isfinite(grid->box[d]) && grid->box[d] > 0.0))
            return fail(PRIPPS_ERR_USAGE, "invalid box side lengths");
    if (N > 0 && (!in->x || !in->y || !in->z || !in->kind))
        return fail(PRIPPS_ERR_USAGE, "coordinates and kinds must be non-NULL");
    if (V > 0 && (!grid->indices || !in->atom_ff || !grid->group))
        return fail(PRIPPS_ERR_USAGE, "indices, atom_ff and group must be non-NULL");
    if (G > 0 && !grid->group_q) return fail(PRIPPS_ERR_USAGE, "group_q must be non-NULL");
    if ((in->water_ff != nullptr) != (in->hydration_weight != nullptr))
        return fail(PRIPPS_ERR_USAGE, "water_ff and hydration_weight go together");
    if (!finite_all(in->x, N) || !finite_all(in->y, N) || !finite_all(in->z, N))
        return fail(PRIPPS_ERR_VALIDATION, "coordinates must be finite");
    const size_t KV = (size_t)K * V;
    if (!finite_all(in->atom_ff, KV) || (in->excluded_ff && !finite_all(in->excluded_ff, KV)) ||
        (in->water_ff && !finite_all(in->water_ff, V)) ||
        (in->hydration_weight && !finite_all(in->hydration_weight, N)))
        return fail(PRIPPS_ERR_VALIDATION, "form factors and weights must be finite");
    for (int64_t i = 0; i < N; ++i)
        if (in->kind[i] < 0 || in->kind[i] >= K)
            return fail(PRIPPS_ERR_VALIDATION, "kind id out of range");
    std::vector<int32_t> gcount(G, 0);
    for (int32_t v = 0; v < V; ++v) {
        const int32_t g = grid->group[v];
        if (g < 0 || g >= G) return fail(PRIPPS_ERR_VALIDATION, "group id out of range");
        ++gcount[g];
    }
    for (int32_t g = 0; g < G; ++g)
        if (gcount[g] == 0) return fail(PRIPPS_ERR_VALIDATION, "empty q group");

    // molecules: contiguous chunks of max(n / num_molecules, 1) atoms, the last
    // one possibly shorter (the reference's `chunks(m)`, src/explicit.rs:222-236;
    // Pripps::intensity's own guards, src/lib.rs:362-403, stay with the caller)
    const int32_t nmol = grid->num_molecules;
    if (nmol < 0) return fail(PRIPPS_ERR_USAGE, "num_molecules must be >= 0");
    const bool self_mode = nmol > 0;
    if (self_mode && (in->excluded_ff || in->hydration_weight))
        return fail(PRIPPS_ERR_USAGE,
                    "--num-molecules supports vacuum runs only (no excluded-volume/hydration model)");
    const int64_t M = self_mode ? std::max<int64_t>(N / nmol, 1) : std::max<int64_t>(N, 1);

    // atoms sorted by class (kind, hydrated) inside each molecule, stable; one
    // segment per class run and molecule
    auto cls_of = [&](int64_t i) {
        const bool wet = in->hydration_weight && in->hydration_weight[i] != 0.0;
        return 2 * in->kind[i] + (wet ? 1 : 0);
    };
    std::vector<int32_t> order(N);
    std::vector<DirectSegment> segs;
    DirectArgs args{};
    for (int64_t m0 = 0; m0 < N; m0 += M) {
        const int64_t m1 = std::min<int64_t>(N, m0 + M);
        std::iota(order.begin() + m0, order.begin() + m1, (int32_t)m0);
        std::stable_sort(order.begin() + m0, order.begin() + m1,
                         [&](int32_t a, int32_t b) { return cls_of(a) < cls_of(b); });
        int prev = -1;
        for (int64_t s = m0; s < m1; ++s) {
            const int cl = cls_of(order[s]);
            if (cl != prev) {
                segs.push_back(DirectSegment{(int32_t)s, (int16_t)(cl / 2), (int16_t)(cl % 2)});
                prev = cl;
            }
        }
        segs.back().flags |= 2;  // ends its molecule
    }
    segs.push_back(DirectSegment{(int32_t)N, 0, 0});  // sentinel

    // plan: 64-vector tiles x atom spans
    EngineAccess acc = engine_access(engine);
    const int n_tiles = (V + 63) / 64;
    int64_t span = 0;
    std::vector<DirectItem> items;
    if (N > 0 && V > 0) {
        const int64_t target_items = 256LL * acc.n_cu;  // several items per wave slot
        span = std::max<int64_t>(256, (N * n_tiles + target_items - 1) / target_items);
        if (self_mode) span = std::max<int64_t>(1, (span + M - 1) / M) * M;
        span = std::min<int64_t>(span, N);
        int sg = 0;
        for (int64_t a0 = 0; a0 < N; a0 += span) {
            while (segs[sg + 1].start <= a0) ++sg;  // the segment holding a0
            for (int t = 0; t < n_tiles; ++t)
                items.push_back(DirectItem{t, (int32_t)a0, (int32_t)std::min<int64_t>(N, a0 + span), sg});
        }
    }
    const int n_items = (int)items.size();
    // CSR of items per tile, in span order
    std::vector<int32_t> tile_ptr(n_tiles + 1, 0), tile_items(n_items);
    for (const auto &it : items) ++tile_ptr[it.tile + 1];
    for (int t = 0; t < n_tiles; ++t) tile_ptr[t + 1] += tile_ptr[t];
    {
        std::vector<int32_t> fill(tile_ptr.begin(), tile_ptr.end() - 1);
        for (int w = 0; w < n_items; ++w) tile_items[fill[items[w].tile]++] = w;
    }
    // CSR of vectors per group, ascending
    std::vector<int32_t> group_ptr(G + 1, 0), group_vec(V);
    for (int32_t g = 0; g < G; ++g) group_ptr[g + 1] = group_ptr[g] + gcount[g];
    {
        std::vector<int32_t> fill(group_ptr.begin(), group_ptr.end() - 1);
        for (int32_t v = 0; v < V; ++v) group_vec[fill[grid->group[v]]++] = v;
    }

    // device layout
    const bool fp64 = precision == PRIPPS_FP64;
    const size_t esz = fp64 ? sizeof(double) : sizeof(float);
    const int V_pad = n_tiles * 64;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes);
        return o;
    };
    // kAtomPad zero atoms after the last one: the FP32 kernel prefetches up
    // to that far past the end of a run (the values are never used)
    const size_t o_atoms = take((size_t)(N + kAtomPad) * 4 * esz);
    const size_t o_nvec = take((size_t)std::max(V_pad, 1) * 4 * esz);
    const size_t o_ffa = take(std::max<size_t>(KV, 1) * sizeof(double));
    const size_t o_ffe = take(in->excluded_ff ? KV * sizeof(double) : 0);
    const size_t o_ffw = take(in->water_ff ? (size_t)V * sizeof(double) : 0);
    const size_t o_items = take((size_t)std::max(n_items, 1) * sizeof(DirectItem));
    const size_t o_tptr = take((size_t)(n_tiles + 1) * sizeof(int32_t));
    const size_t o_titems = take((size_t)std::max(n_items, 1) * sizeof(int32_t));
    const size_t o_gptr = take((size_t)(G + 1) * sizeof(int32_t));
    const size_t o_gvec = take((size_t)std::max(V, 1) * sizeof(int32_t));
    const size_t o_counter = take(sizeof(int));
    const size_t o_segs = take(segs.size() * sizeof(DirectSegment));
    const size_t host_bytes = off;  // everything above is staged from the host
    const size_t o_part = take((size_t)std::max(n_items, 1) * kDirVals * 64 * sizeof(double));
    const size_t o_amp = take((size_t)std::max(V, 1) * kDirVals * sizeof(double));
    const size_t o_out = take((size_t)std::max(G, 1) * 8 * sizeof(double));
    const size_t total = off;

    std::vector<char> host(host_bytes, 0);
    if (N > 0) {
        const double L[3] = {grid->box[0], grid->box[1], grid->box[2]};
        if (fp64) stage_atoms<double>(in, L, order, host.data() + o_atoms);
        else stage_atoms<float>(in, L, order, host.data() + o_atoms);
    }
    if (V > 0) {
        if (fp64) stage_vectors<double>(grid->indices, V, V_pad, host.data() + o_nvec);
        else stage_vectors<float>(grid->indices, V, V_pad, host.data() + o_nvec);
        std::memcpy(host.data() + o_ffa, in->atom_ff, KV * sizeof(double));
        if (in->excluded_ff) std::memcpy(host.data() + o_ffe, in->excluded_ff, KV * sizeof(double));
        if (in->water_ff) std::memcpy(host.data() + o_ffw, in->water_ff, V * sizeof(double));
    }
    if (n_items) std::memcpy(host.data() + o_items, items.data(), n_items * sizeof(DirectItem));
    std::memcpy(host.data() + o_tptr, tile_ptr.data(), tile_ptr.size() * sizeof(int32_t));
    if (n_items) std::memcpy(host.data() + o_titems, tile_items.data(), n_items * sizeof(int32_t));
    std::memcpy(host.data() + o_gptr, group_ptr.data(), group_ptr.size() * sizeof(int32_t));
    if (V) std::memcpy(host.data() + o_gvec, group_vec.data(), V * sizeof(int32_t));
    std::memcpy(host.data() + o_segs, segs.data(), segs.size() * sizeof(DirectSegment));

    auto *m = new pripps_direct_model();
    m->engine = engine;
    m->acc = acc;
    m->fp64 = fp64;
    m->V = V;
    m->G = G;
    m->n_items = n_items;
    m->n_tiles = n_tiles;
    m->n_atoms = N;
    m->has_self = self_mode;
    {
        // |n . t| <= (|nx| + |ny| + |nz|) / 2 with |t| <= 1/2 per axis
        int64_t n1 = 0;
        for (int32_t v = 0; v < V; ++v) {
            const int32_t *ix = grid->indices + 3 * (int64_t)v;
            n1 = std::max<int64_t>(n1, std::llabs(ix[0]) + std::llabs(ix[1]) + std::llabs(ix[2]));
        }
        m->sin_direct = 0.5 * (double)n1 < 255.0;
    }
    m->group_q.assign(grid->group_q, grid->group_q + G);
    m->order = order;
    for (int d = 0; d < 3; ++d) m->L[d] = grid->box[d];

    std::lock_guard<std::mutex> lock(*acc.mu);
    auto abort_build = [&](int rc) {
        if (m->dev) (void)hipFree(m->dev);
        delete m;
        return rc;
    };
    hipError_t err = hipSetDevice(acc.device);
    if (err == hipSuccess) err = hipMalloc(&m->dev, total);
    if (err != hipSuccess) {
        m->dev = nullptr;
        return abort_build(hip_fail(err, "pripps_direct_model_build: device allocation"));
    }
    char *d = static_cast<char *>(m->dev);
    err = hipMemcpyAsync(d, host.data(), host_bytes, hipMemcpyHostToDevice, acc.stream);
    if (err != hipSuccess) return abort_build(hip_fail(err, "hipMemcpyAsync(direct inputs)"));

    args.atoms = d + o_atoms;
    m->d_atoms = d + o_atoms;
    args.segs = reinterpret_cast<const DirectSegment *>(d + o_segs);
    args.nvec = d + o_nvec;
    args.ffa = reinterpret_cast<const double *>(d + o_ffa);
    args.ffe = in->excluded_ff ? reinterpret_cast<const double *>(d + o_ffe) : nullptr;
    args.ffw = in->water_ff ? reinterpret_cast<const double *>(d + o_ffw) : nullptr;
    args.items = reinterpret_cast<const DirectItem *>(d + o_items);
    args.n_items = n_items;
    args.counter = reinterpret_cast<int *>(d + o_counter);
    args.partials = reinterpret_cast<double *>(d + o_part);
    args.V = V;
    args.self = self_mode ? 1 : 0;
    m->args = args;
    m->d_amp = reinterpret_cast<double *>(d + o_amp);
    m->d_out = reinterpret_cast<double *>(d + o_out);
    m->fargs = FinalizeDirectArgs{args.partials, reinterpret_cast<const int32_t *>(d + o_tptr),
                                  reinterpret_cast<const int32_t *>(d + o_titems), m->d_amp, V};
    m->gargs = GroupArgs{m->d_amp, reinterpret_cast<const int32_t *>(d + o_gptr),
                         reinterpret_cast<const int32_t *>(d + o_gvec), m->d_out, G, 0.0, 0.0};

    int per_cu = 0;
    const void *kfn = fp64 ? reinterpret_cast<const void *>(&direct_kernel<double, false>)
                      : m->sin_direct ? reinterpret_cast<const void *>(&direct_kernel<float, true>)
                                      : reinterpret_cast<const void *>(&direct_kernel<float, false>);
    err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 256, 0);
    if (err != hipSuccess) return abort_build(hip_fail(err, "hipOccupancyMaxActiveBlocksPerMultiprocessor"));
    const int waves_needed = n_items;
    m->blocks = std::max(1, std::min(std::max(per_cu, 1) * acc.n_cu, (waves_needed + 3) / 4));

    if (V > 0) {
        err = enqueue(*m, acc.stream, acc.ev0, acc.ev1);
        if (err != hipSuccess) return abort_build(hip_fail(err, "direct kernels"));
    }
    err = hipStreamSynchronize(acc.stream);
    if (err != hipSuccess) return abort_build(hip_fail(err, "hipStreamSynchronize(direct build)"));
    float kms = 0.f;
    if (V > 0 && hipEventElapsedTime(&kms, acc.ev0, acc.ev1) != hipSuccess) kms = 0.f;
    m->stats.kernel_ms = kms;
    m->stats.pair_q_terms = (double)N * (double)V;
    m->stats.span = (int32_t)span;
    m->stats.block = kDirBlock;
    m->stats.q_per_lane = 1;
    m->stats.work_items = n_items;
    m->stats.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    engine_retain(engine);
    *out = m;
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_n_groups(const pripps_direct_model *m, int32_t *n) {
    if (!m || !n) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    *n = m->G;
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_q_values(const pripps_direct_model *m, double *out) {
    if (!m || (!out && m->G)) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    std::copy(m->group_q.begin(), m->group_q.end(), out);
    return PRIPPS_OK;
}

static int run_groups(const pripps_direct_model *m, double vs, double cd, std::vector<double> &h) {
    h.assign((size_t)8 * m->G, 0.0);
    if (m->G == 0) return PRIPPS_OK;
    std::lock_guard<std::mutex> lock(*m->acc.mu);
    DCHK(hipSetDevice(m->acc.device));
    GroupArgs g = m->gargs;
    g.vs = vs;
    g.cd = cd;
    hipLaunchKernelGGL(direct_group_kernel, dim3(m->G), dim3(kGroupThreads), 0, m->acc.stream, g);
    DCHK(hipGetLastError());
    DCHK(hipMemcpyAsync(h.data(), m->d_out, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                        m->acc.stream));
    DCHK(hipStreamSynchronize(m->acc.stream));
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_intensity(const pripps_direct_model *m, double vs, double cd,
                                             double *out_total, double *out_atoms,
                                             double *out_excluded, double *out_hydration) {
    if (!m) return fail(PRIPPS_ERR_USAGE, "model is NULL");
    if (!std::isfinite(vs) || !std::isfinite(cd))
        return fail(PRIPPS_ERR_VALIDATION, "volume_scale and contrast_density must be finite");
    std::vector<double> h;
    int rc = run_groups(m, vs, cd, h);
    if (rc) return rc;
    double *dst[4] = {out_total, out_atoms, out_excluded, out_hydration};
    for (int t = 0; t < 4; ++t)
        if (dst[t]) std::copy(h.begin() + (size_t)t * m->G, h.begin() + (size_t)(t + 1) * m->G, dst[t]);
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_self_term(const pripps_direct_model *m, double *out) {
    if (!m || (!out && m->G)) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    if (!m->has_self)
        return fail(PRIPPS_ERR_USAGE, "the model was built without num_molecules; no self term");
    std::vector<double> h;
    int rc = run_groups(m, 1.0, 0.0, h);
    if (rc) return rc;
    std::copy(h.begin() + (size_t)4 * m->G, h.begin() + (size_t)5 * m->G, out);
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_profiles(const pripps_direct_model *m, double *out) {
    if (!m || (!out && m->G)) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    std::vector<double> h;
    int rc = run_groups(m, 1.0, 0.0, h);
    if (rc) return rc;
    const size_t G = m->G;
    const int src[6] = {1, 2, 3, 5, 6, 7};  // AA EE HH AE AH EH
    for (int p = 0; p < 6; ++p)
        std::copy(h.begin() + src[p] * G, h.begin() + (src[p] + 1) * G, out + p * G);
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_amplitudes(const pripps_direct_model *m, double *out) {
    if (!m || (!out && m->V)) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    if (m->V == 0) return PRIPPS_OK;
    std::lock_guard<std::mutex> lock(*m->acc.mu);
    DCHK(hipSetDevice(m->acc.device));
    DCHK(hipMemcpyAsync(out, m->d_amp, (size_t)m->V * kDirVals * sizeof(double),
                        hipMemcpyDeviceToHost, m->acc.stream));
    DCHK(hipStreamSynchronize(m->acc.stream));
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_stats(const pripps_direct_model *m, pripps_build_stats *out) {
    if (!m || !out) return fail(PRIPPS_ERR_USAGE, "NULL argument");
    *out = m->stats;
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_recompute(pripps_direct_model *m, int32_t repeats,
                                             double *kernel_ms) {
    if (!m) return fail(PRIPPS_ERR_USAGE, "model is NULL");
    if (repeats < 0) return fail(PRIPPS_ERR_USAGE, "repeats must be >= 0");
    std::lock_guard<std::mutex> lock(*m->acc.mu);
    DCHK(hipSetDevice(m->acc.device));
    double sum = 0.0;
    for (int r = 0; r < repeats && m->V > 0; ++r) {
        DCHK(enqueue(*m, m->acc.stream, m->acc.ev0, m->acc.ev1));
        DCHK(hipStreamSynchronize(m->acc.stream));
        float ms = 0.f;
        DCHK(hipEventElapsedTime(&ms, m->acc.ev0, m->acc.ev1));
        sum += ms;
    }
    if (repeats > 0) m->stats.kernel_ms = sum / repeats;
    if (kernel_ms) *kernel_ms = repeats > 0 ? sum / repeats : m->stats.kernel_ms;
    return PRIPPS_OK;
}

extern "C" int pripps_direct_model_free(pripps_direct_model *m) {
    if (!m) return PRIPPS_OK;
    pripps_engine *e = m->engine;
    {
        std::lock_guard<std::mutex> lock(*m->acc.mu);
        (void)hipSetDevice(m->acc.device);
        (void)hipStreamSynchronize(m->acc.stream);
        if (m->dev) (void)hipFree(m->dev);
    }
    delete m;
    engine_unref(e);
    return PRIPPS_OK;
}

// ---------------------------------------------------------------------------
// trajectory frames (reference trajectory_average, src/trajectory.rs:167-251)
// ---------------------------------------------------------------------------
extern "C" int pripps_direct_frames_average(pripps_engine *engine,
                                            const pripps_debye_input *topology,
                                            const pripps_direct_grid *grid, int32_t n_frames,
                                            const double *coords, const double *w,
                                            int32_t stride, int32_t precision, double vs,
                                            double cd, double *out_total, double *out_self) {
    if (!engine || !topology || !grid)
        return fail(PRIPPS_ERR_USAGE, "engine, topology and grid must be non-NULL");
    if (n_frames < 0) return fail(PRIPPS_ERR_VALIDATION, "n_frames must be >= 0");
    if (stride < 1) return fail(PRIPPS_ERR_VALIDATION, "stride must be >= 1");
    if (!std::isfinite(vs) || !std::isfinite(cd))
        return fail(PRIPPS_ERR_VALIDATION, "volume_scale and contrast_density must be finite");
    const int64_t n = topology->n_atoms;
    if (n < 0) return fail(PRIPPS_ERR_VALIDATION, "negative size");
    if (n_frames > 0 && n > 0 && !coords) return fail(PRIPPS_ERR_USAGE, "coords is NULL");
    if (grid->n_groups > 0 && !out_total) return fail(PRIPPS_ERR_USAGE, "out_total is NULL");
    if (out_self && grid->num_molecules <= 0)
        return fail(PRIPPS_ERR_USAGE, "the self term needs num_molecules > 0");
    // weights: the total of all must be usable (read_weights), and so must the
    // total of the frames the stride keeps
    if (w) {
        double sum = 0;
        for (int32_t f = 0; f < n_frames; ++f) sum += w[f];
        if (!std::isnormal(std::fabs(sum)))
            return fail(PRIPPS_ERR_VALIDATION, "sum of weights is zero or infinite");
    }
    std::vector<int32_t> used;
    double wsum = 0;
    for (int32_t f = 0; f < n_frames; f += stride) {
        used.push_back(f);
        wsum += w ? w[f] : 1.0;
    }
    if (!std::isnormal(std::fabs(wsum))) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%g", wsum);
        return fail(PRIPPS_ERR_VALIDATION, "weights of the " + std::to_string(used.size()) +
                                               " averaged frames sum to " + buf +
                                               "; cannot normalize");
    }
    for (size_t k = 0; k < (size_t)n_frames * n * 3; ++k)
        if (!std::isfinite(coords[k])) return fail(PRIPPS_ERR_VALIDATION, "coordinates must be finite");

    // frame 0 of the used set builds the model (plan, tables, first amplitudes)
    std::vector<double> fx(n), fy(n), fz(n);
    auto soa = [&](int32_t f) {
        const double *c = coords + (size_t)f * n * 3;
        for (int64_t i = 0; i < n; ++i) {
            fx[i] = c[3 * i];
            fy[i] = c[3 * i + 1];
            fz[i] = c[3 * i + 2];
        }
    };
    soa(used[0]);
    pripps_debye_input in0 = *topology;
    in0.x = fx.data();
    in0.y = fy.data();
    in0.z = fz.data();
    pripps_direct_model *m = nullptr;
    int rc = pripps_direct_model_build(engine, &in0, grid, precision, &m);
    if (rc) return rc;
    struct Guard {
        pripps_direct_model *m;
        double *acc = nullptr;
        ~Guard() {
            if (acc) {
                (void)hipSetDevice(m->acc.device);
                (void)hipFree(acc);
            }
            pripps_direct_model_free(m);
        }
    } guard{m};
    const int V = m->V, G = m->G;
    std::vector<double> hout((size_t)2 * G, 0.0);
    if (V > 0) {
        std::lock_guard<std::mutex> lock(*m->acc.mu);
        hipStream_t st = m->acc.stream;
        DCHK(hipSetDevice(m->acc.device));
        DCHK(hipMalloc(&guard.acc, (size_t)V * 2 * sizeof(double)));
        hipLaunchKernelGGL(direct_fill_double_kernel, dim3((2 * V + 255) / 256), dim3(256), 0, st,
                           guard.acc, (int64_t)2 * V);
        DCHK(hipGetLastError());
        const size_t esz = m->fp64 ? sizeof(double) : sizeof(float);
        std::vector<char> host((size_t)std::max<int64_t>(n, 1) * 4 * esz);
        for (size_t k = 0; k < used.size(); ++k) {
            const int32_t f = used[k];
            if (k > 0) {
                soa(f);
                pripps_debye_input inf = *topology;
                inf.x = fx.data();
                inf.y = fy.data();
                inf.z = fz.data();
                if (m->fp64) stage_atoms<double>(&inf, m->L, m->order, host.data());
                else stage_atoms<float>(&inf, m->L, m->order, host.data());
                DCHK(hipMemcpyAsync(m->d_atoms, host.data(), (size_t)n * 4 * esz,
                                    hipMemcpyHostToDevice, st));
                DCHK(enqueue(*m, st, nullptr, nullptr));
            }
            hipLaunchKernelGGL(direct_frames_accumulate_kernel, dim3((V + 255) / 256), dim3(256), 0,
                               st, m->d_amp, V, vs, cd, w ? w[f] : 1.0, guard.acc);
            DCHK(hipGetLastError());
            // the host buffer is rewritten for the next frame
            DCHK(hipStreamSynchronize(st));
        }
        if (G > 0) {
            hipLaunchKernelGGL(direct_frames_group_kernel, dim3(G), dim3(kGroupThreads), 0, st,
                               guard.acc, m->gargs.group_ptr, m->gargs.group_vec, G, wsum, m->d_out);
            DCHK(hipGetLastError());
            DCHK(hipMemcpyAsync(hout.data(), m->d_out, hout.size() * sizeof(double),
                                hipMemcpyDeviceToHost, st));
        }
        DCHK(hipStreamSynchronize(st));
    }
    if (G > 0) std::copy(hout.begin(), hout.begin() + G, out_total);
    if (out_self) std::copy(hout.begin() + G, hout.end(), out_self);
    return PRIPPS_OK;
}
