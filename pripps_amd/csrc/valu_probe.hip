// VALU issue-rate probe for gfx950: what one SIMD sustains for the
// instruction shapes the Debye and Direct kernels are made of.  Used to put
// the kernels' measured issue rates (rocprofv3 SQ_INSTS_VALU / GRBM_GUI_ACTIVE)
// against what is achievable, not against a nominal peak.
//
// Every kernel runs independent chains per lane (no memory traffic but one
// store per thread at the end) and prints wave instructions per SIMD-cycle at
// the nominal 2.4 GHz clock.  Under `rocprofv3 --pmc SQ_INSTS_VALU
// GRBM_GUI_ACTIVE` the same runs give the rate per actual cycle.
//
//   fma / add (SGPR)    v_fma_f32 / v_add_f32 with a scalar-register operand
//   add / fmac (VGPR)   the same with vector-register operands only
//   recurrence x4       acc += s; s += d; d = fma(-lam, s, d), four chains:
//                       the Debye inner loop
//   fma literal / const literal and inline-constant operands
//   sin, mix            v_sin_f32 chains, and 2 transcendentals + 6 FMAs: the
//                       Direct kernel's mix
//   recurrence x1 / x2  one or two dependent chains per lane (the dual
//                       kernel's dry loop runs one pair, i.e. one chain)
//   occupancy sweep     add (VGPR) and the recurrences at 1..4 waves per SIMD,
//                       limited with dynamic LDS
#include <hip/hip_runtime.h>

#include <cstdio>

namespace {

constexpr int CH = 8;
constexpr int IT = 4096;
constexpr int kBlocks = 8192;

#define PROBE_KERNEL(name) __global__ void __launch_bounds__(256) name(float *out, float a, float b)

PROBE_KERNEL(k_fma_s) {
    float x[CH];
    for (int c = 0; c < CH; ++c) x[c] = threadIdx.x * 1e-3f + c;
    for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = __builtin_fmaf(x[c], a, b);
    float s = 0;
    for (int c = 0; c < CH; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

PROBE_KERNEL(k_add_s) {
    float x[CH];
    for (int c = 0; c < CH; ++c) x[c] = threadIdx.x * 1e-3f + c;
    for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = x[c] + a;
    float s = 0;
    for (int c = 0; c < CH; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s + b;
}

PROBE_KERNEL(k_add_v) {
    float x[CH], y[CH];
    for (int c = 0; c < CH; ++c) {
        x[c] = threadIdx.x * 1e-3f + c;
        y[c] = threadIdx.x * 1e-6f + c * a;
    }
    for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = x[c] + y[c];
    float s = 0;
    for (int c = 0; c < CH; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s + b;
}

PROBE_KERNEL(k_fmac_v) {
    float x[CH], y[CH], z[CH];
    for (int c = 0; c < CH; ++c) {
        x[c] = threadIdx.x * 1e-3f + c;
        y[c] = a + c * 1e-7f;
        z[c] = b * c;
    }
    for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = __builtin_fmaf(y[c], z[c], x[c]);
    float s = 0;
    for (int c = 0; c < CH; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

PROBE_KERNEL(k_rec) {
    float s[4], d[4], l[4], acc[4];
    for (int c = 0; c < 4; ++c) {
        s[c] = threadIdx.x * 1e-3f + c;
        d[c] = 1e-3f * c;
        l[c] = a * 1e-4f * (c + 1);
        acc[c] = 0;
    }
    for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            acc[c] += s[c];
            s[c] += d[c];
            d[c] = __builtin_fmaf(-l[c], s[c], d[c]);
        }
    float t = 0;
    for (int c = 0; c < 4; ++c) t += acc[c] + s[c];
    out[blockIdx.x * 256 + threadIdx.x] = t + b;
}

// the recurrence with NC independent chains per lane (one chain = one pair of
// the dual kernel's dry loop)
template <int NC>
__global__ void __launch_bounds__(256) k_rec_n(float *out, float a, float b) {
    float s[NC], d[NC], l[NC], acc[4];
    for (int c = 0; c < NC; ++c) {
        s[c] = threadIdx.x * 1e-3f + c;
        d[c] = 1e-3f * c;
        l[c] = a * 1e-4f * (c + 1);
    }
    for (int u = 0; u < 4; ++u) acc[u] = 0;
    for (int i = 0; i < IT / NC; ++i)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc[u] += s[c];
                s[c] += d[c];
                d[c] = __builtin_fmaf(-l[c], s[c], d[c]);
            }
    float t = 0;
    for (int c = 0; c < NC; ++c) t += s[c];
    for (int u = 0; u < 4; ++u) t += acc[u];
    out[blockIdx.x * 256 + threadIdx.x] = t + b;
}

PROBE_KERNEL(k_fma_lit) {
    float x[CH], y[CH];
    for (int c = 0; c < CH; ++c) {
        x[c] = threadIdx.x * 1e-3f + c;
        y[c] = a + c * 1e-7f;
    }
    for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = __builtin_fmaf(x[c], y[c], 0.12345f);
    float s = 0;
    for (int c = 0; c < CH; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s + b;
}

PROBE_KERNEL(k_fma_ic) {
    float x[CH], y[CH];
    for (int c = 0; c < CH; ++c) {
        x[c] = threadIdx.x * 1e-3f + c;
        y[c] = a + c * 1e-7f;
    }
    for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = __builtin_fmaf(x[c], y[c], 2.0f);
    float s = 0;
    for (int c = 0; c < CH; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s + b;
}

PROBE_KERNEL(k_sin) {
    float x[CH];
    for (int c = 0; c < CH; ++c) x[c] = threadIdx.x * 1e-3f + c * 0.1f;
    for (int i = 0; i < IT; ++i)
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = __builtin_amdgcn_sinf(x[c]) * a;
    float s = 0;
    for (int c = 0; c < CH; ++c) s += x[c];
    out[blockIdx.x * 256 + threadIdx.x] = s + b;
}

PROBE_KERNEL(k_mix) {
    float x[CH], y[CH];
    for (int c = 0; c < CH; ++c) {
        x[c] = threadIdx.x * 1e-3f + c * 0.1f;
        y[c] = c;
    }
    for (int i = 0; i < IT / 4; ++i)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const float s1 = __builtin_amdgcn_sinf(x[c]);
            const float s2 = __builtin_amdgcn_cosf(x[c]);
            y[c] = __builtin_fmaf(s1, a, y[c]);
            y[c] = __builtin_fmaf(s2, b, y[c]);
            x[c] = __builtin_fmaf(x[c], a, b);
            x[c] = __builtin_fmaf(x[c], a, b);
            x[c] = __builtin_fmaf(x[c], a, b);
            x[c] = __builtin_fmaf(x[c], a, b);
        }
    float s = 0;
    for (int c = 0; c < CH; ++c) s += x[c] + y[c];
    out[blockIdx.x * 256 + threadIdx.x] = s;
}

struct Probe {
    const char *name;
    void (*f)(float *, float, float);
    double wave_instr_per_thread;  // VALU instructions of the timed loop per thread
};

int run(const Probe &p, float *out, size_t lds, const char *tag) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    float ms = 0.f;
    for (int rep = 0; rep < 3; ++rep) {
        if (hipEventRecord(e0) != hipSuccess) return 1;
        hipLaunchKernelGGL(p.f, dim3(kBlocks), dim3(256), lds, 0, out, 0.999f, 1e-3f);
        if (hipGetLastError() != hipSuccess) return 1;
        if (hipEventRecord(e1) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return 1;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) return 1;
    }
    const double wave_instr = p.wave_instr_per_thread * kBlocks * 4.0;  // 4 waves per block
    printf("%-20s %-14s %8.3f ms  %.3f wave-instr per SIMD-cycle at 2.4 GHz\n", p.name, tag, ms,
           wave_instr / 1024.0 / (ms * 1e-3 * 2.4e9));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

}  // namespace

int main() {
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * kBlocks * 256) != hipSuccess) return 1;
    const double chain = (double)IT * CH;
    const Probe probes[] = {
        {"fma (SGPR)", k_fma_s, chain},          {"add (SGPR)", k_add_s, chain},
        {"add (VGPR)", k_add_v, chain},          {"fmac (VGPR)", k_fmac_v, chain},
        {"recurrence x4", k_rec, 3.0 * 4 * IT},  {"fma literal", k_fma_lit, chain},
        {"fma inline const", k_fma_ic, chain},   {"sin", k_sin, 2.0 * chain},
        {"mix 2 trans + 6 fma", k_mix, 8.0 * CH * (IT / 4)},
        {"recurrence x1", k_rec_n<1>, 3.0 * 4 * IT},
        {"recurrence x2", k_rec_n<2>, 3.0 * 4 * IT},
    };
    int rc = 0;
    for (const auto &p : probes) rc |= run(p, out, 0, "8 waves/SIMD");
    // dynamic LDS per 256-thread block (one wave per SIMD) limits blocks per CU
    const struct { const char *tag; size_t lds; } occ[] = {
        {"1 wave/SIMD", 96 * 1024}, {"2 waves/SIMD", 64 * 1024}, {"3 waves/SIMD", 48 * 1024},
        {"4 waves/SIMD", 36 * 1024}};
    for (const auto &o : occ)
        for (const Probe &p : {probes[2], probes[4], probes[9], probes[10]}) {
            if (hipFuncSetAttribute(reinterpret_cast<const void *>(p.f),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)o.lds) != hipSuccess)
                return 1;
            rc |= run(p, out, o.lds, o.tag);
        }
    (void)hipFree(out);
    return rc;
}
