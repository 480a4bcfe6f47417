// Device-side fit of (volume_scale, contrast_density) against experimental
// data, FP64 throughout.  Restates the reference's src/fit.rs:
//   natural cubic spline (CubicSpline::new / eval, fit.rs:149-189),
//   analytic scale and chi^2/N (compute_scale, chi2_per_point, fit.rs:194-234),
//   box-clamped Nelder-Mead (nelder_mead_2d, fit.rs:263-326, initial_simplex
//   343-351), exhaustive grid (fit_parameters_grid_with_bounds, fit.rs:446-516).
//
// The natural spline is linear in its ordinates, so each of the six cached
// profiles is mapped to the experimental q once (spline_map_kernel); the
// theory curve at any (c1, c2) is then the DebyeTerms::combine of the six
// mapped vectors.  This equals the reference's spline of the combined total
// up to rounding, and keeps every cost evaluation on the device.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>
#include <vector>

#include "internal.h"
#include "pripps_debye.h"

namespace pripps {

// y2 of the natural spline through (q, curve_p), one thread per curve.
__global__ void spline_y2_kernel(const double *__restrict__ prof, const double *__restrict__ xs,
                                 int n, int n_curves, double *__restrict__ y2all,
                                 double *__restrict__ uall) {
    const int p = threadIdx.x;
    if (p >= n_curves) return;
    const double *ys = prof + (size_t)p * n;
    double *y2 = y2all + (size_t)p * n;
    double *u = uall + (size_t)p * n;
    for (int i = 0; i < n; ++i) {
        y2[i] = 0.0;
        u[i] = 0.0;
    }
    if (n < 3) return;
    for (int i = 1; i < n - 1; ++i) {
        const double sig = (xs[i] - xs[i - 1]) / (xs[i + 1] - xs[i - 1]);
        const double pp = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / pp;
        double ui = (ys[i + 1] - ys[i]) / (xs[i + 1] - xs[i]) - (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
        u[i] = (6.0 * ui / (xs[i + 1] - xs[i - 1]) - sig * u[i - 1]) / pp;
    }
    for (int i = n - 2; i >= 0; --i) y2[i] = y2[i] * y2[i + 1] + u[i];
}

// M[p][e] = spline_p(qe[e]) for n_curves curves.
__global__ void spline_map_kernel(const double *__restrict__ prof, const double *__restrict__ y2all,
                                  const double *__restrict__ xs, int n, int n_curves,
                                  const double *__restrict__ qe, int ne, double *__restrict__ M) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const double x = qe[e];
    int mode;  // 0 below, 1 above, 2 inside
    int i = 0;
    if (x <= xs[0]) mode = 0;
    else if (x >= xs[n - 1]) mode = 1;
    else {
        int lo = 0, hi = n;  // partition_point(|v| v < x)
        while (lo < hi) {
            const int mid = lo + (hi - lo) / 2;
            if (xs[mid] < x) lo = mid + 1;
            else hi = mid;
        }
        i = lo;
        mode = i == 0 ? 0 : 2;
    }
    for (int p = 0; p < n_curves; ++p) {
        const double *ys = prof + (size_t)p * n;
        const double *y2 = y2all + (size_t)p * n;
        double v;
        if (mode == 0) v = ys[0];
        else if (mode == 1) v = ys[n - 1];
        else {
            const double x0 = xs[i - 1], x1 = xs[i];
            const double h = x1 - x0;
            const double a = (x1 - x) / h;
            const double b = (x - x0) / h;
            v = a * ys[i - 1] + b * ys[i] +
                ((a * a * a - a) * y2[i - 1] + (b * b * b - b) * y2[i]) * h * h / 6.0;
        }
        M[(size_t)p * ne + e] = v;
    }
}

// Theory at experimental point e: DebyeTerms::combine of the mapped profiles.
__device__ __forceinline__ double theory_at(const double *__restrict__ M, int ne, int e, double c1,
                                            double c2) {
#pragma clang fp contract(off)
    const double aa = M[e], ee = M[ne + e], hh = M[2 * ne + e];
    const double ae = M[3 * ne + e], ah = M[4 * ne + e], eh = M[5 * ne + e];
    const double t1 = aa + c1 * c1 * ee;
    const double t2 = t1 + c2 * c2 * hh;
    const double t3 = t2 - 2.0 * c1 * ae;
    const double t4 = t3 + 2.0 * c2 * ah;
    return t4 - 2.0 * c1 * c2 * eh;
}

struct ExpView {
    const double *M;
    const double *I;
    const double *sig;
    int ne;
};

// chi^2/N with the analytic scale, one thread, sequential sums in the
// reference's order (chi2_per_point, fit.rs:209-234).
__device__ double chi2_serial(const ExpView &v, double c1, double c2, double *scale_out) {
#pragma clang fp contract(off)
    double num = 0.0, den = 0.0;
    for (int e = 0; e < v.ne; ++e) {
        const double ith = theory_at(v.M, v.ne, e, c1, c2);
        const double s2 = v.sig[e] * v.sig[e];
        num = num + ith * v.I[e] / s2;
        den = den + ith * ith / s2;
    }
    const double scale = num / den;
    double chi2 = 0.0;
    for (int e = 0; e < v.ne; ++e) {
        const double ith = theory_at(v.M, v.ne, e, c1, c2);
        const double r = (scale * ith - v.I[e]) / v.sig[e];
        chi2 = chi2 + r * r;
    }
    if (scale_out) *scale_out = scale;
    return chi2 / (double)v.ne;
}

__global__ void fit_grid_kernel(ExpView v, int c1p, int c2p, double c1lo, double c1hi, double c2lo,
                                double c2hi, double *__restrict__ chi2) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= c1p * c2p) return;
    const int i = g / c2p, j = g % c2p;
    const double c1 = c1lo + (c1hi - c1lo) * (double)i / (double)(c1p - 1);
    const double c2 = c2lo + (c2hi - c2lo) * (double)j / (double)(c2p - 1);
    chi2[g] = chi2_serial(v, c1, c2, nullptr);
}

// First finite minimum in row-major (c1-outer) order, as the reference's
// strict `<` scan; then the scale at that point.  One block of 256.
__global__ void fit_grid_pick_kernel(ExpView v, const double *__restrict__ chi2, int c1p, int c2p,
                                     double c1lo, double c1hi, double c2lo, double c2hi,
                                     double *__restrict__ res) {
    __shared__ double bv[256];
    __shared__ int bi[256];
    const int t = threadIdx.x;
    double best = 0.0;
    int besti = -1;
    for (int g = t; g < c1p * c2p; g += 256) {
        const double c = chi2[g];
        if (!isfinite(c)) continue;
        if (besti < 0 || c < best) {
            best = c;
            besti = g;
        }
    }
    bv[t] = best;
    bi[t] = besti;
    __syncthreads();
    if (t == 0) {
        double b = 0.0;
        int idx = -1;
        for (int k = 0; k < 256; ++k) {
            if (bi[k] < 0) continue;
            if (idx < 0 || bv[k] < b || (bv[k] == b && bi[k] < idx)) {
                b = bv[k];
                idx = bi[k];
            }
        }
        if (idx < 0) {
            res[4] = -1.0;
            return;
        }
        const int i = idx / c2p, j = idx % c2p;
        const double c1 = c1lo + (c1hi - c1lo) * (double)i / (double)(c1p - 1);
        const double c2 = c2lo + (c2hi - c2lo) * (double)j / (double)(c2p - 1);
        double scale = 0.0;
        chi2_serial(v, c1, c2, &scale);
        res[0] = c1;
        res[1] = c2;
        res[2] = b;
        res[3] = scale;
        res[4] = 0.0;
    }
}

__device__ __forceinline__ double clampd(double x, double lo, double hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}

struct Vtx {
    double x, y, f;
};

__device__ Vtx nm_eval(const ExpView &v, double x, double y, double lo1, double hi1, double lo2,
                       double hi2) {
    Vtx r;
    r.x = clampd(x, lo1, hi1);
    r.y = clampd(y, lo2, hi2);
    r.f = chi2_serial(v, r.x, r.y, nullptr);
    return r;
}

__device__ void sort3(Vtx *s) {  // stable insertion sort by f
    for (int i = 1; i < 3; ++i) {
        Vtx t = s[i];
        int j = i - 1;
        while (j >= 0 && s[j].f > t.f) {
            s[j + 1] = s[j];
            --j;
        }
        s[j + 1] = t;
    }
}

__device__ double bounded_offset(double start, double lo, double hi, double pref) {
    const double delta = fmin(fabs(pref), (hi - lo) / 2.0);
    if (pref >= 0.0) return start + delta <= hi ? start + delta : start - delta;
    return start - delta >= lo ? start - delta : start + delta;
}

// Nelder-Mead, box-clamped, <= 500 iterations, tolerance 1e-10 on the
// simplex's chi^2 spread.  Single thread: the control flow is inherently
// sequential and each cost is O(n_exp).
__global__ void fit_nm_kernel(ExpView v, double lo1, double hi1, double lo2, double hi2,
                              double *__restrict__ res) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double c1 = clampd(1.0, lo1, hi1), c2 = clampd(0.03, lo2, hi2);
    Vtx s[3];
    s[0] = nm_eval(v, c1, c2, lo1, hi1, lo2, hi2);
    s[1] = nm_eval(v, bounded_offset(c1, lo1, hi1, 0.08), c2, lo1, hi1, lo2, hi2);
    s[2] = nm_eval(v, c1, bounded_offset(c2, lo2, hi2, -0.015), lo1, hi1, lo2, hi2);
    for (int it = 0; it < 500; ++it) {
        sort3(s);
        if (s[2].f - s[0].f < 1e-10) break;
        const double cx = 0.5 * (s[0].x + s[1].x), cy = 0.5 * (s[0].y + s[1].y);
        const Vtx r = nm_eval(v, 2.0 * cx - s[2].x, 2.0 * cy - s[2].y, lo1, hi1, lo2, hi2);
        if (r.f < s[0].f) {
            const Vtx ex = nm_eval(v, 3.0 * cx - 2.0 * s[2].x, 3.0 * cy - 2.0 * s[2].y, lo1, hi1, lo2, hi2);
            s[2] = ex.f < r.f ? ex : r;
        } else if (r.f < s[1].f) {
            s[2] = r;
        } else {
            const Vtx k = nm_eval(v, 0.5 * (cx + s[2].x), 0.5 * (cy + s[2].y), lo1, hi1, lo2, hi2);
            if (k.f < s[2].f) {
                s[2] = k;
            } else {
                for (int i = 1; i < 3; ++i)
                    s[i] = nm_eval(v, 0.5 * (s[i].x + s[0].x), 0.5 * (s[i].y + s[0].y), lo1, hi1, lo2, hi2);
            }
        }
    }
    sort3(s);
    double scale = 0.0;
    chi2_serial(v, s[0].x, s[0].y, &scale);
    res[0] = s[0].x;
    res[1] = s[0].y;
    res[2] = s[0].f;
    res[3] = scale;
    res[4] = 0.0;
}

// Device scratch (doubles) that device_fit needs; the caller owns it.
size_t device_fit_doubles(int nq, int64_t n, int method, int c1p, int c2p) {
    const size_t grid = method == PRIPPS_FIT_GRID ? (size_t)c1p * c2p : 0;
    return (size_t)3 * n + (size_t)12 * nq + (size_t)6 * n + grid + 8;
}

int device_fit(hipStream_t st, double *buf, const double *d_prof, const double *d_q, int nq,
               const double *q, const double *inten, const double *sigma, int64_t n, int method,
               int c1p, int c2p, const double *c1b, const double *c2b, pripps_fit_result *out,
               std::string *err) {
    const int ne = (int)n;
    std::vector<double> host((size_t)3 * ne);
    for (int e = 0; e < ne; ++e) {
        host[e] = q[e];
        host[ne + e] = inten[e];
        host[2 * ne + e] = sigma ? sigma[e] : 1.0;
    }
    const size_t grid = method == PRIPPS_FIT_GRID ? (size_t)c1p * c2p : 0;
    hipError_t e = hipSuccess;
    double *d_exp = buf;
    double *d_y2 = d_exp + host.size();
    double *d_u = d_y2 + (size_t)6 * nq;
    double *d_M = d_u + (size_t)6 * nq;
    double *d_chi = d_M + (size_t)6 * ne;
    double *d_res = d_chi + grid;
    double res[5] = {0, 0, 0, 0, -2.0};
    ExpView v{d_M, d_exp + ne, d_exp + 2 * ne, ne};
    do {
        if ((e = hipMemcpyAsync(d_exp, host.data(), host.size() * 8, hipMemcpyHostToDevice, st)) != hipSuccess) break;
        hipLaunchKernelGGL(spline_y2_kernel, dim3(1), dim3(64), 0, st, d_prof, d_q, nq, kProfiles,
                           d_y2, d_u);
        hipLaunchKernelGGL(spline_map_kernel, dim3((ne + 255) / 256), dim3(256), 0, st, d_prof, d_y2,
                           d_q, nq, kProfiles, (const double *)d_exp, ne, d_M);
        if (method == PRIPPS_FIT_GRID) {
            hipLaunchKernelGGL(fit_grid_kernel, dim3((unsigned)((grid + 255) / 256)), dim3(256), 0, st, v,
                               c1p, c2p, c1b[0], c1b[1], c2b[0], c2b[1], d_chi);
            hipLaunchKernelGGL(fit_grid_pick_kernel, dim3(1), dim3(256), 0, st, v, (const double *)d_chi,
                               c1p, c2p, c1b[0], c1b[1], c2b[0], c2b[1], d_res);
        } else {
            hipLaunchKernelGGL(fit_nm_kernel, dim3(1), dim3(64), 0, st, v, c1b[0], c1b[1], c2b[0],
                               c2b[1], d_res);
        }
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = hipMemcpyAsync(res, d_res, sizeof(res), hipMemcpyDeviceToHost, st)) != hipSuccess) break;
        e = hipStreamSynchronize(st);
    } while (0);
    if (e != hipSuccess) {
        *err = std::string("fit: ") + hipGetErrorString(e);
        return PRIPPS_ERR_DEVICE;
    }
    if (res[4] != 0.0) {
        *err = "grid fit did not find any finite χ² value";
        return PRIPPS_ERR_INTERNAL;
    }
    out->volume_scale = res[0];
    out->contrast_density = res[1];
    out->chi2 = res[2];
    out->scale = res[3];
    return PRIPPS_OK;
}

// The theory curve at experimental q as write_fit_csv evaluates it
// (src/fit.rs:548-560): the natural spline of the combined total I(q) over
// the model's grid, evaluated at qe.  FP64 on the device.
// Device scratch (doubles) that device_spline_total needs; the caller owns it.
size_t device_spline_total_doubles(int nq, int64_t n) {
    return (size_t)4 * nq + (size_t)2 * nq + (size_t)2 * n + 8;
}

int device_spline_total(hipStream_t st, double *buf, const double *d_prof, const double *d_q,
                        int nq, double vs, double cd, const double *qe, int64_t n, double *out,
                        std::string *err) {
    const int ne = (int)n;
    hipError_t e = hipSuccess;
    double *d_tot = buf;             // [4][nq]: combine output, row 0 = total
    double *d_y2 = d_tot + (size_t)4 * nq;
    double *d_u = d_y2 + nq;
    double *d_qe = d_u + nq;
    double *d_out = d_qe + ne;
    do {
        if ((e = hipMemcpyAsync(d_qe, qe, (size_t)ne * 8, hipMemcpyHostToDevice, st)) != hipSuccess) break;
        if ((e = launch_combine(d_prof, nq, vs, cd, d_tot, st)) != hipSuccess) break;
        hipLaunchKernelGGL(spline_y2_kernel, dim3(1), dim3(64), 0, st, (const double *)d_tot, d_q, nq,
                           1, d_y2, d_u);
        hipLaunchKernelGGL(spline_map_kernel, dim3((ne + 255) / 256), dim3(256), 0, st,
                           (const double *)d_tot, (const double *)d_y2, d_q, nq, 1,
                           (const double *)d_qe, ne, d_out);
        if ((e = hipGetLastError()) != hipSuccess) break;
        if ((e = hipMemcpyAsync(out, d_out, (size_t)ne * 8, hipMemcpyDeviceToHost, st)) != hipSuccess) break;
        e = hipStreamSynchronize(st);
    } while (0);
    if (e != hipSuccess) {
        *err = std::string("fit curve: ") + hipGetErrorString(e);
        return PRIPPS_ERR_DEVICE;
    }
    return PRIPPS_OK;
}

}  // namespace pripps
