/*
 * CPU oracle for the fit half of the Debye API -- TEST INFRASTRUCTURE ONLY.
 *
 * From-scratch FP64 restatement of the reference's src/fit.rs, used only by
 * tests/ to check the engine's device-side fit.  Functions and the reference
 * code they restate:
 *   spline_init / spline_eval   CubicSpline::new / eval   src/fit.rs:149-189
 *   oracle_compute_scale        compute_scale             src/fit.rs:194-205
 *   oracle_chi2_per_point       chi2_per_point            src/fit.rs:209-234
 *   nelder_mead_2d              nelder_mead_2d            src/fit.rs:263-326
 *   bounded_offset              bounded_offset            src/fit.rs:328-341
 *   oracle_initial_simplex      initial_simplex           src/fit.rs:343-351
 *   oracle_fit_nelder_mead      fit_parameters_with_bounds      src/fit.rs:365-411
 *   oracle_fit_grid             fit_parameters_grid_with_bounds src/fit.rs:446-516
 *   validate_bounds             validate_bounds           src/fit.rs:533-545
 * The theory curve is supplied by a callback (c1, c2) -> total[nq], the
 * analogue of the reference's `Fn(f64, f64) -> Intensity` closure.
 *
 * Pinned by tests/test_fit_oracle.py against the reference's own fit tests
 * (src/fit.rs:566-704): spline on linear data, grid minimum (1.25, 0.04)
 * with chi2 < 1e-24, custom c2 bounds (1.25, 1.2), bounds validation,
 * initial simplex inside custom bounds.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef void (*oracle_theory_fn)(double c1, double c2, void *ctx, double *total_out);

typedef struct {
    const double *xs, *ys;
    int64_t n;
    double *y2;
} spline_t;

static void spline_init(spline_t *s, const double *xs, const double *ys, int64_t n) {
    s->xs = xs;
    s->ys = ys;
    s->n = n;
    s->y2 = calloc((size_t)(n ? n : 1), sizeof(double));
    if (n < 3) return;
    double *u = calloc((size_t)n, sizeof(double));
    for (int64_t i = 1; i < n - 1; ++i) {
        double sig = (xs[i] - xs[i - 1]) / (xs[i + 1] - xs[i - 1]);
        double p = sig * s->y2[i - 1] + 2.0;
        s->y2[i] = (sig - 1.0) / p;
        u[i] = (ys[i + 1] - ys[i]) / (xs[i + 1] - xs[i]) - (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
        u[i] = (6.0 * u[i] / (xs[i + 1] - xs[i - 1]) - sig * u[i - 1]) / p;
    }
    for (int64_t i = n - 2; i >= 0; --i) s->y2[i] = s->y2[i] * s->y2[i + 1] + u[i];
    free(u);
}

static double spline_eval(const spline_t *s, double x) {
    const double *xs = s->xs, *ys = s->ys;
    int64_t n = s->n;
    if (x <= xs[0]) return ys[0];
    if (x >= xs[n - 1]) return ys[n - 1];
    /* partition_point(|v| v < x): first index with xs[i] >= x */
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (xs[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    int64_t i = lo;
    if (i == 0) return ys[0];
    double x0 = xs[i - 1], x1 = xs[i], y0 = ys[i - 1], y1 = ys[i];
    double d0 = s->y2[i - 1], d1 = s->y2[i];
    double h = x1 - x0;
    double a = (x1 - x) / h;
    double b = (x - x0) / h;
    return a * y0 + b * y1 + ((a * a * a - a) * d0 + (b * b * b - b) * d1) * h * h / 6.0;
}

/* Spline evaluation exported for tests. */
void oracle_spline_eval(const double *xs, const double *ys, int64_t n, const double *x,
                        int64_t nx, double *out) {
    spline_t s;
    spline_init(&s, xs, ys, n);
    for (int64_t k = 0; k < nx; ++k) out[k] = spline_eval(&s, x[k]);
    free(s.y2);
}

double oracle_compute_scale(const double *tq, const double *ttot, int64_t nq, const double *q,
                            const double *inten, const double *sigma, int64_t n) {
    spline_t s;
    spline_init(&s, tq, ttot, nq);
    double num = 0.0, den = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        double ith = spline_eval(&s, q[k]);
        double s2 = sigma[k] * sigma[k];
        num = num + ith * inten[k] / s2;
        den = den + ith * ith / s2;
    }
    free(s.y2);
    return num / den;
}

double oracle_chi2_per_point(const double *tq, const double *ttot, int64_t nq, const double *q,
                             const double *inten, const double *sigma, int64_t n) {
    spline_t s;
    spline_init(&s, tq, ttot, nq);
    double *ith = malloc(sizeof(double) * (size_t)(n ? n : 1));
    for (int64_t k = 0; k < n; ++k) ith[k] = spline_eval(&s, q[k]);
    double num = 0.0, den = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        double s2 = sigma[k] * sigma[k];
        num = num + ith[k] * inten[k] / s2;
        den = den + ith[k] * ith[k] / s2;
    }
    double scale = num / den;
    double chi2 = 0.0;
    for (int64_t k = 0; k < n; ++k) {
        double r = (scale * ith[k] - inten[k]) / sigma[k];
        chi2 += r * r;
    }
    free(ith);
    free(s.y2);
    return chi2 / (double)n;
}

typedef struct {
    oracle_theory_fn fn;
    void *ctx;
    const double *tq;
    int64_t nq;
    double *buf;
    const double *q, *inten, *sigma;
    int64_t n;
} cost_ctx;

static double cost(cost_ctx *c, double c1, double c2) {
    c->fn(c1, c2, c->ctx, c->buf);
    return oracle_chi2_per_point(c->tq, c->buf, c->nq, c->q, c->inten, c->sigma, c->n);
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

static int validate_bounds(double lo, double hi) {
    if (!isfinite(lo) || !isfinite(hi)) return -1; /* "bounds must be finite" */
    if (lo >= hi) return -2;                         /* "lower bound must be less than upper" */
    return 0;
}

static double bounded_offset(double start, double lo, double hi, double preferred) {
    double delta = fmin(fabs(preferred), (hi - lo) / 2.0);
    if (preferred >= 0.0) return start + delta <= hi ? start + delta : start - delta;
    return start - delta >= lo ? start - delta : start + delta;
}

void oracle_initial_simplex(const double *c1b, const double *c2b, double *out6) {
    double c1 = clampd(1.0, c1b[0], c1b[1]);
    double c2 = clampd(0.03, c2b[0], c2b[1]);
    out6[0] = c1;
    out6[1] = c2;
    out6[2] = bounded_offset(c1, c1b[0], c1b[1], 0.08);
    out6[3] = c2;
    out6[4] = c1;
    out6[5] = bounded_offset(c2, c2b[0], c2b[1], -0.015);
}

typedef struct { double x, y, f; } vtx;

/* Stable sort of three vertices by f (Rust slice::sort_by is stable). */
static void sort3(vtx *s) {
    for (int i = 1; i < 3; ++i) {
        vtx t = s[i];
        int j = i - 1;
        while (j >= 0 && s[j].f > t.f) {
            s[j + 1] = s[j];
            --j;
        }
        s[j + 1] = t;
    }
}

static vtx nm_eval(cost_ctx *c, double x, double y, const double *b1, const double *b2) {
    vtx v;
    v.x = clampd(x, b1[0], b1[1]);
    v.y = clampd(y, b2[0], b2[1]);
    v.f = cost(c, v.x, v.y);
    return v;
}

/* Returns 0 on success, or a negative validation code. out4 = c1, c2, chi2, scale. */
int oracle_fit_nelder_mead(oracle_theory_fn fn, void *ctx, const double *tq, int64_t nq,
                           const double *q, const double *inten, const double *sigma, int64_t n,
                           const double *c1b, const double *c2b, double *out4) {
    int rc = validate_bounds(c1b[0], c1b[1]);
    if (rc) return rc;
    rc = validate_bounds(c2b[0], c2b[1]);
    if (rc) return rc - 10;
    cost_ctx c = {fn, ctx, tq, nq, malloc(sizeof(double) * (size_t)nq), q, inten, sigma, n};
    double init[6];
    oracle_initial_simplex(c1b, c2b, init);
    vtx s[3];
    for (int k = 0; k < 3; ++k) s[k] = nm_eval(&c, init[2 * k], init[2 * k + 1], c1b, c2b);
    for (int it = 0; it < 500; ++it) {
        sort3(s);
        if (s[2].f - s[0].f < 1e-10) break;
        double cx = 0.5 * (s[0].x + s[1].x), cy = 0.5 * (s[0].y + s[1].y);
        vtx r = nm_eval(&c, 2.0 * cx - s[2].x, 2.0 * cy - s[2].y, c1b, c2b);
        if (r.f < s[0].f) {
            vtx e = nm_eval(&c, 3.0 * cx - 2.0 * s[2].x, 3.0 * cy - 2.0 * s[2].y, c1b, c2b);
            s[2] = e.f < r.f ? e : r;
        } else if (r.f < s[1].f) {
            s[2] = r;
        } else {
            vtx k = nm_eval(&c, 0.5 * (cx + s[2].x), 0.5 * (cy + s[2].y), c1b, c2b);
            if (k.f < s[2].f) {
                s[2] = k;
            } else {
                for (int i = 1; i < 3; ++i)
                    s[i] = nm_eval(&c, 0.5 * (s[i].x + s[0].x), 0.5 * (s[i].y + s[0].y), c1b, c2b);
            }
        }
    }
    sort3(s);
    fn(s[0].x, s[0].y, ctx, c.buf);
    out4[0] = s[0].x;
    out4[1] = s[0].y;
    out4[2] = s[0].f;
    out4[3] = oracle_compute_scale(tq, c.buf, nq, q, inten, sigma, n);
    free(c.buf);
    return 0;
}

int oracle_fit_grid(oracle_theory_fn fn, void *ctx, const double *tq, int64_t nq, const double *q,
                    const double *inten, const double *sigma, int64_t n, int64_t c1_points,
                    int64_t c2_points, const double *c1b, const double *c2b, double *out4) {
    if (c1_points < 2 || c2_points < 2) return -20; /* "at least 2 points per axis" */
    int rc = validate_bounds(c1b[0], c1b[1]);
    if (rc) return rc;
    rc = validate_bounds(c2b[0], c2b[1]);
    if (rc) return rc - 10;
    cost_ctx c = {fn, ctx, tq, nq, malloc(sizeof(double) * (size_t)nq), q, inten, sigma, n};
    int have = 0;
    double bc1 = 0, bc2 = 0, bchi = 0;
    for (int64_t i = 0; i < c1_points; ++i) {
        double c1 = c1b[0] + (c1b[1] - c1b[0]) * (double)i / (double)(c1_points - 1);
        for (int64_t j = 0; j < c2_points; ++j) {
            double c2 = c2b[0] + (c2b[1] - c2b[0]) * (double)j / (double)(c2_points - 1);
            double chi = cost(&c, c1, c2);
            if (!isfinite(chi)) continue;
            if (!have || chi < bchi) {
                have = 1;
                bc1 = c1;
                bc2 = c2;
                bchi = chi;
            }
        }
    }
    if (!have) {
        free(c.buf);
        return -30; /* "grid fit did not find any finite chi2 value" */
    }
    fn(bc1, bc2, ctx, c.buf);
    out4[0] = bc1;
    out4[1] = bc2;
    out4[2] = bchi;
    out4[3] = oracle_compute_scale(tq, c.buf, nq, q, inten, sigma, n);
    free(c.buf);
    return 0;
}
