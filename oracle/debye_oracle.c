/*
 * CPU oracle for the exact Debye path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is a from-scratch FP64 restatement of the reference crate's
 * exact Debye sums (mlund/pripps).  It is the checker for the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library never links or calls it.
 *
 * Functions and the reference code they restate:
 *   oracle_linear_q_values      src/q_sampler.rs:24-29
 *   oracle_sinc                 src/debye.rs:17-24
 *   oracle_self_intensity_at_q  src/debye.rs:214-232
 *   oracle_cross_intensity_at_q src/debye.rs:236-250
 *   oracle_build_model          src/debye.rs:78-136 (per-q body 93-128),
 *                               with the per-atom form factors of
 *                               src/formfactor.rs:528-544 (by kind) and
 *                               src/solvent/mod.rs:357-374 (weighted set)
 *   oracle_combine              src/debye.rs:186-195
 *   oracle_eval_gaussian        src/formfactor.rs:369-375
 *   oracle_eval_single_gaussian src/formfactor.rs:196
 *   oracle_eval_polynomial      src/formfactor.rs:379-385
 *   oracle_interpolate          src/formfactor.rs:389-408
 *   oracle_united_table         src/formfactor.rs:642-698 (one orbit set)
 *   oracle_fraser_params        src/solvent/fraser.rs:48-57
 *
 * Two summation modes for the pair sums:
 *   ORACLE_REFERENCE_ORDER  the reference's own loop nest: three self sums and
 *                           three cross sums per q, one sequential f64
 *                           accumulator each, i-outer / j-inner, libm sin.
 *                           Used for the CPU baseline timing and order parity.
 *   ORACLE_COMPENSATED      one pass over i<j using the colocated-set identity
 *                           sum_ij fa_i fb_j s_ij = sum_i fa_i fb_i
 *                             + sum_{i<j} (fa_i fb_j + fa_j fb_i) s_ij,
 *                           every accumulator a double-double (TwoSum), so the
 *                           only error left is the per-term rounding of sin and
 *                           sqrt.  This is the <=1e-12 truth.
 *
 * Pinning (see tests/test_oracle.py): q grid endpoints and nq=1
 * (src/q_sampler.rs:173-185); q=0 closed form (sum f)^2
 * (src/multipole/mod.rs:907-940); two atoms 2+2 sinc(q r)
 * (src/multipole/mod.rs:942-967); one atom f^2; disabled solvent sets give
 * zero EE/HH/cross; translation/rotation/permutation invariance; the two
 * modes against each other and against mpmath brute force on tiny inputs;
 * Cromer-Mann s^2 value 6 exp(-20*0.25/(16 pi^2)) (src/formfactor.rs:759-774);
 * table interpolation values (src/formfactor.rs:729-740); united resolution
 * at q=0 gives 1.4 (src/formfactor.rs:860-891); Fraser F(0)=rho V
 * (src/solvent/fraser.rs:142-152).
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  FMA
 * contraction is off because the reference (Rust) never contracts a*b+c.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_REFERENCE_ORDER 0
#define ORACLE_COMPENSATED 1

/* Profile order follows DebyeTerms (src/debye.rs:176-184). */
enum { P_AA = 0, P_EE = 1, P_HH = 2, P_AE = 3, P_AH = 4, P_EH = 5, NPROF = 6 };

void oracle_linear_q_values(double qmin, double qmax, int64_t nq, double *out) {
    /* q_i = qmin + (qmax - qmin) * i / max(nq - 1, 1), evaluated exactly in
     * this order: ((qmax - qmin) * i) / denom, then + qmin. */
    int64_t d = nq - 1 > 1 ? nq - 1 : 1;
    double denom = (double)d;
    for (int64_t i = 0; i < nq; ++i) out[i] = qmin + (qmax - qmin) * (double)i / denom;
}

double oracle_sinc(double x) {
    const double nearly_zero = 1e-8;
    if (fabs(x) < nearly_zero) return 1.0;
    return sin(x) / x;
}

static inline double dist3(const double *a, const double *b) {
    double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

double oracle_self_intensity_at_q(const double *pos, const double *ff, int64_t n, double q) {
    double diagonal = 0.0;
    for (int64_t i = 0; i < n; ++i) diagonal += ff[i] * ff[i];
    if (n < 2) return diagonal;
    double off = 0.0;
    for (int64_t i = 0; i < n - 1; ++i) {
        double fi = ff[i];
        const double *pi = pos + 3 * i;
        for (int64_t j = i + 1; j < n; ++j) {
            double r = dist3(pi, pos + 3 * j);
            off += fi * ff[j] * oracle_sinc(q * r);
        }
    }
    return fma(2.0, off, diagonal); /* 2.0f64.mul_add(off_diagonal, diagonal) */
}

double oracle_cross_intensity_at_q(const double *pa, const double *fa, int64_t na,
                                   const double *pb, const double *fb, int64_t nb, double q) {
    double total = 0.0;
    for (int64_t i = 0; i < na; ++i) {
        for (int64_t j = 0; j < nb; ++j) {
            total += fa[i] * fb[j] * oracle_sinc(q * dist3(pa + 3 * i, pb + 3 * j));
        }
    }
    return total;
}

/* ---- double-double accumulator (Knuth TwoSum) ---- */
typedef struct { double hi, lo; } dd_t;

static inline void dd_add(dd_t *a, double x) {
    double s = a->hi + x;
    double bp = s - a->hi;
    double err = (a->hi - (s - bp)) + (x - bp);
    a->lo += err;
    /* renormalise occasionally is unnecessary: lo stays tiny relative to hi
     * for the term counts used here (< 2^40). */
    a->hi = s;
}

static inline double dd_value(dd_t a) { return a.hi + a.lo; }

static inline void dd_merge(dd_t *a, dd_t b) {
    dd_add(a, b.hi);
    dd_add(a, b.lo);
}

/*
 * Per-atom form factors at q index m for the three colocated sets:
 *   atoms     fA_i = atom_ff[kind_i * nq + m]            (formfactor_by_index)
 *   excluded  fE_i = excl_ff[kind_i * nq + m]  or absent (FoXS/Fraser dummies
 *             depend only on the kind's displaced volume)
 *   hydration fH_i = w_i * water_ff[m]         or absent (DummySet::weighted)
 */
static void fill_ff(const int32_t *kind, int64_t n, int64_t nq, int64_t m,
                    const double *atom_ff, const double *excl_ff, const double *water_ff,
                    const double *weights, double *fa, double *fe, double *fh) {
    for (int64_t i = 0; i < n; ++i) {
        fa[i] = atom_ff[(int64_t)kind[i] * nq + m];
        if (excl_ff) fe[i] = excl_ff[(int64_t)kind[i] * nq + m];
        if (water_ff) fh[i] = weights[i] * water_ff[m];
    }
}

static void build_reference_order(const double *pos, const int32_t *kind, int64_t n,
                                  const double *q, int64_t nq, const double *atom_ff,
                                  const double *excl_ff, const double *water_ff,
                                  const double *weights, int nthreads, double *out) {
    int64_t ne = excl_ff ? n : 0;  /* a disabled set has no dummies at all */
    int64_t nh = water_ff ? n : 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t m = 0; m < nq; ++m) {
        double *fa = malloc(sizeof(double) * (size_t)(n ? n : 1));
        double *fe = malloc(sizeof(double) * (size_t)(n ? n : 1));
        double *fh = malloc(sizeof(double) * (size_t)(n ? n : 1));
        fill_ff(kind, n, nq, m, atom_ff, excl_ff, water_ff, weights, fa, fe, fh);
        double qq = q[m];
        out[P_AA * nq + m] = oracle_self_intensity_at_q(pos, fa, n, qq);
        out[P_EE * nq + m] = oracle_self_intensity_at_q(pos, fe, ne, qq);
        out[P_HH * nq + m] = oracle_self_intensity_at_q(pos, fh, nh, qq);
        out[P_AE * nq + m] = oracle_cross_intensity_at_q(pos, fa, n, pos, fe, ne, qq);
        out[P_AH * nq + m] = oracle_cross_intensity_at_q(pos, fa, n, pos, fh, nh, qq);
        out[P_EH * nq + m] = oracle_cross_intensity_at_q(pos, fe, ne, pos, fh, nh, qq);
        free(fa);
        free(fe);
        free(fh);
    }
}

/* Compensated mode: rows i are distributed over threads; each thread keeps
 * double-double accumulators for every (profile, q); the thread partials
 * are merged in thread order, so the result does not depend on scheduling
 * beyond the last ulp of the double-double. */
static void build_compensated(const double *pos, const int32_t *kind, int64_t n,
                              const double *q, int64_t nq, const double *atom_ff,
                              const double *excl_ff, const double *water_ff,
                              const double *weights, int nthreads, double *out) {
    int has_e = excl_ff != NULL, has_h = water_ff != NULL;
    if (nthreads < 1) nthreads = 1;
    dd_t *part = calloc((size_t)nthreads * NPROF * (size_t)nq, sizeof(dd_t));
#pragma omp parallel num_threads(nthreads)
    {
        int t = 0;
#ifdef _OPENMP
        extern int omp_get_thread_num(void);
        t = omp_get_thread_num();
#endif
        dd_t *acc = part + (size_t)t * NPROF * (size_t)nq;
        double *s = malloc(sizeof(double) * (size_t)(nq ? nq : 1));
        /* diagonal terms (i == j, r = 0, sinc = 1) */
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            for (int64_t m = 0; m < nq; ++m) {
                double a = atom_ff[(int64_t)kind[i] * nq + m];
                double e = has_e ? excl_ff[(int64_t)kind[i] * nq + m] : 0.0;
                double h = has_h ? weights[i] * water_ff[m] : 0.0;
                dd_add(&acc[P_AA * nq + m], a * a);
                dd_add(&acc[P_EE * nq + m], e * e);
                dd_add(&acc[P_HH * nq + m], h * h);
                dd_add(&acc[P_AE * nq + m], a * e);
                dd_add(&acc[P_AH * nq + m], a * h);
                dd_add(&acc[P_EH * nq + m], e * h);
            }
        }
        /* off-diagonal pairs i < j, rows interleaved for balance */
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < n; ++i) {
            const double *pi = pos + 3 * i;
            for (int64_t j = i + 1; j < n; ++j) {
                double r = dist3(pi, pos + 3 * j);
                for (int64_t m = 0; m < nq; ++m) s[m] = oracle_sinc(q[m] * r);
                for (int64_t m = 0; m < nq; ++m) {
                    double ai = atom_ff[(int64_t)kind[i] * nq + m];
                    double aj = atom_ff[(int64_t)kind[j] * nq + m];
                    double ei = has_e ? excl_ff[(int64_t)kind[i] * nq + m] : 0.0;
                    double ej = has_e ? excl_ff[(int64_t)kind[j] * nq + m] : 0.0;
                    double hi = has_h ? weights[i] * water_ff[m] : 0.0;
                    double hj = has_h ? weights[j] * water_ff[m] : 0.0;
                    double sm = s[m];
                    dd_add(&acc[P_AA * nq + m], 2.0 * ai * aj * sm);
                    if (has_e) {
                        dd_add(&acc[P_EE * nq + m], 2.0 * ei * ej * sm);
                        dd_add(&acc[P_AE * nq + m], (ai * ej + aj * ei) * sm);
                    }
                    if (has_h) {
                        dd_add(&acc[P_HH * nq + m], 2.0 * hi * hj * sm);
                        dd_add(&acc[P_AH * nq + m], (ai * hj + aj * hi) * sm);
                    }
                    if (has_e && has_h) dd_add(&acc[P_EH * nq + m], (ei * hj + ej * hi) * sm);
                }
            }
        }
        free(s);
    }
    for (int64_t k = 0; k < NPROF * nq; ++k) {
        dd_t tot = {0.0, 0.0};
        for (int t = 0; t < nthreads; ++t) dd_merge(&tot, part[(size_t)t * NPROF * (size_t)nq + k]);
        out[k] = dd_value(tot);
    }
    free(part);
}

/*
 * Six cached partial profiles, layout out[6][nq] in DebyeTerms order
 * (atoms, excluded, hydration, atom_excluded, atom_hydration,
 * excluded_hydration).  pos is AoS [n][3].  excl_ff == NULL disables the
 * excluded set; water_ff == NULL (or weights == NULL) disables hydration.
 * Returns 0 on success, -1 on bad arguments.
 */
int oracle_build_model(const double *pos, const int32_t *kind, int64_t n, int32_t n_kinds,
                       const double *q, int64_t nq, const double *atom_ff,
                       const double *excl_ff, const double *water_ff, const double *weights,
                       int mode, int nthreads, double *out) {
    if (n < 0 || nq < 0 || (n > 0 && (!pos || !kind || !atom_ff)) || (nq > 0 && !q) || !out)
        return -1;
    for (int64_t i = 0; i < n; ++i)
        if (kind[i] < 0 || kind[i] >= n_kinds) return -1;
    if (!weights) water_ff = NULL;
    if (nthreads < 1) nthreads = 1;
    memset(out, 0, sizeof(double) * NPROF * (size_t)nq);
    if (mode == ORACLE_REFERENCE_ORDER)
        build_reference_order(pos, kind, n, q, nq, atom_ff, excl_ff, water_ff, weights, nthreads, out);
    else
        build_compensated(pos, kind, n, q, nq, atom_ff, excl_ff, water_ff, weights, nthreads, out);
    return 0;
}

/* DebyeTerms::combine (src/debye.rs:186-195), same operation order. */
double oracle_combine(const double *t6, double vs, double cd) {
    return t6[P_AA] + vs * vs * t6[P_EE] + cd * cd * t6[P_HH] - 2.0 * vs * t6[P_AE] +
           2.0 * cd * t6[P_AH] - 2.0 * vs * cd * t6[P_EH];
}

/* ---- form factors ---- */
double oracle_eval_gaussian(double q, const double *a, const double *b, int n, double c) {
    const double pi = 3.14159265358979323846;
    double s2 = q * q / (16.0 * pi * pi);
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += a[i] * exp(-b[i] * s2);
    return sum + c;
}

double oracle_eval_single_gaussian(double a, double b, double q) { return a * exp(-b * q * q); }

/* Rust's f64::powi lowers to binary exponentiation by repeated squaring. */
static double powi_binary(double a, int b) {
    int recip = b < 0;
    double r = 1.0;
    if (recip) b = -b;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

double oracle_eval_polynomial(const double *coeff, int n, double q) {
    double sum = 0.0;
    for (int i = 0; i < n; ++i) sum += coeff[i] * powi_binary(q, i);
    return sum;
}

double oracle_interpolate(const double *xs, const double *ys, int64_t n, double x) {
    int64_t low = 0, high = n - 1;
    while (low < high) {
        int64_t mid = (low + high) / 2;
        if (x < xs[mid]) high = mid;
        else if (x > xs[mid]) low = mid + 1;
        else return ys[mid];
    }
    double x0 = xs[low - 1], y0 = ys[low - 1], x1 = xs[low], y1 = ys[low];
    return y0 + (x - x0) * (y1 - y0) / (x1 - x0);
}

/*
 * United-atom table: 101 points on [qmin, qmax] of the centre form factor
 * plus sum_o weight_o * F_o(q) * sinc(q * offset_o).  Centre and orbit form
 * factors are passed pre-tabulated on the same 101-point grid (the caller
 * evaluates them), so this restates only the grid and the orbit sum.
 */
void oracle_united_grid(double qmin, double qmax, double *qs /*[101]*/) {
    for (int i = 0; i <= 100; ++i) qs[i] = qmin + (qmax - qmin) * (double)i / 100.0;
}

void oracle_united_table(const double *qs, const double *center /*[101]*/, int n_orbit,
                         const double *orbit_ff /*[n_orbit][101]*/, const double *offset,
                         const double *weight, double *out /*[101]*/) {
    for (int i = 0; i <= 100; ++i) out[i] = center[i];
    for (int o = 0; o < n_orbit; ++o)
        for (int i = 0; i <= 100; ++i)
            out[i] += weight[o] * orbit_ff[o * 101 + i] * oracle_sinc(qs[i] * offset[o]);
}

/* Fraser dummy: a = rho * V, b = V^(2/3) / denominator. */
void oracle_fraser_params(double volume, double rho, double denominator, double *a, double *b) {
    *a = rho * volume;
    *b = pow(volume, 2.0 / 3.0) / denominator;
}
