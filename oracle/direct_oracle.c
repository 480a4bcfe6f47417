/*
 * CPU oracle for the Direct Fourier transform -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch restatement of the reference crate's explicit q-vector
 * transform (mlund/pripps).  It is the checker for the HIP Direct kernels:
 * only tests/ and bench.py may load it; the product library never does.
 *
 * Functions and the reference code they restate:
 *   oracle_direct_rows  per q-vector amplitudes and self term:
 *     - q-vector = (2 pi n) ./ L           src/explicit.rs:228, 279-281, 294
 *     - complex_amplitude (sequential fold, q.r, libm cos/sin)
 *                                          src/explicit.rs:136-146
 *     - vacuum rows: molecules are contiguous chunks of n_atoms / N atoms,
 *       A = sum_m F_m, self = sum_m |F_m|^2  src/explicit.rs:217-244
 *     - solvent rows: A, E, H amplitudes   src/explicit.rs:283-326
 *     - PhaseTable (powers of exp(i 2 pi x_a / L_a) by repeated complex
 *       multiplication, conjugated for negative indices)
 *                                          src/explicit.rs:80-133
 *
 * Modes:
 *   ORACLE_DIRECT_REFERENCE  the reference's arithmetic: q.r in radians, libm
 *                            cos/sin, one sequential f64 accumulator per sum.
 *   ORACLE_DIRECT_TABLE      the reference's PhaseTable path.
 *   ORACLE_DIRECT_EXTENDED   the truth for tolerance checks: the phase in
 *                            revolutions n . t with t = x / L reduced mod 1,
 *                            cos/sin and every sum in long double (64-bit
 *                            mantissa on x86-64).
 *
 * Pinning (tests/test_direct_host.py): the reference's
 * phase_table_matches_sin_cos (12 atoms, lattice grid, 4 molecules, 1e-10,
 * src/explicit.rs:480-515) across the three modes; one atom gives f^2 at
 * every vector; the whole-box self term equals |A|^2; N identical molecules
 * at the same position give self = |A|^2 / N; translation by a lattice vector
 * leaves every amplitude unchanged; inversion conjugates them.
 *
 * Output per vector, [n_vectors][7]: Re A, Im A, Re E, Im E, Re H, Im H, self.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_DIRECT_REFERENCE 0
#define ORACLE_DIRECT_TABLE 1
#define ORACLE_DIRECT_EXTENDED 2

static const double kPi = 3.14159265358979323846;

typedef struct {
    double re, im;
} cplx;

static cplx cmul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}

/* exp(i 2 pi n x / L) for n = 0..stride-1 by successive multiplication. */
static void phase_powers(double x, double L, int64_t stride, cplx *out) {
    const double ang = 2.0 * kPi * x / L;
    const cplx step = {cos(ang), sin(ang)};
    cplx p = {1.0, 0.0};
    for (int64_t k = 0; k < stride; ++k) {
        out[k] = p;
        p = cmul(p, step);
    }
}

/* Amplitude over atoms [a0, a1) with per-atom form factors f(atom). */
typedef double (*ff_fn)(const void *ctx, int64_t atom);

struct ff_ctx {
    const int32_t *kind;
    const double *table; /* [K][V] or [V] (shared) */
    const double *weight; /* per-atom multiplier or NULL */
    int32_t V;
    int32_t v;
    int shared; /* 1: table[v] for every atom (water) */
};

static double ff_at(const struct ff_ctx *c, int64_t atom) {
    double f = c->shared ? c->table[c->v] : c->table[(int64_t)c->kind[atom] * c->V + c->v];
    if (c->weight) f = c->weight[atom] * f;
    return f;
}

static void amp_reference(const double *x, const double *y, const double *z, int64_t a0,
                          int64_t a1, const double q[3], const struct ff_ctx *c, double *re,
                          double *im) {
    double r = 0.0, i = 0.0;
    for (int64_t a = a0; a < a1; ++a) {
        const double qr = q[0] * x[a] + q[1] * y[a] + q[2] * z[a];
        const double f = ff_at(c, a);
        r = r + f * cos(qr);
        i = i + f * sin(qr);
    }
    *re = r;
    *im = i;
}

static void amp_table(const cplx *const pw[3], int64_t stride, int64_t a0, int64_t a1,
                      const int32_t n[3], const struct ff_ctx *c, double *re, double *im) {
    cplx s = {0.0, 0.0};
    for (int64_t a = a0; a < a1; ++a) {
        cplx t[3];
        for (int d = 0; d < 3; ++d) {
            const int32_t k = n[d] < 0 ? -n[d] : n[d];
            t[d] = pw[d][a * stride + k];
            if (n[d] < 0) t[d].im = -t[d].im;
        }
        const double f = ff_at(c, a);
        cplx v = {f * t[0].re, f * t[0].im};
        v = cmul(cmul(v, t[1]), t[2]);
        s.re += v.re;
        s.im += v.im;
    }
    *re = s.re;
    *im = s.im;
}

static void amp_extended(const long double *t, int64_t a0, int64_t a1, const int32_t n[3],
                         const struct ff_ctx *c, long double *re, long double *im) {
    long double r = 0.0L, i = 0.0L;
    const long double two_pi = 6.283185307179586476925286766559L;
    for (int64_t a = a0; a < a1; ++a) {
        long double p = (long double)n[0] * t[3 * a] + (long double)n[1] * t[3 * a + 1] +
                        (long double)n[2] * t[3 * a + 2];
        p -= nearbyintl(p);
        const long double f = (long double)ff_at(c, a);
        r += f * cosl(two_pi * p);
        i += f * sinl(two_pi * p);
    }
    *re = r;
    *im = i;
}

/* n_mol: 0 -> whole box (one chunk, the self term is then |A|^2); N > 0 ->
 * chunks of max(n_atoms / N, 1) atoms, as the reference's `chunks(m)`. */
int oracle_direct_rows(int64_t n_atoms, const double *x, const double *y, const double *z,
                       const int32_t *kind, int32_t n_kinds, const double box[3],
                       int32_t n_vectors, const int32_t *indices, const double *atom_ff,
                       const double *excluded_ff, const double *water_ff,
                       const double *hydration_weight, int32_t n_mol, int mode, double *out) {
    (void)n_kinds;
    const int64_t nm = n_mol > 0 ? n_mol : 1;
    int64_t m = n_atoms / nm;
    if (m < 1) m = 1;
    cplx *pw[3] = {NULL, NULL, NULL};
    long double *t = NULL;
    int64_t stride = 0;
    if (mode == ORACLE_DIRECT_TABLE) {
        int32_t nmax = 0;
        for (int64_t k = 0; k < 3 * (int64_t)n_vectors; ++k) {
            const int32_t a = indices[k] < 0 ? -indices[k] : indices[k];
            if (a > nmax) nmax = a;
        }
        stride = nmax + 1;
        for (int d = 0; d < 3; ++d) {
            pw[d] = (cplx *)malloc(sizeof(cplx) * (size_t)(n_atoms * stride + 1));
            if (!pw[d]) return -1;
        }
        for (int64_t a = 0; a < n_atoms; ++a) {
            phase_powers(x[a], box[0], stride, pw[0] + a * stride);
            phase_powers(y[a], box[1], stride, pw[1] + a * stride);
            phase_powers(z[a], box[2], stride, pw[2] + a * stride);
        }
    } else if (mode == ORACLE_DIRECT_EXTENDED) {
        t = (long double *)malloc(sizeof(long double) * (size_t)(3 * n_atoms + 1));
        if (!t) return -1;
        const double *p[3] = {x, y, z};
        for (int64_t a = 0; a < n_atoms; ++a)
            for (int d = 0; d < 3; ++d) {
                const long double u = (long double)p[d][a] / (long double)box[d];
                t[3 * a + d] = u - nearbyintl(u);
            }
    }

#pragma omp parallel for schedule(dynamic, 4)
    for (int32_t v = 0; v < n_vectors; ++v) {
        const int32_t *n = indices + 3 * (int64_t)v;
        double q[3];
        for (int d = 0; d < 3; ++d) q[d] = (2.0 * kPi * (double)n[d]) / box[d];
        struct ff_ctx ca = {kind, atom_ff, NULL, n_vectors, v, 0};
        struct ff_ctx ce = {kind, excluded_ff, NULL, n_vectors, v, 0};
        struct ff_ctx ch = {kind, water_ff, hydration_weight, n_vectors, v, 1};
        double *o = out + 7 * (int64_t)v;
        if (mode == ORACLE_DIRECT_EXTENDED) {
            long double Ar = 0, Ai = 0, self = 0, er, ei;
            for (int64_t a0 = 0; a0 < n_atoms; a0 += m) {
                const int64_t a1 = a0 + m < n_atoms ? a0 + m : n_atoms;
                amp_extended(t, a0, a1, n, &ca, &er, &ei);
                Ar += er;
                Ai += ei;
                self += er * er + ei * ei;
            }
            o[0] = (double)Ar;
            o[1] = (double)Ai;
            o[6] = (double)self;
            o[2] = o[3] = o[4] = o[5] = 0.0;
            if (excluded_ff) {
                amp_extended(t, 0, n_atoms, n, &ce, &er, &ei);
                o[2] = (double)er;
                o[3] = (double)ei;
            }
            if (water_ff && hydration_weight) {
                amp_extended(t, 0, n_atoms, n, &ch, &er, &ei);
                o[4] = (double)er;
                o[5] = (double)ei;
            }
            continue;
        }
        double Ar = 0, Ai = 0, self = 0, re, im;
        for (int64_t a0 = 0; a0 < n_atoms; a0 += m) {
            const int64_t a1 = a0 + m < n_atoms ? a0 + m : n_atoms;
            if (mode == ORACLE_DIRECT_TABLE) amp_table((const cplx *const *)pw, stride, a0, a1, n, &ca, &re, &im);
            else amp_reference(x, y, z, a0, a1, q, &ca, &re, &im);
            Ar += re;
            Ai += im;
            self += re * re + im * im;
        }
        o[0] = Ar;
        o[1] = Ai;
        o[6] = self;
        o[2] = o[3] = o[4] = o[5] = 0.0;
        if (excluded_ff) {
            if (mode == ORACLE_DIRECT_TABLE) amp_table((const cplx *const *)pw, stride, 0, n_atoms, n, &ce, &re, &im);
            else amp_reference(x, y, z, 0, n_atoms, q, &ce, &re, &im);
            o[2] = re;
            o[3] = im;
        }
        if (water_ff && hydration_weight) {
            if (mode == ORACLE_DIRECT_TABLE) amp_table((const cplx *const *)pw, stride, 0, n_atoms, n, &ch, &re, &im);
            else amp_reference(x, y, z, 0, n_atoms, q, &ch, &re, &im);
            o[4] = re;
            o[5] = im;
        }
    }
    for (int d = 0; d < 3; ++d) free(pw[d]);
    free(t);
    return 0;
}
