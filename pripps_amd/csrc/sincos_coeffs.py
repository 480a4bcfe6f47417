"""Fit the polynomial coefficients of ``sincos_rev`` in debye_kernels.hip, and
print the FP64 Taylor tables of ``sincos_rev_f64`` / ``sin_rev_f64``.

sin(2 pi t) and cos(2 pi t) on |t| <= 1/8 revolution:

    sin = t (S1 + S3 t^2 + S5 t^4 + S7 t^6 + S9 t^8)  relative error
    cos = 1 + C2 t^2 + C4 t^4 + C6 t^6 + C8 t^8        absolute error

S1 = 2 pi and C2 = -2 pi^2 are fixed and split into float hi + lo parts, so the
leading terms carry no rounding bias.  The remaining coefficients are fitted by
Lawson's iteratively reweighted least squares (which converges to the minimax
fit) and rounded to float one at a time, lowest order first, each rounding
followed by a refit of the higher-order coefficients.  Run:

    python pripps_amd/csrc/sincos_coeffs.py
"""
import mpmath as mp
import numpy as np

mp.mp.dps = 40
TWO_PI = mp.mpf(2) * mp.pi


def f32(x) -> float:
    return float(np.float32(float(x)))


def split(x):
    hi = f32(x)
    lo = f32(mp.mpf(x) - mp.mpf(hi))
    return hi, lo


def lawson(A, b, w0, iters=400):
    """Weighted minimax of A c ~ b (rows weighted by w0) by Lawson's method."""
    lam = np.full(len(b), 1.0 / len(b))
    c = None
    for _ in range(iters):
        sw = np.sqrt(lam) * w0
        c, *_ = np.linalg.lstsq(A * sw[:, None], b * sw, rcond=None)
        err = np.abs(A @ c - b) * w0
        lam = lam * err
        lam /= lam.sum()
    return c


def fit_sin():
    t = np.array([float(mp.mpf(1) / 8 * mp.sin(mp.pi / 2 * (k + 0.5) / 4000)) for k in range(4000)])
    s1hi, s1lo = split(TWO_PI)
    exact = np.array([float(mp.sin(TWO_PI * mp.mpf(x))) for x in t])
    lead = np.array([float(mp.sin(TWO_PI * mp.mpf(x)) - (mp.mpf(s1hi) + mp.mpf(s1lo)) * mp.mpf(x))
                     for x in t])
    fixed = []
    for n_free in (4, 3, 2, 1):
        # residual after the fixed coefficients, in units of t^3
        rem = lead.copy()
        for j, cf in enumerate(fixed):
            rem -= cf * t ** (3 + 2 * j)
        start = len(fixed)
        A = np.stack([t ** (3 + 2 * (start + j)) for j in range(n_free)], axis=1)
        c = lawson(A, rem, 1.0 / np.abs(exact))
        fixed.append(f32(c[0]))
    coef = [s1hi, s1lo] + fixed
    # error of the float coefficients in exact arithmetic (relative)
    p = (mp.mpf(s1hi) + mp.mpf(s1lo))
    val = np.array([float(p * mp.mpf(x) + sum(mp.mpf(cf) * mp.mpf(x) ** (3 + 2 * j)
                                          for j, cf in enumerate(fixed))) for x in t])
    rel = np.max(np.abs(val - exact) / np.abs(exact))
    return coef, rel


def fit_cos():
    t = np.array([float(mp.mpf(1) / 8 * mp.sin(mp.pi / 2 * (k + 0.5) / 4000)) for k in range(4000)])
    c2 = -2 * mp.pi ** 2
    c2hi, c2lo = split(c2)
    exact = np.array([float(mp.cos(TWO_PI * mp.mpf(x))) for x in t])
    lead = np.array([float(mp.cos(TWO_PI * mp.mpf(x)) - 1 - (mp.mpf(c2hi) + mp.mpf(c2lo)) * mp.mpf(x) ** 2)
                     for x in t])
    fixed = []
    for n_free in (3, 2, 1):
        rem = lead.copy()
        for j, cf in enumerate(fixed):
            rem -= cf * t ** (4 + 2 * j)
        start = len(fixed)
        A = np.stack([t ** (4 + 2 * (start + j)) for j in range(n_free)], axis=1)
        c = lawson(A, rem, np.ones_like(t))
        fixed.append(f32(c[0]))
    coef = [c2hi, c2lo] + fixed
    val = np.array([float(1 + (mp.mpf(c2hi) + mp.mpf(c2lo)) * mp.mpf(x) ** 2 +
                          sum(mp.mpf(cf) * mp.mpf(x) ** (4 + 2 * j) for j, cf in enumerate(fixed)))
                    for x in t])
    err = np.max(np.abs(val - exact))
    return coef, err


def taylor_f64():
    """FP64 tables of sincos_rev_f64 / sin_rev_f64: the Taylor coefficients
    (-1)^n (2 pi)^m / m! of sin(2 pi t) and cos(2 pi t) - 1, rounded to double,
    with the truncation bound of each table on its interval."""
    def c(m):
        return (-1) ** ((m - 1) // 2 if m % 2 else m // 2) * TWO_PI ** m / mp.factorial(m)
    quarter_sin = [float(c(m)) for m in range(1, 18, 2)]
    quarter_cos = [float(c(m)) for m in range(2, 17, 2)]
    half_sin = [float(c(m)) for m in range(1, 24, 2)]
    bounds = {"quarter sine (|t| <= 1/8, to t^17)": (mp.pi / 4) ** 19 / mp.factorial(19),
              "quarter cosine - 1 (|t| <= 1/8, to t^16)": (mp.pi / 4) ** 18 / mp.factorial(18),
              "half sine (|t| <= 1/4, to t^23)": (mp.pi / 2) ** 25 / mp.factorial(25)}
    return quarter_sin, quarter_cos, half_sin, bounds


if __name__ == "__main__":
    qs, qc, hs, bounds = taylor_f64()
    for name, tab in (("kDSinQ", qs), ("kDCosQ", qc), ("kDSinH", hs)):
        print(f"constexpr double {name}[{len(tab)}] = {{" + ", ".join(repr(v) for v in tab) + "};")
    for k, v in bounds.items():
        print(f"// truncation bound, {k}: {mp.nstr(v, 3)}")
    sc, se = fit_sin()
    cc, ce = fit_cos()
    names_s = ("S1_HI", "S1_LO", "S3", "S5", "S7", "S9")
    names_c = ("C2_HI", "C2_LO", "C4", "C6", "C8")
    for n, v in zip(names_s, sc):
        print(f"constexpr float k{n} = {v:.9e}f;")
    print(f"// max relative error of the sine polynomial: {se:.2e}")
    for n, v in zip(names_c, cc):
        print(f"constexpr float k{n} = {v:.9e}f;")
    print(f"// max absolute error of the cosine polynomial: {ce:.2e}")
