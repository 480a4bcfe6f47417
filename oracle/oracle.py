"""ctypes front-end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Loads ``oracle/liboracle.so`` (built by ``oracle/Makefile``).  Only tests/,
``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this
module, always as the checker; the engine in ``pripps_amd`` never does.

Every function here forwards to the C restatement in ``debye_oracle.c`` /
``fit_oracle.c``, whose headers list the reference file:line each part
restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

REFERENCE_ORDER = 0
COMPENSATED = 1

PROFILE_NAMES = ("atoms", "excluded", "hydration", "atom_excluded", "atom_hydration",
                 "excluded_hydration")

_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
THEORY_FN = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_double, ctypes.c_void_p, _D)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_linear_q_values.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int64, _D]
        L.oracle_sinc.argtypes = [ctypes.c_double]
        L.oracle_sinc.restype = ctypes.c_double
        L.oracle_self_intensity_at_q.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_double]
        L.oracle_self_intensity_at_q.restype = ctypes.c_double
        L.oracle_cross_intensity_at_q.argtypes = [_D, _D, ctypes.c_int64, _D, _D, ctypes.c_int64,
                                                  ctypes.c_double]
        L.oracle_cross_intensity_at_q.restype = ctypes.c_double
        L.oracle_build_model.argtypes = [_D, _I32, ctypes.c_int64, ctypes.c_int32, _D,
                                         ctypes.c_int64, _D, _D, _D, _D, ctypes.c_int,
                                         ctypes.c_int, _D]
        L.oracle_build_model.restype = ctypes.c_int
        L.oracle_combine.argtypes = [_D, ctypes.c_double, ctypes.c_double]
        L.oracle_combine.restype = ctypes.c_double
        L.oracle_eval_gaussian.argtypes = [ctypes.c_double, _D, _D, ctypes.c_int, ctypes.c_double]
        L.oracle_eval_gaussian.restype = ctypes.c_double
        L.oracle_eval_single_gaussian.argtypes = [ctypes.c_double] * 3
        L.oracle_eval_single_gaussian.restype = ctypes.c_double
        L.oracle_eval_polynomial.argtypes = [_D, ctypes.c_int, ctypes.c_double]
        L.oracle_eval_polynomial.restype = ctypes.c_double
        L.oracle_interpolate.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_double]
        L.oracle_interpolate.restype = ctypes.c_double
        L.oracle_united_grid.argtypes = [ctypes.c_double, ctypes.c_double, _D]
        L.oracle_united_table.argtypes = [_D, _D, ctypes.c_int, _D, _D, _D, _D]
        L.oracle_fraser_params.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D]
        L.oracle_spline_eval.argtypes = [_D, _D, ctypes.c_int64, _D, ctypes.c_int64, _D]
        L.oracle_compute_scale.argtypes = [_D, _D, ctypes.c_int64, _D, _D, _D, ctypes.c_int64]
        L.oracle_compute_scale.restype = ctypes.c_double
        L.oracle_chi2_per_point.argtypes = [_D, _D, ctypes.c_int64, _D, _D, _D, ctypes.c_int64]
        L.oracle_chi2_per_point.restype = ctypes.c_double
        L.oracle_initial_simplex.argtypes = [_D, _D, _D]
        L.oracle_fit_nelder_mead.argtypes = [THEORY_FN, ctypes.c_void_p, _D, ctypes.c_int64, _D, _D,
                                             _D, ctypes.c_int64, _D, _D, _D]
        L.oracle_fit_nelder_mead.restype = ctypes.c_int
        L.oracle_fit_grid.argtypes = [THEORY_FN, ctypes.c_void_p, _D, ctypes.c_int64, _D, _D, _D,
                                      ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _D, _D, _D]
        L.oracle_fit_grid.restype = ctypes.c_int
        L.oracle_direct_rows.argtypes = [ctypes.c_int64, _D, _D, _D, _I32, ctypes.c_int32, _D,
                                         ctypes.c_int32, _I32, _D, _D, _D, _D, ctypes.c_int32,
                                         ctypes.c_int, _D]
        L.oracle_direct_rows.restype = ctypes.c_int
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def linear_q_values(qmin: float, qmax: float, nq: int) -> np.ndarray:
    out = np.empty(nq, dtype=np.float64)
    lib().oracle_linear_q_values(qmin, qmax, nq, out.ctypes.data_as(_D))
    return out


def sinc(x: float) -> float:
    return lib().oracle_sinc(x)


def self_intensity_at_q(pos, ff, q: float) -> float:
    pos, pp = _d(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    ff, fp = _d(ff)
    return lib().oracle_self_intensity_at_q(pp, fp, len(ff), q)


def cross_intensity_at_q(pa, fa, pb, fb, q: float) -> float:
    pa, pap = _d(np.asarray(pa, dtype=np.float64).reshape(-1, 3))
    fa, fap = _d(fa)
    pb, pbp = _d(np.asarray(pb, dtype=np.float64).reshape(-1, 3))
    fb, fbp = _d(fb)
    return lib().oracle_cross_intensity_at_q(pap, fap, len(fa), pbp, fbp, len(fb), q)


def build_model(pos, kind, q, atom_ff, excluded_ff=None, water_ff=None, weights=None,
                mode: int = COMPENSATED, nthreads: int | None = None) -> np.ndarray:
    """Six partial profiles [6][nq] (DebyeTerms order) -- src/debye.rs:78-136."""
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    kind = np.ascontiguousarray(kind, dtype=np.int32)
    q, qp = _d(q)
    atom_ff = np.ascontiguousarray(atom_ff, dtype=np.float64)
    n_kinds = atom_ff.shape[0]
    nq = len(q)
    keep = []

    def opt(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return a.ctypes.data_as(_D)

    out = np.zeros((6, nq), dtype=np.float64)
    if nthreads is None:
        nthreads = os.cpu_count() or 1
    rc = lib().oracle_build_model(pos.ctypes.data_as(_D), kind.ctypes.data_as(_I32), len(kind),
                                  n_kinds, qp, nq, atom_ff.ctypes.data_as(_D), opt(excluded_ff),
                                  opt(water_ff), opt(weights), mode, nthreads,
                                  out.ctypes.data_as(_D))
    if rc != 0:
        raise ValueError("oracle_build_model: invalid arguments")
    return out


def combine(profiles: np.ndarray, vs: float, cd: float) -> np.ndarray:
    """DebyeTerms::combine per q (src/debye.rs:186-195)."""
    p = np.ascontiguousarray(profiles, dtype=np.float64)
    out = np.empty(p.shape[1])
    col = np.empty(6)
    for m in range(p.shape[1]):
        col[:] = p[:, m]
        out[m] = lib().oracle_combine(col.ctypes.data_as(_D), vs, cd)
    return out


def eval_gaussian(q, ab, c):
    a = np.ascontiguousarray([x[0] for x in ab], dtype=np.float64)
    b = np.ascontiguousarray([x[1] for x in ab], dtype=np.float64)
    return lib().oracle_eval_gaussian(q, a.ctypes.data_as(_D), b.ctypes.data_as(_D), len(a), c)


def eval_single_gaussian(a, b, q):
    return lib().oracle_eval_single_gaussian(a, b, q)


def eval_polynomial(coeff, q):
    c, cp = _d(coeff)
    return lib().oracle_eval_polynomial(cp, len(c), q)


def interpolate(xs, ys, x):
    xs, xp = _d(xs)
    ys, yp = _d(ys)
    return lib().oracle_interpolate(xp, yp, len(xs), x)


def united_table(qmin, qmax, center_fn, orbits):
    """orbits: list of (form-factor callable, offset, weight)."""
    qs = np.empty(101)
    lib().oracle_united_grid(qmin, qmax, qs.ctypes.data_as(_D))
    center = np.array([center_fn(x) for x in qs])
    off = np.array([o[1] for o in orbits], dtype=np.float64)
    w = np.array([o[2] for o in orbits], dtype=np.float64)
    offf = np.ascontiguousarray([[o[0](x) for x in qs] for o in orbits], dtype=np.float64).ravel()
    out = np.empty(101)
    lib().oracle_united_table(qs.ctypes.data_as(_D), center.ctypes.data_as(_D), len(orbits),
                              offf.ctypes.data_as(_D), off.ctypes.data_as(_D),
                              w.ctypes.data_as(_D), out.ctypes.data_as(_D))
    return qs, out


def fraser_params(volume, rho, denominator):
    a = ctypes.c_double()
    b = ctypes.c_double()
    lib().oracle_fraser_params(volume, rho, denominator, ctypes.byref(a), ctypes.byref(b))
    return a.value, b.value


def spline_eval(xs, ys, x):
    xs, xp = _d(xs)
    ys, yp = _d(ys)
    x, xxp = _d(np.atleast_1d(x))
    out = np.empty(len(x))
    lib().oracle_spline_eval(xp, yp, len(xs), xxp, len(x), out.ctypes.data_as(_D))
    return out


def chi2_per_point(tq, ttot, q, inten, sigma):
    arrs = [_d(a) for a in (tq, ttot, q, inten, sigma)]
    return lib().oracle_chi2_per_point(arrs[0][1], arrs[1][1], len(arrs[0][0]), arrs[2][1],
                                       arrs[3][1], arrs[4][1], len(arrs[2][0]))


def compute_scale(tq, ttot, q, inten, sigma):
    arrs = [_d(a) for a in (tq, ttot, q, inten, sigma)]
    return lib().oracle_compute_scale(arrs[0][1], arrs[1][1], len(arrs[0][0]), arrs[2][1],
                                      arrs[3][1], arrs[4][1], len(arrs[2][0]))


def initial_simplex(c1b, c2b):
    out = np.empty(6)
    a, ap = _d(c1b)
    b, bp = _d(c2b)
    lib().oracle_initial_simplex(ap, bp, out.ctypes.data_as(_D))
    return [(out[0], out[1]), (out[2], out[3]), (out[4], out[5])]


FIT_ERRORS = {
    -1: "c1 bounds must be finite", -2: "c1 lower bound must be less than upper bound",
    -11: "c2 bounds must be finite", -12: "c2 lower bound must be less than upper bound",
    -20: "grid fit requires at least 2 points per axis",
    -30: "grid fit did not find any finite chi2 value",
}


def fit(theory, tq, q, inten, sigma=None, method="nelder-mead", c1_points=151, c2_points=101,
        c1_bounds=(0.5, 2.0), c2_bounds=(0.0, 0.1)):
    """Restated fit.rs fit over a theory callable (c1, c2) -> total[nq]."""
    tq, tqp = _d(tq)
    nq = len(tq)
    q, qp = _d(q)
    inten, ip = _d(inten)
    if sigma is None or len(sigma) == 0:
        sigma = np.ones(len(q))
    sigma, sp = _d(sigma)
    c1b, c1p = _d(c1_bounds)
    c2b, c2p = _d(c2_bounds)
    out = np.empty(4)

    def cb(c1, c2, _ctx, buf):
        t = np.asarray(theory(c1, c2), dtype=np.float64)
        ctypes.memmove(buf, t.ctypes.data, 8 * nq)

    fn = THEORY_FN(cb)
    if method == "grid":
        rc = lib().oracle_fit_grid(fn, None, tqp, nq, qp, ip, sp, len(q), c1_points, c2_points,
                                   c1p, c2p, out.ctypes.data_as(_D))
    else:
        rc = lib().oracle_fit_nelder_mead(fn, None, tqp, nq, qp, ip, sp, len(q), c1p, c2p,
                                          out.ctypes.data_as(_D))
    if rc != 0:
        raise ValueError(FIT_ERRORS.get(rc, f"fit error {rc}"))
    return tuple(out)


# ---------------------------------------------------------------------------
# Direct Fourier transform (direct_oracle.c)
# ---------------------------------------------------------------------------
DIRECT_REFERENCE = 0
DIRECT_TABLE = 1
DIRECT_EXTENDED = 2


def direct_rows(pos, kind, box, indices, atom_ff, excluded_ff=None, water_ff=None, weights=None,
                num_molecules=0, mode=DIRECT_EXTENDED):
    """Per q-vector [n_vectors][7]: Re A, Im A, Re E, Im E, Re H, Im H, self
    (direct_oracle.c: src/explicit.rs:136-146, 217-244, 283-326)."""
    pos = np.ascontiguousarray(np.asarray(pos, dtype=np.float64).reshape(-1, 3))
    x, y, z = (np.ascontiguousarray(pos[:, d]) for d in range(3))
    kind = np.ascontiguousarray(kind, dtype=np.int32)
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int32).reshape(-1, 3))
    V = len(idx)
    box = np.ascontiguousarray(np.broadcast_to(np.asarray(box, dtype=np.float64), (3,)))
    ffa = np.ascontiguousarray(atom_ff, dtype=np.float64)
    ffe = None if excluded_ff is None else np.ascontiguousarray(excluded_ff, dtype=np.float64)
    ffw = None if water_ff is None else np.ascontiguousarray(water_ff, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    K = ffa.shape[0] if ffa.ndim == 2 else 1
    out = np.zeros((V, 7))

    def ptr(a):
        return None if a is None else a.ctypes.data_as(_D)

    rc = lib().oracle_direct_rows(len(x), ptr(x), ptr(y), ptr(z), kind.ctypes.data_as(_I32), K,
                                  ptr(box), V, idx.ctypes.data_as(_I32), ptr(ffa), ptr(ffe),
                                  ptr(ffw), ptr(w), int(num_molecules), int(mode), ptr(out))
    if rc != 0:
        raise MemoryError("oracle_direct_rows")
    return out
