"""Direct Fourier transform of a periodic box: q-vector grids, grouping by
|q|, and the cached model (reference ``src/explicit.rs``, ``src/q_grid.rs``,
``src/gpu.rs``; Python ``Pripps.direct_model`` / ``intensity(model="direct")``,
``pripps-py/src/lib.rs:441-483, 485-595``).

The amplitudes are computed on the GPU by ``pripps_direct_model_build``
(``include/pripps_direct.h``); this module builds the q-vectors, the form-factor
tables at each vector's |q|, and the grouping of vectors into output rows.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _native as N
from .errors import UsageError

TWO_PI = 2.0 * math.pi

# The 13 low-index directions of the [100], [110] and [111] families, one of
# each antipodal pair (src/explicit.rs:44-58).
MILLER_INDEX = np.array([
    (1, 0, 0), (0, 1, 0), (0, 0, 1),
    (1, 1, 0), (0, 1, 1), (1, 0, 1),
    (-1, 1, 0), (-1, 0, 1), (0, -1, 1),
    (1, 1, 1), (-1, 1, 1), (1, -1, 1), (1, 1, -1),
], dtype=np.int32)

EXACT_BIN_LIMIT = 20_000.0          # src/q_grid.rs:170
GOLDEN_ANGLE = 2.399_963_229_728_653  # src/q_grid.rs:303


def rust_round(x):
    """f64::round: nearest integer, ties away from zero (exact for every x)."""
    x = np.asarray(x, dtype=np.float64)
    t = np.trunc(x)
    return np.where(np.abs(x - t) >= 0.5, t + np.sign(x), t)


def miller_indices(p_max: int) -> np.ndarray:
    """p (h, k, l) for the 13 Miller directions and p = 1..p_max, direction-major
    (src/explicit.rs:268-277)."""
    p = np.arange(1, int(p_max) + 1, dtype=np.int32)
    return (MILLER_INDEX[:, None, :] * p[None, :, None]).reshape(-1, 3)


def vector_norms(indices: np.ndarray, box_sides) -> np.ndarray:
    """|q| of q = (2 pi n) ./ L, as the reference forms it (2 pi n first, then the
    component division, then sqrt(x^2 + y^2 + z^2); src/explicit.rs:228, 279-281)."""
    n = np.asarray(indices, dtype=np.float64).reshape(-1, 3)
    L = np.asarray(box_sides, dtype=np.float64)
    q = (TWO_PI * n) / L
    return np.sqrt((q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1]) + q[:, 2] * q[:, 2])


def q_bin(q):
    """(q 1e6).round() as i64 (src/explicit.rs:388-392)."""
    return rust_round(np.asarray(q, dtype=np.float64) * 1e6).astype(np.int64)


def group_by_qnorm(qnorm: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Groups of equal |q| bin: the group of each vector and each group's q (its
    first vector's |q|), groups sorted by q (average_by_qnorm,
    src/explicit.rs:399-420)."""
    bins = q_bin(qnorm)
    first = {}
    for v, b in enumerate(bins.tolist()):
        first.setdefault(b, v)
    reps = sorted(first.items(), key=lambda kv: qnorm[kv[1]])
    row = {b: g for g, (b, _) in enumerate(reps)}
    group = np.array([row[b] for b in bins.tolist()], dtype=np.int32)
    group_q = np.array([qnorm[v] for _, v in reps], dtype=np.float64)
    return group, group_q


def average_duplicates(rows: Sequence[Tuple[float, float]]) -> List[Tuple[float, float]]:
    """Mean of the values of equal-|q| rows, sorted by q (src/explicit.rs:422-430)."""
    if not rows:
        return []
    q = np.array([r[0] for r in rows], dtype=np.float64)
    y = np.array([r[1] for r in rows], dtype=np.float64)
    group, group_q = group_by_qnorm(q)
    out = []
    for g in range(len(group_q)):
        vals = y[group == g]
        acc = 0.0
        for v in vals:
            acc += v
        out.append((float(group_q[g]), acc / len(vals)))
    return out


# ---------------------------------------------------------------------------
# lattice grids (src/q_grid.rs)
# ---------------------------------------------------------------------------
def _bin_width(L) -> float:
    return TWO_PI / max(L)


def _expected_count(L, k: int) -> float:
    w = _bin_width(L)
    vol = L[0] * L[1] * L[2]
    kw = k * w
    c = TWO_PI
    return TWO_PI * (kw * kw) * w * vol / (c * c * c)


def _largest_exact_bin(L) -> int:
    k = 1
    while _expected_count(L, k + 1) <= EXACT_BIN_LIMIT:
        k += 1
    return k


def _norm(L, n) -> float:
    s = 0.0
    for a in range(3):
        t = TWO_PI * float(n[a]) / L[a]
        s += t * t
    return math.sqrt(s)


def _upper_half(n) -> bool:
    h, k, l = n
    return h > 0 or (h == 0 and (k > 0 or (k == 0 and l > 0)))


def _direction_key(n) -> float:
    x, y, z = (float(c) for c in n)
    norm = math.sqrt(x * x + y * y + z * z)
    band = math.floor((z / norm + 1.0) * 8.0)
    return band * 10.0 + math.atan2(y, x)


def _spread_subset(shell, cap: int):
    if len(shell) <= cap:
        return list(shell)
    order = sorted(shell, key=lambda v: _direction_key(v[0]))  # stable, as sort_by
    return [order[i * len(order) // cap] for i in range(cap)]


def _exact_shells(L, q_max: float):
    w = _bin_width(L)
    last_exact = _largest_exact_bin(L)
    q_max = min(q_max, (last_exact + 0.5) * w)
    n_max = [int(math.floor(q_max * (1.0 + 1e-12) * L[a] / TWO_PI)) for a in range(3)]
    h, k, l = np.meshgrid(np.arange(-n_max[0], n_max[0] + 1), np.arange(-n_max[1], n_max[1] + 1),
                          np.arange(-n_max[2], n_max[2] + 1), indexing="ij")
    n = np.stack([h.ravel(), k.ravel(), l.ravel()], axis=1)
    up = (n[:, 0] > 0) | ((n[:, 0] == 0) & ((n[:, 1] > 0) | ((n[:, 1] == 0) & (n[:, 2] > 0))))
    n = n[up]
    t = (TWO_PI * n.astype(np.float64)) / np.asarray(L, dtype=np.float64)
    q = np.sqrt((t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1]) + t[:, 2] * t[:, 2])
    b = rust_round(q / w).astype(np.int64)
    keep = (q <= q_max * (1.0 + 1e-12)) & (b <= last_exact)
    shells = {}
    for vec, qq, bb in zip(n[keep].tolist(), q[keep].tolist(), b[keep].tolist()):
        shells.setdefault(bb, []).append((tuple(vec), qq))
    return [shells[bb] for bb in sorted(shells)]


def _shell_holds(L, a, b, n) -> bool:
    length = _norm(L, n)
    return a <= length < b


def _shell_any_vector(L, a, b):
    step = [TWO_PI / L[d] for d in range(3)]
    n_max = [int(math.floor(b / s)) for s in step]
    for h in range(0, n_max[0] + 1):
        for k in range(-n_max[1], n_max[1] + 1):
            planar = (h * step[0]) ** 2 + (k * step[1]) ** 2
            if planar >= b * b:
                continue
            l0 = int(math.ceil(math.sqrt(max(a * a - planar, 0.0)) / step[2]))
            for l in range(l0 - 1, l0 + 2):
                n = (h, k, l) if _upper_half((h, k, l)) else (-h, -k, -l)
                if _upper_half(n) and _shell_holds(L, a, b, n):
                    return n
    return None


def _sampled_shell(L, k: int, q_max: float, cap: int):
    w = _bin_width(L)
    a = (k - 0.5) * w
    b = min((k + 0.5) * w, q_max)
    if b <= a:
        return None
    a3, b3 = a * a * a, b * b * b
    a4, b4 = (a * a) * (a * a), (b * b) * (b * b)
    q = 0.75 * (b4 - a4) / (b3 - a3)
    b_shell = min((k + 0.5) * w, q_max * (1.0 + 1e-12))
    wanted = min(cap, int(EXACT_BIN_LIMIT))
    attempts = 2 * wanted + 16
    i = np.arange(attempts, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / attempts
    ang = i * GOLDEN_ANGLE
    s, c = np.sin(ang), np.cos(ang)
    rho = np.sqrt(1.0 - z * z)
    t = i * 0.618_033_988_749_895
    t = t - np.trunc(t)
    r = np.cbrt(a3 + t * (b3 - a3))
    d = np.stack([rho * c, rho * s, z], axis=1)
    n = np.stack([rust_round(r * d[:, ax] * L[ax] / TWO_PI) for ax in range(3)], axis=1)
    n = n.astype(np.int64)
    seen = set()
    vectors = []
    for vec in n.tolist():
        vec = tuple(vec)
        if not _upper_half(vec):
            vec = (-vec[0], -vec[1], -vec[2])
        if _upper_half(vec) and _shell_holds(L, a, b_shell, vec) and vec not in seen:
            seen.add(vec)
            vectors.append(vec)
    if not vectors:
        v = _shell_any_vector(L, a, b_shell)
        return None if v is None else (q, [v])
    m = min(wanted, len(vectors))
    return q, [vectors[i * len(vectors) // m] for i in range(m)]


class QGrid:
    """Which q-vectors a periodic box is sampled at (reference ``QGrid``,
    src/q_grid.rs:13-158): ``miller(p_max)`` (13 directions x p, grouped by
    exact |q| for any box) or ``lattice(box_sides, q_max, max_per_bin)`` (every
    reciprocal-lattice vector up to q_max in |q| bins of width 2 pi / max(L),
    large bins thinned to an even spread, bins beyond ~20,000 vectors
    sampled)."""

    def __init__(self, kind: str, p_max: int = 0, box_sides=None, vectors=None, bins=None,
                 q=None):
        self._kind = kind
        self._p_max = p_max
        self._box = None if box_sides is None else tuple(float(x) for x in box_sides)
        self._vectors = vectors
        self._bins = bins
        self._q = q

    @staticmethod
    def miller(p_max: int) -> "QGrid":
        return QGrid("miller", p_max=int(p_max))

    @staticmethod
    def lattice(box_sides, q_max: float, max_per_bin: int) -> "QGrid":
        L = [float(x) for x in np.broadcast_to(np.asarray(box_sides, dtype=np.float64), (3,))]
        if not all(math.isfinite(x) and x > 0.0 for x in L):
            raise UsageError(f"invalid box sides {L}")
        if not (math.isfinite(q_max) and q_max > 0.0):
            raise UsageError(f"invalid q_max {q_max} Å⁻¹")
        shells = _exact_shells(L, float(q_max))
        if not shells:
            raise UsageError(f"q_max {q_max} Å⁻¹ is below the shortest lattice vector of box {L}")
        cap = max(int(max_per_bin), 1)
        vectors, bins, qs = [], [], []
        for shell in shells:
            s = 0.0
            for _, qq in shell:
                s += qq
            qs.append(s / len(shell))
            for vec, _ in _spread_subset(shell, cap):
                vectors.append(vec)
                bins.append(len(qs) - 1)
        w = _bin_width(L)
        first_sampled = _largest_exact_bin(L) + 1
        last = int(rust_round(q_max / w))
        for k in range(first_sampled, last + 1):
            got = _sampled_shell(L, k, float(q_max), cap)
            if got is not None:
                qs.append(got[0])
                for vec in got[1]:
                    vectors.append(vec)
                    bins.append(len(qs) - 1)
        return QGrid("lattice", box_sides=L, vectors=np.array(vectors, dtype=np.int32).reshape(-1, 3),
                     bins=np.array(bins, dtype=np.int32), q=np.array(qs, dtype=np.float64))

    @property
    def kind(self) -> str:
        return self._kind

    def q(self) -> np.ndarray:
        """Bin q-values, ascending; empty for Miller grids."""
        return np.zeros(0) if self._kind == "miller" else self._q.copy()

    def vectors_per_bin(self) -> List[int]:
        if self._kind == "miller":
            return []
        return np.bincount(self._bins, minlength=len(self._q)).tolist()

    def indices(self) -> np.ndarray:
        if self._kind == "miller":
            return miller_indices(self._p_max)
        return self._vectors.copy()

    def check_box(self, box_sides) -> None:
        if self._kind == "miller":
            return
        ok = all(abs(b - l) <= 1e-9 * l for b, l in zip(box_sides, self._box))
        if not ok:
            raise UsageError(f"box {list(box_sides)} differs from the {list(self._box)} the q-grid "
                             "was built for")

    def groups(self, box_sides) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(indices [V][3], |q| [V], group [V], group_q [G]) for a box."""
        self.check_box(box_sides)
        idx = self.indices()
        qn = vector_norms(idx, box_sides)
        if self._kind == "miller":
            group, group_q = group_by_qnorm(qn)
        else:
            group, group_q = self._bins.copy(), self._q.copy()
        return idx, qn, group, group_q


def parse_box(box_sides) -> Tuple[float, float, float]:
    """A float (cubic box) or a 3-sequence of floats (pripps-py/src/lib.rs:766-777)."""
    if isinstance(box_sides, (int, float, np.floating, np.integer)):
        v = float(box_sides)
        return (v, v, v)
    try:
        vals = [float(x) for x in box_sides]
    except TypeError:
        vals = []
    if len(vals) != 3:
        raise ValueError("`box_sides` must be a float (cubic) or a 3-sequence of floats")
    return tuple(vals)


def check_box(box) -> None:
    if not all(math.isfinite(b) and b > 0.0 for b in box):
        raise UsageError(f"invalid box side lengths {list(box)}")


class NativeDirectModel:
    """Owner of a ``pripps_direct_model`` handle."""

    def __init__(self, handle, engine):
        self._h = handle
        self._engine = engine  # keeps the engine alive
        n = ctypes.c_int32()
        N.check(N.load().pripps_direct_model_n_groups(self._h, ctypes.byref(n)))
        self.n_groups = n.value

    @property
    def handle(self):
        return self._h

    def q_values(self) -> np.ndarray:
        out = np.zeros(self.n_groups)
        N.check(N.load().pripps_direct_model_q_values(self._h, N.dptr(out)))
        return out

    def intensity(self, vs: float, cd: float):
        G = self.n_groups
        out = np.zeros((4, G))
        ptrs = [out[k].ctypes.data_as(N._D) for k in range(4)]
        N.check(N.load().pripps_direct_model_intensity(self._h, float(vs), float(cd), *ptrs))
        return out[0], out[1], out[2], out[3]

    def self_term(self) -> np.ndarray:
        out = np.zeros(self.n_groups)
        N.check(N.load().pripps_direct_model_self_term(self._h, N.dptr(out)))
        return out

    def profiles(self) -> np.ndarray:
        out = np.zeros((6, self.n_groups))
        N.check(N.load().pripps_direct_model_profiles(self._h, N.dptr(out)))
        return out

    def amplitudes(self, n_vectors: int) -> np.ndarray:
        out = np.zeros((n_vectors, 7))
        N.check(N.load().pripps_direct_model_amplitudes(self._h, N.dptr(out)))
        return out

    def stats(self) -> N.BuildStats:
        s = N.BuildStats()
        N.check(N.load().pripps_direct_model_stats(self._h, ctypes.byref(s)))
        return s

    def recompute(self, repeats: int) -> float:
        ms = ctypes.c_double()
        N.check(N.load().pripps_direct_model_recompute(self._h, int(repeats), ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self._h:
            N.load().pripps_direct_model_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _grid_struct(box, indices, group, group_q, num_molecules):
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.int32).reshape(-1, 3))
    grp = np.ascontiguousarray(group, dtype=np.int32)
    gq = np.ascontiguousarray(group_q, dtype=np.float64)
    grid = N.DirectGrid((ctypes.c_double * 3)(*[float(b) for b in box]), len(idx),
                        idx.ctypes.data_as(N._I32), int(num_molecules), len(gq),
                        grp.ctypes.data_as(N._I32), gq.ctypes.data_as(N._D))
    return grid, (idx, grp, gq)


def frames_average(engine, inp: N.PreparedInput, box, indices, group, group_q, frames,
                   weights=None, stride: int = 1, num_molecules: int = 0,
                   precision: str = "fp32", volume_scale: float = 1.0,
                   contrast_density: float = 0.0):
    """(total [G], self [G] or None): weighted average over frames
    (``pripps_direct_frames_average``)."""
    prec = N.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    grid, keep = _grid_struct(box, indices, group, group_q, num_molecules)
    f = np.ascontiguousarray(frames, dtype=np.float64)
    w = None if weights is None else np.ascontiguousarray(weights, dtype=np.float64)
    if w is not None and len(w) != len(f):
        from .errors import ValidationError
        raise ValidationError(f"{len(w)} weights for {len(f)} frames")
    G = len(keep[2])
    total = np.zeros(G)
    self_term = np.zeros(G) if num_molecules else None
    N.check(N.load().pripps_direct_frames_average(
        engine.handle, ctypes.byref(inp.c), ctypes.byref(grid), len(f), N.dptr(f), N.dptr(w),
        int(stride), prec, float(volume_scale), float(contrast_density), N.dptr(total),
        N.dptr(self_term)))
    return total, self_term


def build_native(engine, inp: N.PreparedInput, box, indices, group, group_q,
                 num_molecules: int = 0, precision: str = "fp32") -> NativeDirectModel:
    prec = N.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    grid, keep = _grid_struct(box, indices, group, group_q, num_molecules)
    h = ctypes.c_void_p()
    N.check(N.load().pripps_direct_model_build(engine.handle, ctypes.byref(inp.c),
                                               ctypes.byref(grid), prec, ctypes.byref(h)))
    return NativeDirectModel(h, engine)
