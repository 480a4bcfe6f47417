"""Seeded synthetic workloads for the exact Debye engine.

Every workload is a set of globules:

* positions uniform in a sphere of radius R by sequential random addition
  with a minimum pair separation of 1.4 A (a candidate is kept iff it is at
  least 1.4 A from every atom kept before it);
* elements C 63 % / N 17 % / O 19 % / S 1 %, exact counts by largest
  remainder, randomly ordered;
* hydration weights w ~ U[0, 1] for atoms within 3 A of the globule's convex
  hull, 0 otherwise.

Named workloads (sizes are the ones the benchmarks quote):

=========  ==========================================================
vacuum1k   N = 1000, R = 17 A, q uniform [0, 0.5], nq = 500, vacuum
solvent10k N = 10000, R = 37 A, q uniform [0, 0.5], nq = 500, FoXS
           excluded volume + surface hydration weights
helix60k   6 x 10000-atom globules (R = 37 A) on a helix of pitch 60 A:
           3/4 turn per globule, helix radius 45 A (centre spacing
           77.9 A > 2R + 1.4 A, so globules never overlap); hydration
           shell per globule; q log-spaced on [0.005, 1.0], nq = 1000
batch512   512 globules, N log-uniform on [500, 20000], R = 17 (N/1000)^(1/3)
           A, Gaussian jitter sigma = 0.5 A per coordinate, own uniform
           q grid on [0, qmax] with nq ~ U{100..1000}, qmax ~ U[0.3, 0.8]
served     request stream: N log-uniform [300, 30000], nq ~ U{200..800},
           q on [0, 0.5], solvent on, 5 intensity() knob pairs each
=========  ==========================================================
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

ELEMENTS = ("C", "N", "O", "S")
FRACTIONS = (0.63, 0.17, 0.19, 0.01)
MIN_SEPARATION = 1.4
SHELL = 3.0


@dataclass
class Workload:
    name: str
    names: List[str]                 # element per atom
    pos: np.ndarray                  # [N, 3] float64
    q: np.ndarray                    # [nq] float64
    weights: Optional[np.ndarray]    # [N] hydration weights, None = vacuum
    excluded: str = "foxs"           # excluded-volume model name
    meta: dict = field(default_factory=dict)

    @property
    def n_atoms(self) -> int:
        return len(self.names)

    @property
    def pair_q_terms(self) -> int:
        n = self.n_atoms
        return n * (n - 1) // 2 * len(self.q)


def linear_q(qmin: float, qmax: float, nq: int) -> np.ndarray:
    """q_i = qmin + (qmax - qmin) i / max(nq - 1, 1) (src/q_sampler.rs:24-29)."""
    denom = float(max(nq - 1, 1))
    return np.array([qmin + (qmax - qmin) * i / denom for i in range(nq)], dtype=np.float64)


def log_q(qmin: float, qmax: float, nq: int) -> np.ndarray:
    return np.exp(np.linspace(math.log(qmin), math.log(qmax), nq))


def element_names(n: int, rng: np.random.Generator) -> List[str]:
    raw = [f * n for f in FRACTIONS]
    counts = [int(math.floor(x)) for x in raw]
    order = sorted(range(len(raw)), key=lambda k: -(raw[k] - counts[k]))
    for k in order[: n - sum(counts)]:
        counts[k] += 1
    names = []
    for e, c in zip(ELEMENTS, counts):
        names += [e] * c
    perm = rng.permutation(n)
    return [names[p] for p in perm]


def globule_positions(n: int, radius: float, rng: np.random.Generator) -> np.ndarray:
    """Sequential random addition in a sphere with a 1.4 A hard core."""
    from scipy.spatial import cKDTree

    kept = np.empty((0, 3))
    tree = None
    batch = max(256, n // 4)
    while len(kept) < n:
        cand = rng.uniform(-radius, radius, size=(batch, 3))
        cand = cand[np.einsum("ij,ij->i", cand, cand) <= radius * radius]
        if tree is not None and len(cand):
            d, _ = tree.query(cand, k=1, distance_upper_bound=MIN_SEPARATION)
            cand = cand[d >= MIN_SEPARATION]
        if len(cand):
            # greedy in draw order inside the batch
            bad = np.zeros(len(cand), dtype=bool)
            pairs = cKDTree(cand).query_pairs(MIN_SEPARATION, output_type="ndarray")
            if len(pairs):
                pairs = pairs[np.argsort(pairs[:, 1], kind="stable")]
                by_later = {}
                for a, b in pairs:
                    by_later.setdefault(int(b), []).append(int(a))
                for b in range(len(cand)):
                    for a in by_later.get(b, ()):
                        if not bad[a]:
                            bad[b] = True
                            break
            cand = cand[~bad]
            kept = np.vstack([kept, cand[: n - len(kept)]])
            tree = cKDTree(kept)
    return kept


def surface_weights(pos: np.ndarray, rng: np.random.Generator, shell: float = SHELL) -> np.ndarray:
    """w ~ U[0, 1] for atoms within `shell` of the convex hull, else 0."""
    from scipy.spatial import ConvexHull

    hull = ConvexHull(pos)
    eq = hull.equations  # n.x + d <= 0 inside, |n| = 1
    depth = -(pos @ eq[:, :3].T + eq[:, 3])
    dist = depth.min(axis=1)
    w = rng.uniform(0.0, 1.0, size=len(pos))
    return np.where(dist <= shell, w, 0.0)


def globule(n: int, radius: float, seed: int, solvent: bool = True, jitter: float = 0.0):
    rng = np.random.default_rng(seed)
    pos = globule_positions(n, radius, rng)
    names = element_names(n, rng)
    w = surface_weights(pos, rng) if solvent else None
    if jitter > 0.0:
        pos = pos + rng.normal(0.0, jitter, size=pos.shape)
    return names, pos, w


def radius_for(n: int) -> float:
    return 17.0 * (n / 1000.0) ** (1.0 / 3.0)


def vacuum1k(seed: int = 1) -> Workload:
    names, pos, _ = globule(1000, 17.0, seed, solvent=False)
    return Workload("vacuum1k", names, pos, linear_q(0.0, 0.5, 500), None, excluded="none")


def solvent10k(seed: int = 2, n: int = 10000, nq: int = 500) -> Workload:
    radius = 37.0 if n == 10000 else radius_for(n)
    names, pos, w = globule(n, radius, seed, solvent=True)
    return Workload("solvent10k" if n == 10000 else f"solvent{n}", names, pos,
                    linear_q(0.0, 0.5, nq), w, meta={"vs": 1.0, "cd": 0.03})


def helix60k(seed: int = 3, per_globule: int = 10000, nq: int = 1000) -> Workload:
    names: List[str] = []
    parts, ws = [], []
    pitch, rad, turn = 60.0, 45.0, 0.75
    radius = 37.0 if per_globule == 10000 else radius_for(per_globule)
    for g in range(6):
        theta = 2.0 * math.pi * turn * g
        centre = np.array([rad * math.cos(theta), rad * math.sin(theta), pitch * turn * g])
        nm, p, w = globule(per_globule, radius, seed * 1000 + g, solvent=True)
        names += nm
        parts.append(p + centre)
        ws.append(w)
    return Workload("helix60k", names, np.vstack(parts), log_q(0.005, 1.0, nq),
                    np.concatenate(ws), meta={"helix": {"pitch": pitch, "radius": rad,
                                                       "turn_per_globule": turn}})


def batch(count: int = 512, seed: int = 4, n_lo: int = 500, n_hi: int = 20000,
          nq_lo: int = 100, nq_hi: int = 1000, qmax_lo: float = 0.3, qmax_hi: float = 0.8,
          jitter: float = 0.5) -> List[Workload]:
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), size=count)).astype(int)
    nqs = rng.integers(nq_lo, nq_hi + 1, size=count)
    qmaxs = rng.uniform(qmax_lo, qmax_hi, size=count)
    out = []
    for s in range(count):
        names, pos, w = globule(int(ns[s]), radius_for(int(ns[s])), seed * 100000 + s,
                                solvent=True, jitter=jitter)
        out.append(Workload(f"batch{s}", names, pos, linear_q(0.0, float(qmaxs[s]), int(nqs[s])), w))
    return out


def request_stream(count: int, seed: int = 5, n_lo: int = 300, n_hi: int = 30000,
                   nq_lo: int = 200, nq_hi: int = 800):
    """Sizes and knobs of a served request stream (structures are built lazily
    with `globule(n, radius_for(n), seed)`)."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(n_lo), math.log(n_hi), size=count)).astype(int)
    nqs = rng.integers(nq_lo, nq_hi + 1, size=count)
    knobs = np.stack([rng.uniform(0.95, 1.12, size=(count, 5)),
                      rng.uniform(-2.0, 4.0, size=(count, 5))], axis=-1)
    seeds = rng.integers(0, 2**31 - 1, size=count)
    return ns, nqs, knobs, seeds
