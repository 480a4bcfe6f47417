"""GPU parity of the Chebyshev pair kernels (FP32 and FP64, q-strips of up
to 256 points with a per-pair Jacobi-Anger expansion) against the CPU oracle,
on uniform and non-uniform grids, with and without solvent, at the edges
(empty and one-atom structures, coincident atoms, q = 0, one-point strips),
and bitwise batch == alone.  Tolerance: envelope error 1e-5 (FP32
compensated) and 1e-12 (FP64), as in test_gpu_parity.py."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import random_cluster
from test_gpu_parity import build, envelope_error, make_case, oracle_profiles

pytestmark = pytest.mark.gpu

CHEB = 256
TOL = {"fp32": 1e-5, "fp64": 1e-12}
PREC = pytest.mark.parametrize("precision", ["fp32", "fp64"])


def _cheb(**kw):
    from pripps_amd._native import tuning
    return tuning(q_per_lane=CHEB, **kw)


@PREC
@pytest.mark.parametrize("n,nq,qmax,solvent", [
    (300, 37, 0.5, True), (700, 130, 0.5, True), (2000, 300, 0.5, True),
    (1500, 600, 0.8, False), (2500, 500, 0.5, True), (64, 257, 0.5, True),
    (65, 256, 0.5, False), (5, 9, 0.3, True), (2, 3, 0.5, True)])
def test_cheb_uniform_matches_oracle(engine, oracle, precision, n, nq, qmax, solvent):
    case = make_case(n, 300 + n, nq=nq, qmax=qmax, solvent=solvent)
    m = build(engine, case, precision, _cheb())
    assert m.stats().q_per_lane == CHEB
    err = envelope_error(m.profiles(), oracle_profiles(oracle, case))
    assert err.max() <= TOL[precision], err.max()


@PREC
@pytest.mark.parametrize("n,nq,lo,hi", [(1200, 400, 0.005, 1.0), (600, 50, 0.01, 0.6),
                                        (900, 1000, 0.005, 1.0)])
def test_cheb_log_grid_matches_oracle(engine, oracle, precision, n, nq, lo, hi):
    q = np.exp(np.linspace(np.log(lo), np.log(hi), nq))
    case = make_case(n, 400 + n, q=q)
    m = build(engine, case, precision, _cheb())
    err = envelope_error(m.profiles(), oracle_profiles(oracle, case))
    assert err.max() <= TOL[precision], err.max()


@PREC
def test_cheb_unsorted_grid(engine, oracle, precision):
    """q values in no particular order: each strip's centre and half-width
    come from its own minimum and maximum."""
    rng = np.random.default_rng(5)
    q = rng.permutation(np.linspace(0.0, 0.6, 300))
    case = make_case(800, 77, q=q)
    err = envelope_error(build(engine, case, precision, _cheb()).profiles(),
                         oracle_profiles(oracle, case))
    assert err.max() <= TOL[precision], err.max()


@PREC
def test_cheb_edges(engine, oracle, precision):
    pos = random_cluster(50, 17)
    pos = np.vstack([pos, pos[:5]])  # five duplicated atoms: r = 0, sinc = 1
    rng = np.random.default_rng(3)
    kind = rng.integers(0, 2, size=len(pos)).astype(np.int32)
    for q in (np.array([0.0]), np.array([0.3]), np.array([0.0, 0.25]),
              np.array([1e-4, 0.7, 0.1])):
        ffa = np.array([[6.0 + k for _ in q] for k in range(2)])
        case = (pos, kind, q, ffa, ffa * 0.4, np.full(len(q), 5.0), np.linspace(0, 1, len(pos)))
        err = envelope_error(build(engine, case, precision, _cheb()).profiles(),
                             oracle_profiles(oracle, case))
        assert err.max() <= TOL[precision], (q, err.max())
    for n in (0, 1):
        case = make_case(n, 3, nq=20)
        got = build(engine, case, precision, _cheb()).profiles()
        assert np.all(np.isfinite(got))
        assert envelope_error(got, oracle_profiles(oracle, case)).max() <= TOL[precision]


@PREC
def test_cheb_many_kinds_and_blocks(engine, oracle, precision):
    case = make_case(900, 19, kinds=12, nq=200)
    for block in (16, 64, 128):
        got = build(engine, case, precision, _cheb(block=block)).profiles()
        err = envelope_error(got, oracle_profiles(oracle, case))
        assert err.max() <= TOL[precision], (block, err.max())


@PREC
def test_cheb_batch_equals_single(engine, precision):
    """Structures of different sizes, grids and expansion orders in one
    batch: each is bitwise the same as built alone."""
    from pripps_amd._native import PreparedInput
    cases = [make_case(n, 500 + n, nq=nq, qmax=qm) for n, nq, qm in
             ((90, 41, 0.5), (1400, 300, 0.8), (5, 7, 0.3), (2600, 500, 0.45), (0, 10, 0.5),
              (700, 640, 0.6))]
    t = _cheb(block=64, span=256)
    batch = engine.build_batch([PreparedInput(*c) for c in cases], precision, t)
    for c, bm in zip(cases, batch):
        assert np.array_equal(build(engine, c, precision, t).profiles(), bm.profiles())


@pytest.mark.parametrize("precision,tune", [("fp32", None), ("fp64", None), ("fp32", "q64")])
def test_graph_replays_match_build(engine, precision, tune):
    """recompute(graph=True) replays the build's kernel launches as one
    captured HIP graph per run.  Every replay -- with eager replays,
    intensity() calls and host synchronisations in between -- reproduces the
    build's profiles bit for bit; the Chebyshev plan (counters reset by the
    kernel), the FP64 recurrence plan and an FP32 recurrence plan (counter
    zeroed by a fill kernel inside the graph) alike.  A later build is
    replayed from a new capture, and the earlier model can no longer be
    replayed."""
    from pripps_amd._native import tuning
    from pripps_amd.errors import UsageError
    t = tuning(q_per_lane=64) if tune == "q64" else None
    case = make_case(1800, 41, nq=300)
    m = build(engine, case, precision, t)
    ref = m.profiles()
    for k in range(3):
        _, total = m.recompute(2 + k, graph=True)
        assert total > 0
        assert np.array_equal(m.profiles(), ref), k
        m.recompute(1)
        m.intensity(1.0, 0.0)
        assert np.array_equal(m.profiles(), ref), k
    other = make_case(900, 42, nq=120)
    m2 = build(engine, other, precision, t)
    ref2 = m2.profiles()
    m2.recompute(3, graph=True)
    assert np.array_equal(m2.profiles(), ref2)
    assert np.array_equal(m.profiles(), ref)  # the earlier model's profiles stay
    with pytest.raises(UsageError):
        m.recompute(1, graph=True)
