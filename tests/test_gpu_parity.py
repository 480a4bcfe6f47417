"""GPU parity of the Debye engine against the CPU oracle (through the C ABI).

Error measure for profile P_XY at q: |P_gpu - P_oracle| / sqrt(|P_XX| |P_YY|).
For the three self profiles this is the plain relative error; for the cross
profiles, which change sign, it is relative to their Cauchy-Schwarz envelope
(|P_XY| <= sqrt(P_XX P_YY)).  Tolerances (the bar of the precision modes):
  fp32 (compensated)  1e-5
  fp64, fp64_plain    1e-12
"""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from conftest import REFDATA, random_cluster

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-5, "fp64": 1e-12, "fp64_plain": 1e-12}
PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def envelope_error(got, ref):
    diag = np.abs(ref[:3])
    err = np.zeros_like(ref)
    for p, (x, y) in enumerate(PAIRS):
        env = np.sqrt(diag[x] * diag[y])
        env = np.where(env > 0, env, 1.0)
        err[p] = np.abs(got[p] - ref[p]) / env
    return err


def make_case(n, seed, nq=37, qmax=0.5, q=None, solvent=True, kinds=4, excluded=True):
    rng = np.random.default_rng(seed)
    pos = random_cluster(n, seed) if n > 0 else np.zeros((0, 3))
    kind = rng.integers(0, kinds, size=n).astype(np.int32)
    if q is None:
        q = np.array([0.0 + (qmax - 0.0) * i / float(max(nq - 1, 1)) for i in range(nq)])
    ffa = np.array([[6.0 + k - 2.0 * x * x for x in q] for k in range(kinds)])
    ffe = np.array([[0.334 * (10 + 3 * k) * math.exp(-0.3 * x * x) for x in q] for k in range(kinds)]) \
        if (solvent and excluded) else None
    water = np.array([10.0 - 3.0 * x for x in q]) if solvent else None
    w = np.where(rng.uniform(size=n) < 0.3, rng.uniform(size=n), 0.0) if solvent else None
    return pos, kind, q, ffa, ffe, water, w


def build(engine, case, precision, tuning=None):
    from pripps_amd._native import PreparedInput
    pos, kind, q, ffa, ffe, water, w = case
    inp = PreparedInput(pos, kind, q, ffa, ffe, water, w)
    return engine.build(inp, precision, tuning)


def oracle_profiles(oracle, case):
    pos, kind, q, ffa, ffe, water, w = case
    return oracle.build_model(pos, kind, q, ffa, ffe, water, w, mode=oracle.COMPENSATED)


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp64_plain"])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 63, 64, 65, 130, 700])
def test_small_structures(engine, oracle, precision, n):
    case = make_case(n, 100 + n)
    print(f"case n={n} precision={precision}", flush=True)
    m = build(engine, case, precision)
    got = m.profiles()
    ref = oracle_profiles(oracle, case)
    err = envelope_error(got, ref)
    assert err.max() <= TOL[precision], (err.max(), np.unravel_index(err.argmax(), err.shape))


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp64_plain"])
def test_nonuniform_q(engine, oracle, precision):
    q = np.exp(np.linspace(math.log(0.005), math.log(1.0), 45))
    case = make_case(300, 7, q=q)
    print(f"nonuniform precision={precision}", flush=True)
    got = build(engine, case, precision).profiles()
    err = envelope_error(got, oracle_profiles(oracle, case))
    assert err.max() <= TOL[precision], err.max()


@pytest.mark.parametrize("precision,n,nq,want_q", [
    # Both precisions take the Chebyshev kernel (256) from 3e7 pair-q terms
    # (N^2 nq / 2); below, FP32 64-point strips when nq > 32, else 32, and
    # FP64 32
    ("fp32", 500, 240, 256), ("fp32", 500, 239, 64), ("fp32", 1000, 61, 256),
    ("fp32", 1000, 59, 64), ("fp32", 700, 20, 32), ("fp32", 2000, 150, 256),
    ("fp32", 2048, 32, 256),
    ("fp64", 1000, 59, 32), ("fp64", 1000, 61, 256), ("fp64", 500, 240, 256),
    ("fp64", 500, 239, 32), ("fp64", 2048, 20, 256),
])
def test_default_plan_boundaries(engine, oracle, precision, n, nq, want_q):
    """The default kernel and strip length at their boundaries in N and nq."""
    case = make_case(n, 40 + nq, nq=nq)
    print(f"plan boundary n={n} nq={nq} precision={precision}", flush=True)
    m = build(engine, case, precision)
    assert m.stats().q_per_lane == want_q
    err = envelope_error(m.profiles(), oracle_profiles(oracle, case))
    assert err.max() <= TOL[precision], (err.max(), np.unravel_index(err.argmax(), err.shape))


@pytest.mark.parametrize("precision,qpl", [("fp32", 64), ("fp32", 128), ("fp32", 256),
                                           ("fp64", 32), ("fp64", 64)])
def test_large_phases_and_coarse_grid(engine, oracle, precision, qpl):
    """q up to 6 1/A on a coarse grid over an 80 A wide cluster: phases q r / 2 pi
    up to ~80 revolutions (the seeds' exact reduction of large arguments), and a
    half step h = dq r / 2 above 1/8 turn for most pairs (the seeds' reduced
    half-step path instead of the small-h polynomial)."""
    from pripps_amd._native import tuning
    pos, kind, q, ffa, ffe, water, w = make_case(600, 23, nq=130, qmax=6.0)
    pos = pos * (40.0 / np.max(np.linalg.norm(pos - pos.mean(axis=0), axis=1)))
    case = (pos, kind, q, ffa, ffe, water, w)
    dq = q[1] - q[0]
    span = np.max(np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=2))
    assert dq * span / (4 * math.pi) > 0.125  # the host cannot prove small h
    print(f"large phases q_per_lane={qpl} precision={precision}", flush=True)
    m = build(engine, case, precision, tuning(q_per_lane=qpl))
    assert m.stats().q_per_lane == qpl
    err = envelope_error(m.profiles(), oracle_profiles(oracle, case))
    print(f"large phases q_per_lane={qpl} precision={precision}: max envelope error {err.max():.2e}")
    assert err.max() <= TOL[precision], (err.max(), np.unravel_index(err.argmax(), err.shape))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("qpl", [32, 64, 128])
def test_strip_shapes(engine, oracle, precision, qpl):
    from pripps_amd._native import tuning
    if precision == "fp64" and qpl == 128:
        pytest.skip("fp64 strips hold at most 64 q-points")
    case = make_case(400, 11, nq=200)
    print(f"strip q_per_lane={qpl} precision={precision}", flush=True)
    got = build(engine, case, precision, tuning(q_per_lane=qpl, block=48, span=128)).profiles()
    err = envelope_error(got, oracle_profiles(oracle, case))
    assert err.max() <= TOL[precision], err.max()


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp64_plain"])
def test_solvent_variants(engine, oracle, precision):
    for solvent, excluded in ((False, False), (True, False), (True, True)):
        case = make_case(200, 13, solvent=solvent, excluded=excluded)
        got = build(engine, case, precision).profiles()
        ref = oracle_profiles(oracle, case)
        if not solvent:
            assert np.all(got[1:] == 0.0)
        if not excluded:
            assert np.all(got[1] == 0.0) and np.all(got[3] == 0.0) and np.all(got[5] == 0.0)
        err = envelope_error(got, ref)
        assert err.max() <= TOL[precision], (solvent, excluded, err.max())


@pytest.mark.parametrize("precision", ["fp32", "fp64", "fp64_plain"])
def test_coincident_atoms_and_q_edges(engine, oracle, precision):
    pos = random_cluster(50, 17)
    pos = np.vstack([pos, pos[:5]])  # five duplicated atoms: r = 0, sinc = 1
    rng = np.random.default_rng(3)
    kind = rng.integers(0, 2, size=len(pos)).astype(np.int32)
    for q in (np.array([0.0]), np.array([0.3]), np.array([0.0, 0.25]), np.array([1e-4, 0.7, 0.1])):
        ffa = np.array([[6.0 + k for _ in q] for k in range(2)])
        case = (pos, kind, q, ffa, ffa * 0.4, np.full(len(q), 5.0), np.linspace(0, 1, len(pos)))
        got = build(engine, case, precision).profiles()
        err = envelope_error(got, oracle_profiles(oracle, case))
        assert err.max() <= TOL[precision], (q, err.max())


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_many_kinds(engine, oracle, precision):
    case = make_case(500, 19, kinds=12)
    got = build(engine, case, precision).profiles()
    err = envelope_error(got, oracle_profiles(oracle, case))
    assert err.max() <= TOL[precision], err.max()


def test_intensity_is_reference_combine(engine, oracle):
    case = make_case(300, 23)
    m = build(engine, case, "fp32")
    prof = m.profiles()
    for vs, cd in ((1.0, 0.03), (0.0, 0.0), (1.7, 0.0), (0.95, -2.0), (1.12, 4.0)):
        total, atoms, excluded, hydration = m.intensity(vs, cd)
        assert np.array_equal(total, oracle.combine(prof, vs, cd))
        assert np.array_equal(atoms, prof[0]) and np.array_equal(excluded, prof[1])
        assert np.array_equal(hydration, prof[2])
    assert np.array_equal(m.intensity(0.0, 0.0)[0], prof[0])  # intensity(0,0) == atoms


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_equals_single(engine, precision):
    from pripps_amd._native import PreparedInput, tuning
    cases = [make_case(n, 200 + n, nq=nq, qmax=qm) for n, nq, qm in
             ((90, 41, 0.5), (400, 130, 0.8), (5, 7, 0.3), (1300, 64, 0.45), (0, 10, 0.5))]
    t = tuning(q_per_lane=32 if precision == "fp64" else 64, block=64, span=256)
    inputs = [PreparedInput(*c) for c in cases]
    batch = engine.build_batch(inputs, precision, t)
    for c, bm in zip(cases, batch):
        single = build(engine, c, precision, t)
        assert np.array_equal(single.profiles(), bm.profiles())


def test_reference_fit_known_answers(engine):
    """src/fit.rs:619-682 through the device fit: a model whose combine gives
    total = [1, c1, 10 c2] at q = [0, 1, 2]."""
    from pripps_amd._native import FIT_GRID, FIT_NELDER_MEAD
    from pripps_amd.errors import ValidationError
    q = np.array([0.0, 1.0, 2.0])
    prof = np.zeros((6, 3))
    prof[0, 0] = 1.0
    prof[3, 1] = -0.5  # -2 c1 AE = c1
    prof[4, 2] = 5.0   # 2 c2 AH = 10 c2
    m = engine.from_profiles(q, prof)
    r = m.fit(q, [1.0, 1.25, 0.4], None, FIT_GRID, 7, 6)
    assert abs(r.volume_scale - 1.25) < 1e-12 and abs(r.contrast_density - 0.04) < 1e-12
    assert r.chi2 < 1e-24
    r = m.fit(q, [1.0, 1.25, 12.0], None, FIT_GRID, 7, 6, (0.5, 2.0), (0.0, 2.0))
    assert abs(r.volume_scale - 1.25) < 1e-12 and abs(r.contrast_density - 1.2) < 1e-12
    assert r.chi2 < 1e-24
    with pytest.raises(ValidationError, match="at least 2 points"):
        m.fit(q, [1.0, 1.25, 0.4], None, FIT_GRID, 1, 6)
    with pytest.raises(ValidationError, match="lower bound"):
        m.fit(q, [1.0, 1.25, 12.0], None, FIT_GRID, 7, 6, (0.5, 2.0), (2.0, 0.0))
    with pytest.raises(ValidationError, match="lower bound"):
        m.fit(q, [1.0, 1.25, 0.4], None, FIT_NELDER_MEAD, c1_bounds=(2.0, 0.5))


@pytest.mark.parametrize("method", ["grid", "nelder-mead"])
def test_fit_matches_oracle(engine, oracle, method):
    from pripps_amd._native import FIT_GRID, FIT_NELDER_MEAD
    case = make_case(500, 29, nq=120, qmax=0.5)
    m = build(engine, case, "fp64")
    prof = m.profiles()
    q = case[2]
    exp = np.loadtxt(os.path.join(REFDATA, "lyzexp.dat"), comments="#")
    qe = exp[:, 0][exp[:, 0] <= 0.5]
    # synthetic "experiment": the model at (1.03, 0.05) times 2.5e-3 plus structure
    tot = oracle.combine(prof, 1.03, 0.05)
    ie = np.interp(qe, q, tot) * 2.5e-3 * (1.0 + 0.01 * np.sin(40 * qe))
    se = 0.01 * ie + 1e-6
    code = FIT_GRID if method == "grid" else FIT_NELDER_MEAD
    r = m.fit(qe, ie, se, code, 31, 21)
    ref = oracle.fit(lambda c1, c2: oracle.combine(prof, c1, c2), q, qe, ie, se,
                     method=method, c1_points=31, c2_points=21)
    if method == "grid":
        assert r.volume_scale == ref[0] and r.contrast_density == ref[1]
        assert abs(r.chi2 - ref[2]) <= 1e-10 * abs(ref[2])
    else:
        assert abs(r.volume_scale - ref[0]) < 1e-6 and abs(r.contrast_density - ref[1]) < 1e-6
        assert abs(r.chi2 - ref[2]) <= 1e-8 * abs(ref[2])
    assert abs(r.scale - ref[3]) <= 1e-8 * abs(ref[3])


def test_public_api_end_to_end(engine, oracle, tmp_path):
    """Pripps.debye_model on an XYZ file with FoXS form factors, excluded
    volume and explicit hydration weights, against the oracle on the same
    host tables."""
    import pripps_amd
    from pripps_amd.structure import write_xyz
    from pripps_amd.synth import globule
    names, pos, w = globule(800, 20.0, 5)
    path = tmp_path / "g.xyz"
    write_xyz(path, names, pos)
    p = pripps_amd.Pripps(os.path.join(REFDATA, "foxs_formfactors.yaml"), engine=engine)
    model = p.debye_model(str(path), qmin=0.0, qmax=0.5, nq=101, excluded="foxs",
                          hydration="weights", hydration_weights=w, precision="fp32")
    it = model.intensity(1.0, 0.03)
    assert np.array_equal(it.q, model.q_values)
    ids = p.names_to_ids(names)
    inp = p.prepare_input(pos, ids, it.q, excluded="foxs", hydration="weights",
                          hydration_weights=w)
    ref = oracle.build_model(pos, inp.kind, inp.q, inp.atom_ff, inp.excluded_ff, inp.water_ff,
                             inp.weights)
    err = envelope_error(model.profiles(), ref)
    assert err.max() <= 1e-5, err.max()
    fit = model.fit(it.q[5:], it.total[5:] * 3.0, method="grid", c1_points=16, c2_points=11)
    assert abs(fit.scale - 3.0) < 1e-3
