"""GPU tests of the public Debye API against the reference's own API tests
and the oracle:

* one-shot Pripps.intensity(model="debye") equals the cached model, total
  and contributions bit for bit (src/lib.rs:535-566), and Fraser's
  volume_scale is applied once (src/lib.rs:598-633);
* the model's q grid is the oracle's linear grid bit for bit
  (src/q_sampler.rs:24-29);
* per-kind excluded volume (src/solvent/per_kind_excluded.rs:16-39) on 4LZT
  against the oracle fed the YAML's `excluded_solvent` polynomials;
* the device Nelder-Mead on a problem with a known minimum and against the
  oracle's Nelder-Mead; a grid fit larger than the engine's kept fit scratch
  followed by a small one, both against the oracle;
* the trajectory average against the weighted mean of oracle profiles;
* the serving loop's request map under concurrent submits and waits.

Tolerances: FP32 compensated 1e-5, FP64 1e-12 (envelope error, see
test_gpu_parity.py)."""
from __future__ import annotations

import os
import threading

import numpy as np
import pytest

from conftest import REFDATA
from test_gpu_parity import envelope_error

pytestmark = pytest.mark.gpu


def _lysozyme(engine):
    import pripps_amd
    p = pripps_amd.Pripps(os.path.join(REFDATA, "poly_amino_acid.yaml"), engine=engine)
    xyz = os.path.join(REFDATA, "4lzt.xyz")
    st = pripps_amd.structure_from_file(xyz, p.formfactors.kinds())
    return p, xyz, st


def _surface_weights(n, seed=17):
    rng = np.random.default_rng(seed)
    return np.where(rng.uniform(size=n) < 0.4, rng.uniform(size=n), 0.0)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_debye_one_shot_equals_cached_model(engine, precision):
    """debye_solvent_matches_cached_model_for_matching_knobs (src/lib.rs:535-566),
    with explicit hydration weights in place of the SASA tessellation: the
    one-shot intensity at (1.0, 0.03) is the cached model's, bit for bit."""
    p, xyz, st = _lysozyme(engine)
    w = _surface_weights(len(st.ids))
    kw = dict(excluded="fraser", hydration="weights", hydration_weights=w, precision=precision)
    one = p.intensity(xyz, model="debye", qmin=0.0, qmax=0.3, nq=20, volume_scale=1.0,
                      contrast_density=0.03, **kw)
    cached = p.debye_model(xyz, qmin=0.0, qmax=0.3, nq=20, **kw).intensity(1.0, 0.03)
    for name in ("q", "total", "atoms", "excluded", "hydration"):
        assert np.array_equal(getattr(one, name), getattr(cached, name)), name
    assert one.atoms is not None and np.all(one.hydration > 0)
    # vacuum one-shot: no contributions, the total is the atoms profile
    vac = p.intensity(xyz, model="debye", qmin=0.0, qmax=0.3, nq=20, precision=precision)
    assert vac.atoms is None and vac.excluded is None and vac.hydration is None
    m = p.debye_model(xyz, qmin=0.0, qmax=0.3, nq=20, precision=precision)
    assert np.array_equal(vac.total, m.profiles()[0])


def test_fraser_volume_scale_is_not_applied_twice(engine):
    """fraser_volume_scale_is_not_applied_twice (src/lib.rs:598-633), Debye
    half: one-shot with volume_scale 1.7 equals the unit-scale cached model's
    intensity(1.7, 0.0), bit for bit."""
    p, xyz, _ = _lysozyme(engine)
    one = p.intensity(xyz, model="debye", qmin=0.0, qmax=0.2, nq=8, excluded="fraser",
                      volume_scale=1.7)
    unit = p.debye_model(xyz, qmin=0.0, qmax=0.2, nq=8, excluded="fraser")
    assert np.array_equal(one.total, unit.intensity(1.7, 0.0).total)
    assert not np.array_equal(one.total, unit.intensity(1.7 * 1.7, 0.0).total)


def test_model_q_values_are_the_oracle_grid(engine, oracle):
    p, xyz, _ = _lysozyme(engine)
    for qmin, qmax, nq in ((0.0, 0.5, 500), (0.0, 0.3, 20), (0.013, 0.7, 77)):
        m = p.debye_model(xyz, qmin=qmin, qmax=qmax, nq=nq)
        want = oracle.linear_q_values(qmin, qmax, nq)
        assert m.q_values.tobytes() == want.tobytes()
        assert m.intensity().q.tobytes() == want.tobytes()


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("fp64", 1e-12)])
def test_per_kind_excluded_matches_oracle(engine, oracle, precision, tol):
    """excluded="per-kind" on 4LZT with poly_amino_acid.yaml: every residue's
    dummy carries its kind's `excluded_solvent` form factor.  The oracle gets
    its own table, evaluated from the YAML coefficients with the oracle's
    polynomial (src/formfactor.rs:379-385)."""
    from test_host import _yaml_tagged
    p, xyz, st = _lysozyme(engine)
    raw = _yaml_tagged(os.path.join(REFDATA, "poly_amino_acid.yaml"))
    w = _surface_weights(len(st.ids), 23)
    m = p.debye_model(xyz, qmin=0.0, qmax=0.5, nq=101, excluded="per-kind", hydration="weights",
                      hydration_weights=w, precision=precision)
    used = sorted(set(st.ids.tolist()))
    kind = np.array([used.index(i) for i in st.ids], dtype=np.int32)
    q = oracle.linear_q_values(0.0, 0.5, 101)
    ffa = p.formfactors.table(used, q)
    ffe = np.array([[oracle.eval_polynomial(
        raw[p.formfactors.entries[k].kind.name]["excluded_solvent"]["coeff"], x) for x in q]
        for k in used])
    water = np.array([p.formfactors.form_factor_by_name("H2O").eval(float(x)) for x in q])
    ref = oracle.build_model(st.pos, kind, q, ffa, ffe, water, w)
    err = envelope_error(m.profiles(), ref)
    assert err.max() <= tol, err.max()
    assert np.all(ref[1] > 0)


def _theory3_model(engine):
    """A model whose combine is total = [1, c1, 10 c2] at q = [0, 1, 2]."""
    q = np.array([0.0, 1.0, 2.0])
    prof = np.zeros((6, 3))
    prof[0, 0] = 1.0
    prof[3, 1] = -0.5  # -2 c1 AE = c1
    prof[4, 2] = 5.0   # 2 c2 AH = 10 c2
    return q, prof, engine.from_profiles(q, prof)


@pytest.mark.parametrize("target", [(1.25, 0.04), (0.71, 0.083), (1.93, 0.002)])
def test_device_nelder_mead_known_minimum(engine, oracle, target):
    """chi2/N = 0 at the generating point; the device simplex contracts onto it
    to the precision of the reference's stopping rule (spread < 1e-10 in
    chi2/N), and follows the oracle's simplex (same steps, same order)."""
    from pripps_amd._native import FIT_NELDER_MEAD
    q, prof, m = _theory3_model(engine)
    data = np.array([1.0, target[0], 10.0 * target[1]])
    r = m.fit(q, data, None, FIT_NELDER_MEAD)
    assert r.chi2 < 1e-10
    assert abs(r.volume_scale - target[0]) < 3e-5 and abs(r.contrast_density - target[1]) < 3e-6
    ref = oracle.fit(lambda c1, c2: oracle.combine(prof, c1, c2), q, q, data)
    assert abs(r.volume_scale - ref[0]) < 1e-9 and abs(r.contrast_density - ref[1]) < 1e-9
    # a target outside the box ends exactly on the bound
    r = m.fit(q, [1.0, 3.0, 0.4], None, FIT_NELDER_MEAD)
    assert r.volume_scale == 2.0
    r = m.fit(q, [1.0, 1.0, 0.3], None, FIT_NELDER_MEAD)
    assert (r.volume_scale, r.contrast_density, r.chi2) == (1.0, 0.03, 0.0)


def test_grid_fit_beyond_kept_scratch_then_small(engine, oracle):
    """A 300 x 300 grid (90,000 chi2 values, more than the engine's initial
    fit scratch) and then a 31 x 21 grid on the same model: both are the
    oracle's grid fit."""
    from pripps_amd._native import FIT_GRID
    from test_gpu_parity import build, make_case
    case = make_case(400, 41, nq=90, qmax=0.5)
    m = build(engine, case, "fp64")
    prof = m.profiles()
    q = case[2]
    qe = np.linspace(0.01, 0.49, 60)
    tot = oracle.combine(prof, 1.07, 0.045)
    ie = np.interp(qe, q, tot) * 1e-3 * (1.0 + 0.02 * np.sin(30 * qe))
    se = 0.01 * np.abs(ie) + 1e-7

    def theory(c1, c2):
        return (prof[0] + c1 * c1 * prof[1] + c2 * c2 * prof[2] - 2.0 * c1 * prof[3]
                + 2.0 * c2 * prof[4] - 2.0 * c1 * c2 * prof[5])

    for n1, n2 in ((300, 300), (31, 21)):
        r = m.fit(qe, ie, se, FIT_GRID, n1, n2)
        ref = oracle.fit(theory, q, qe, ie, se, method="grid", c1_points=n1, c2_points=n2)
        assert r.volume_scale == ref[0] and r.contrast_density == ref[1], (n1, n2)
        assert abs(r.chi2 - ref[2]) <= 1e-10 * abs(ref[2])
        assert abs(r.scale - ref[3]) <= 1e-9 * abs(ref[3])


@pytest.mark.parametrize("precision,tol", [("fp32", 1e-5), ("fp64", 1e-12)])
def test_trajectory_average_matches_oracle(engine, oracle, foxs, precision, tol):
    """trajectory_average (src/trajectory.rs:167-251) with stride 2 and
    per-frame weights indexed by the original frame: the weighted mean of the
    oracle's profiles of frames 0, 2, 4, 6."""
    from pripps_amd.synth import globule
    names, pos, w = globule(600, 11.0, 31)
    ids = foxs.names_to_ids(names)
    rng = np.random.default_rng(6)
    frames = pos[None] + rng.normal(0, 0.6, size=(7,) + pos.shape)
    weights = rng.uniform(0.3, 1.7, size=7)
    q = np.linspace(0.0, 0.45, 140)
    m = foxs.trajectory_average_from_arrays(ids, frames, weights=weights, stride=2, q_values=q,
                                            excluded="foxs", hydration="weights",
                                            hydration_weights=w, precision=precision,
                                            max_frames_per_launch=2)
    assert m.stats().frames == 4
    inp = foxs.prepare_input(frames[0], ids, q, excluded="foxs", hydration="weights",
                             hydration_weights=w)
    ref = np.zeros((6, len(q)))
    for f in (0, 2, 4, 6):
        ref += weights[f] * oracle.build_model(frames[f], inp.kind, q, inp.atom_ff,
                                               inp.excluded_ff, inp.water_ff, inp.weights)
    ref /= weights[[0, 2, 4, 6]].sum()
    err = envelope_error(m.profiles(), ref)
    assert err.max() <= tol, err.max()


def test_server_waits_survive_concurrent_submits(engine):
    """One thread waits for a request while another submits 300 more (the
    request map rehashes several times meanwhile); a second concurrent wait
    for the waited id is refused with a usage error, and every request
    completes with its alone result."""
    from pripps_amd._native import PreparedInput
    from pripps_amd.errors import UsageError
    from pripps_amd.server import Server
    from test_gpu_parity import make_case
    # three large requests keep the lanes busy, so the waited one (the
    # fourth) is still queued while the small ones arrive
    bigs = [PreparedInput(*make_case(20000, 3 + k, nq=600)) for k in range(4)]
    big = bigs[-1]
    tiny = [PreparedInput(*make_case(3 + k % 5, 100 + k, nq=4 + k % 3)) for k in range(300)]
    results, errors = {}, []
    with Server(engine, "fp32") as srv:
        first = [srv.submit(b) for b in bigs[:-1]]
        rid = srv.submit(big)
        outcome = {}

        def waiter(key):
            try:
                outcome[key] = srv.wait(rid)[0]
            except UsageError as exc:
                outcome[key] = exc

        t1 = threading.Thread(target=waiter, args=("a",))
        t1.start()
        ids = [srv.submit(x) for x in tiny]
        waiter("b")
        t1.join()
        got = [v for v in outcome.values() if not isinstance(v, Exception)]
        refused = [v for v in outcome.values() if isinstance(v, Exception)]
        assert len(got) == 1 and len(refused) == 1, outcome
        assert "already being waited for" in str(refused[0]) or "unknown request" in str(refused[0])
        for k, r in enumerate(ids):
            try:
                results[k] = srv.wait(r)[0].profiles()
            except Exception as exc:  # noqa: BLE001
                errors.append((k, exc))
        big_prof = got[0].profiles()
        for r in first:
            srv.wait(r)
    assert not errors, errors[:3]
    from pripps_amd.engine import Engine
    fresh = Engine(0)
    assert np.array_equal(big_prof, fresh.build(big, "fp32").profiles())
    for k in (0, 7, 150, 299):
        assert np.array_equal(results[k], fresh.build(tiny[k], "fp32").profiles())
