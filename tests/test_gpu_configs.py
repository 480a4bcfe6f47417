"""Parity at the benchmark configurations' full sizes (GPU).

Sizes are the configurations' own; the oracle is evaluated on a few sampled
q where the full grid would take minutes, and the remaining q are checked
through the FP32-vs-FP64 agreement of the GPU modes.  Tolerances: FP32
compensated 1e-5, FP64 1e-12 (envelope error, see test_gpu_parity.py).
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import REFDATA
from test_gpu_parity import envelope_error

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def foxs_engine(engine):
    import pripps_amd
    return pripps_amd.Pripps(os.path.join(REFDATA, "foxs_formfactors.yaml"), engine=engine)


def oracle_at(oracle, inp, sel):
    return oracle.build_model(np.column_stack([inp.x, inp.y, inp.z]), inp.kind, inp.q[sel],
                              inp.atom_ff[:, sel],
                              None if inp.excluded_ff is None else inp.excluded_ff[:, sel],
                              None if inp.water_ff is None else inp.water_ff[sel], inp.weights,
                              mode=oracle.COMPENSATED, nthreads=min(16, os.cpu_count() or 1))


def test_vacuum1k_fp64(engine, oracle, foxs_engine):
    from pripps_amd.synth import vacuum1k
    wl = vacuum1k()
    inp = foxs_engine.prepare_workload(wl)
    assert inp.excluded_ff is None and inp.water_ff is None
    ref = oracle_at(oracle, inp, np.arange(len(wl.q)))
    for prec, tol in (("fp64", 1e-12), ("fp64_plain", 1e-12), ("fp32", 1e-5)):
        print(f"vacuum1k {prec}", flush=True)
        got = engine.build(inp, prec).profiles()
        err = envelope_error(got, ref)
        assert err.max() <= tol, (prec, err.max())


def test_solvent10k(engine, oracle, foxs_engine):
    from pripps_amd.synth import solvent10k
    wl = solvent10k()
    inp = foxs_engine.prepare_workload(wl)
    sel = np.linspace(1, len(wl.q) - 1, 6).round().astype(int)
    ref = oracle_at(oracle, inp, sel)
    p32 = engine.build(inp, "fp32").profiles()
    p64 = engine.build(inp, "fp64").profiles()
    assert envelope_error(p32[:, sel], ref).max() <= 1e-5
    assert envelope_error(p64[:, sel], ref).max() <= 1e-12
    assert envelope_error(p32, p64).max() <= 1e-5
    m = engine.build(inp, "fp32")
    total = m.intensity(1.0, 0.03)[0]
    assert np.array_equal(total, oracle.combine(m.profiles(), 1.0, 0.03))


def test_helix60k_log_q(engine, oracle, foxs_engine):
    from pripps_amd.synth import helix60k
    wl = helix60k()
    inp = foxs_engine.prepare_workload(wl)
    sel = np.array([150, 900])
    print("helix60k oracle", flush=True)
    ref = oracle_at(oracle, inp, sel)
    print("helix60k fp64", flush=True)
    p64 = engine.build(inp, "fp64").profiles()
    print("helix60k fp32", flush=True)
    p32 = engine.build(inp, "fp32").profiles()
    assert envelope_error(p64[:, sel], ref).max() <= 1e-12
    assert envelope_error(p32[:, sel], ref).max() <= 1e-5
    assert envelope_error(p32, p64).max() <= 1e-5


def test_batch_subset_alone_equals_in_mix(engine, oracle, foxs_engine):
    from pripps_amd.synth import batch
    wls = batch(count=24)
    inps = [foxs_engine.prepare_workload(w) for w in wls]
    b32 = engine.build_batch(inps, "fp32")
    b64 = engine.build_batch(inps, "fp64")
    k = int(np.argmax([w.n_atoms for w in wls]))
    # the largest fills a launch alone, so its units are weighted by wave age
    for i in sorted({0, 5, 11, 23, k}):
        alone = engine.build(inps[i], "fp32").profiles()
        assert np.array_equal(alone, b32[i].profiles()), i
    assert np.array_equal(engine.build(inps[k], "fp64").profiles(), b64[k].profiles())
    for m32, m64 in zip(b32, b64):
        assert envelope_error(m32.profiles(), m64.profiles()).max() <= 1e-5
    sel = np.array([1, len(wls[k].q) // 2, len(wls[k].q) - 1])
    ref = oracle_at(oracle, inps[k], sel)
    assert envelope_error(b64[k].profiles()[:, sel], ref).max() <= 1e-12


def test_server_alone_equals_in_mix_and_admission(engine, foxs_engine):
    from pripps_amd.errors import OutOfDeviceMemoryError
    from pripps_amd.server import Server
    from pripps_amd.synth import Workload, globule, linear_q, radius_for
    rng = np.random.default_rng(8)
    inps = []
    for s in range(12):
        n = int(rng.integers(300, 4000))
        names, pos, w = globule(n, radius_for(n), 1000 + s)
        wl = Workload(f"r{s}", names, pos, linear_q(0.0, 0.5, int(rng.integers(200, 800))), w)
        inps.append(foxs_engine.prepare_workload(wl))
    with Server(engine, "fp32") as srv:
        ids = [srv.submit(i) for i in inps]
        res = [srv.wait(r) for r in ids]
    for inp, (m, lat) in zip(inps, res):
        assert lat > 0
        assert np.array_equal(m.profiles(), engine.build(inp, "fp32").profiles())
        total = m.intensity(1.05, 0.5)[0]
        assert np.all(np.isfinite(total))
    # admission: a budget smaller than one request fails that request only.
    # A 500-atom, 100-q request needs well under 1 MB of device memory; the
    # 20,000-atom, 800-q one over 10 MB (its coefficient records alone).
    snames, spos, sw = globule(500, radius_for(500), 76)
    tiny = foxs_engine.prepare_workload(Workload("small", snames, spos, linear_q(0, 0.5, 100), sw))
    with Server(engine, "fp32", memory_budget=4_000_000) as srv:
        small = srv.submit(tiny)
        m, _ = srv.wait(small)
        assert m.nq == 100
        bnames, bpos, bw = globule(20000, radius_for(20000), 77)
        big = foxs_engine.prepare_workload(Workload("big", bnames, bpos, linear_q(0, 0.5, 800), bw))
        rid = srv.submit(big)
        with pytest.raises(OutOfDeviceMemoryError):
            srv.wait(rid)


def test_server_bad_request_fails_alone(engine, foxs_engine):
    """A request whose coordinates hold a NaN fails with a validation error;
    the requests batched with it are built again one by one and are the same
    as built alone.  A second concurrent wait for one id is refused."""
    from pripps_amd.errors import NotReadyError, UsageError, ValidationError
    from pripps_amd.server import Server
    from pripps_amd.synth import Workload, globule, linear_q, radius_for
    rng = np.random.default_rng(9)
    inps = []
    for s in range(8):
        n = int(rng.integers(300, 2000))
        names, pos, w = globule(n, radius_for(n), 2000 + s)
        wl = Workload(f"r{s}", names, pos, linear_q(0.0, 0.5, int(rng.integers(100, 400))), w)
        inps.append(foxs_engine.prepare_workload(wl))
    bad = 5
    inps[bad].y[17] = np.nan
    with Server(engine, "fp32") as srv:
        ids = [srv.submit(i) for i in inps]
        got = {}
        for k, rid in enumerate(ids):
            if k == bad:
                with pytest.raises(ValidationError, match="finite"):
                    srv.wait(rid)
            else:
                got[k] = srv.wait(rid)[0]
        # one waiter at a time: a wait that timed out leaves the id waitable
        slow = srv.submit(inps[0])
        try:
            m, _ = srv.wait(slow, timeout_ms=0)  # finished already
        except NotReadyError:
            m, _ = srv.wait(slow)
        assert np.array_equal(m.profiles(), got[0].profiles())
        with pytest.raises(UsageError, match="unknown request"):
            srv.wait(slow)
    for k, m in got.items():
        assert np.array_equal(m.profiles(), engine.build(inps[k], "fp32").profiles())


def test_served_stream_c5(engine, oracle, foxs_engine):
    """The served configuration at its own sizes: 40 requests of the request
    stream (N log-U[300, 30000], the largest at 29,013 atoms; nq U[200, 800];
    q on [0, 0.5]; FoXS solvent with surface hydration weights), offered as a
    Poisson stream at 2,000 requests/s -- above the loop's saturation, so
    requests queue and batch.  Each is followed by five intensity() calls at
    random knobs.  A sample of results equals the same structure built alone
    on a fresh engine, bit for bit, and the largest request is within 1e-5 of
    the oracle at three q."""
    import time
    from pripps_amd.engine import Engine
    from pripps_amd.server import Server
    from pripps_amd.synth import ELEMENTS, globule, linear_q, radius_for, request_stream
    ns, nqs, knobs, seeds = request_stream(40, seed=21)
    assert ns.max() >= 25000
    lut = foxs_engine.names_to_ids(list(ELEMENTS))
    inps = []
    for n, nq, sd in zip(ns, nqs, seeds):
        names, pos, w = globule(int(n), radius_for(int(n)), int(sd), solvent=True)
        idx = np.array([ELEMENTS.index(e) for e in names])
        inps.append(foxs_engine.prepare_input(pos, lut[idx], linear_q(0.0, 0.5, int(nq)),
                                              excluded="foxs", hydration="weights",
                                              hydration_weights=w))
    print("c5: 40 requests prepared", flush=True)
    arrivals = np.cumsum(np.random.default_rng(3).exponential(1.0 / 2000.0, size=len(inps)))
    totals = {}
    with Server(engine, "fp32") as srv:
        t0 = time.perf_counter()
        ids = []
        for k, inp in enumerate(inps):
            delay = t0 + arrivals[k] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(inp))
        models = []
        for k, rid in enumerate(ids):
            m, lat = srv.wait(rid)
            assert lat > 0
            for vs, cd in knobs[k]:
                tot = m.intensity(vs, cd)[0]
                assert np.all(np.isfinite(tot))
            totals[k] = tot
            models.append(m)
        wall = time.perf_counter() - t0
        print(f"c5: 40 requests served in {wall:.2f} s", flush=True)
        profs = [m.profiles() for m in models]
    fresh = Engine(0)
    big = int(np.argmax(ns))
    sample = sorted({big, *np.linspace(0, len(inps) - 1, 8).round().astype(int).tolist()})
    for k in sample:
        alone = fresh.build(inps[k], "fp32")
        assert np.array_equal(profs[k], alone.profiles()), k
        assert np.array_equal(totals[k], alone.intensity(*knobs[k][-1])[0]), k
    sel = np.array([1, int(nqs[big]) // 2, int(nqs[big]) - 1])
    ref = oracle_at(oracle, inps[big], sel)
    assert envelope_error(profs[big][:, sel], ref).max() <= 1e-5
