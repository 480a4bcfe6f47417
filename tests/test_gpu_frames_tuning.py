"""GPU tests of the trajectory-frames average (batched launch + HIP graph
replay) and of the autotuner's tuning table."""
from __future__ import annotations

import os

import numpy as np
import pytest

from conftest import REFDATA, random_cluster

pytestmark = pytest.mark.gpu


def _two_frames(n, scale):
    """The reference test's synthetic frames (src/lib.rs:991-1015): positions
    in nm, x = 0.10 i, y = 0.15 (i mod 7), z = 0.20 (i mod 5), times scale."""
    i = np.arange(n, dtype=np.float32)
    pos = np.stack([np.float32(scale) * np.float32(0.10) * i,
                    np.float32(scale) * np.float32(0.15) * (i % 7),
                    np.float32(scale) * np.float32(0.20) * (i % 5)], axis=1)
    return pos.astype(np.float64)


def test_trajectory_stride_and_weights(engine, tmp_path):
    """Restates trajectory_stride_and_weights (src/lib.rs:1017-1099): a weight
    file selecting one frame gives that frame alone, uniform weights give the
    mean of the two frames, and stride 2 keeps frame 0 renormalised."""
    import pripps_amd
    p = pripps_amd.Pripps(os.path.join(REFDATA, "poly_amino_acid.yaml"), engine=engine)
    template = os.path.join(REFDATA, "4lzt.xyz")
    st = pripps_amd.structure_from_file(template, p.formfactors.kinds())
    n = len(st.ids)
    frames = np.stack([_two_frames(n, 1.0), _two_frames(n, 1.2)])  # nm

    def write(name, w):
        path = tmp_path / name
        path.write_text(f"{w[0]}\n{w[1]}\n")
        return path

    def run(weights, stride, precision="fp64"):
        m = p.trajectory_average(template, frames, qmin=0.05, qmax=0.3, nq=8, units="nm",
                                 weights=weights, stride=stride, precision=precision)
        return m.intensity(1.0, 0.03).total, m

    i_f0, m0 = run(write("w0.txt", [1.0, 0.0]), 1)
    i_f1, _ = run(write("w1.txt", [0.0, 1.0]), 1)
    i_avg, _ = run(None, 1)
    i_s2, ms = run(None, 2)

    def close(a, b):
        return np.all(np.abs(a - b) <= 1e-9 * np.maximum(np.abs(a), 1.0))

    assert not close(i_f0, i_f1)
    assert close(i_avg, 0.5 * (i_f0 + i_f1))
    assert close(i_s2, i_f0)
    assert ms.stats().frames == 1 and m0.stats().frames == 2
    # a frame alone through the ordinary build gives the same profiles bit for bit
    alone = p.debye_model_from_arrays(frames[0] * 10.0, st.ids, qmin=0.05, qmax=0.3, nq=8,
                                      precision="fp64")
    assert np.array_equal(alone.profiles(), m0.profiles())
    for prec in ("fp32",):
        a32, _ = run(write("w2.txt", [1.0, 0.0]), 1, prec)
        b32 = p.debye_model_from_arrays(frames[0] * 10.0, st.ids, qmin=0.05, qmax=0.3, nq=8,
                                        precision=prec).intensity(1.0, 0.03).total
        assert np.array_equal(a32, b32)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_frames_graph_replays_match_per_frame_builds(engine, foxs, precision):
    """Seven jittered frames of a solvated globule in chunks of two: three full
    chunks replay one captured HIP graph, the last runs eagerly.  Every result
    matches per-frame builds averaged on the host, and the chunking does not
    change the answer (same accumulation order)."""
    from pripps_amd.synth import globule
    names, pos, w = globule(700, 11.5, 91)
    ids = foxs.names_to_ids(names)
    rng = np.random.default_rng(5)
    frames = pos[None] + rng.normal(0, 0.5, size=(7,) + pos.shape)
    weights = rng.uniform(0.2, 2.0, size=7)
    q = np.linspace(0.0, 0.45, 150)
    kw = dict(q_values=q, excluded="foxs", hydration="weights", hydration_weights=w,
              precision=precision)
    ref = np.zeros((6, len(q)))
    for f in range(7):
        m = foxs.debye_model_from_arrays(frames[f], ids, **kw)
        ref += weights[f] * m.profiles()
    ref /= weights.sum()
    results = []
    for chunk in (2, 0, 2, 3):
        m = foxs.trajectory_average_from_arrays(ids, frames, weights=weights,
                                                max_frames_per_launch=chunk, **kw)
        s = m.stats()
        assert s.frames == 7
        if chunk == 2:
            assert s.frames_per_launch == 2 and s.graph_launches == 3
        if chunk == 0:
            assert s.frames_per_launch == 7 and s.graph_launches == 1
        prof = m.profiles()
        scale = np.abs(ref).max(axis=1, keepdims=True)
        # FP64: the same plans, so only the averaging order differs.  FP32
        # runs the Chebyshev kernel, whose strips follow a bound of the
        # largest pair distance; the frames path takes one bound over all
        # frames, so a frame's strips may differ from its own build's
        assert np.abs(prof - ref).max() <= (1e-13 if precision == "fp64" else 1e-6) * scale.max()
        results.append(prof)
        # eager work between graph launches
        foxs.debye_model_from_arrays(frames[1], ids, **kw).profiles()
    for r in results[1:]:
        assert np.array_equal(r, results[0])


def test_frames_validation(engine, foxs):
    from pripps_amd.errors import ValidationError
    pos = random_cluster(50, 3)
    ids = np.zeros(50, dtype=np.int64)
    frames = np.stack([pos, pos + 0.1])
    kw = dict(qmin=0.0, qmax=0.3, nq=10)
    with pytest.raises(ValidationError, match="zero or infinite"):
        foxs.trajectory_average_from_arrays(ids, frames, weights=[1.0, -1.0], **kw)
    with pytest.raises(ValidationError, match="averaged frames sum to 0"):
        foxs.trajectory_average_from_arrays(ids, frames, weights=[0.0, 1.0], stride=2, **kw)
    with pytest.raises(ValidationError, match="stride"):
        foxs.trajectory_average_from_arrays(ids, frames, stride=0, **kw)
    with pytest.raises(ValidationError, match="does not match"):
        foxs.trajectory_average_from_arrays(ids, frames[:, :40], **kw)
    with pytest.raises(ValidationError, match="number of weights"):
        foxs.trajectory_average_from_arrays(ids, frames, weights=[1.0, 1.0, 1.0], **kw)
    with pytest.raises(ValidationError):
        foxs.trajectory_average_from_arrays(ids, frames[:0], **kw)
    # a bad frame in a later chunk (40 frames run as chunks of 16): its
    # preparation fails while the first chunk's kernels are in flight
    many = np.stack([pos + 0.01 * k for k in range(40)])
    many[25, 7, 1] = np.nan
    with pytest.raises(ValidationError, match="frame 25.*finite"):
        foxs.trajectory_average_from_arrays(ids, many, **kw)
    # the engine is still usable, and the good frames average as before
    good = foxs.trajectory_average_from_arrays(ids, many[:24], **kw)
    again = foxs.trajectory_average_from_arrays(ids, many[:24], **kw)
    assert np.array_equal(good.intensity(1.0, 0.0).total, again.intensity(1.0, 0.0).total)


def test_autotune_table_applies_and_keeps_batches_deterministic(foxs):
    """The autotuner records one entry per size class; builds without explicit
    tuning then use it, bitwise like an explicit build with the winning plan,
    and alone equals in-batch."""
    from pripps_amd import _native as N
    from pripps_amd.engine import Engine
    from pripps_amd.errors import ValidationError
    eng = Engine(0)  # fresh table
    names_pos = [random_cluster(1500, 21), random_cluster(400, 22)]
    inps = [foxs.prepare_input(p, np.arange(len(p)) % 4, np.linspace(0, 0.5, 200))
            for p in names_pos]
    best, ms = eng.autotune(inps[0], "fp32", repeats=1)
    # span 0: a Chebyshev plan (q_per_lane 256) that sizes its work units itself
    assert ms > 0 and best.q_per_lane in (32, 64, 128, 256)
    assert best.span >= 64 or (best.q_per_lane == 256 and best.span == 0)
    table = eng.tuning_table()
    assert len(table) == 1
    t0 = table[0]
    assert t0["precision"] == N.FP32_COMPENSATED and t0["uniform"] == 1
    assert t0["log2_atoms"] == round(np.log2(1500)) and t0["log2_nq"] == 8
    auto = eng.build(inps[0], "fp32")
    s = auto.stats()
    assert (s.q_per_lane, s.block, s.span) == (best.q_per_lane, best.block, best.span)
    expl = eng.build(inps[0], "fp32", best)
    assert np.array_equal(auto.profiles(), expl.profiles())
    batch = eng.build_batch(inps, "fp32")
    assert np.array_equal(batch[0].profiles(), auto.profiles())
    assert np.array_equal(batch[1].profiles(), eng.build(inps[1], "fp32").profiles())
    # a saved table restores on another engine
    other = Engine(0)
    other.set_tuning_entry(t0)
    assert other.tuning_table() == table
    assert np.array_equal(other.build(inps[0], "fp32").profiles(), auto.profiles())
    bad = dict(t0, q_per_lane=128, uniform=0)
    with pytest.raises(ValidationError):
        other.set_tuning_entry(bad)


def test_fp32_variants_agree(engine, foxs):
    """Inner-loop variants 1 (paired loads), 3 (default) and 4 (path fixed per
    chunk) order their FP32 operations differently; each stays within the
    FP32 tolerance of the FP64 result, and each is reproducible."""
    from pripps_amd import _native as N
    from test_gpu_parity import envelope_error
    pos = random_cluster(900, 8)
    inp = foxs.prepare_input(pos, np.arange(900) % 4, np.linspace(0, 0.5, 128),
                             excluded="foxs")
    p64 = engine.build(inp, "fp64").profiles()
    for v in (1, 3, 4):
        t = N.tuning(q_per_lane=64, variant=v)
        a = engine.build(inp, "fp32", t).profiles()
        assert envelope_error(a, p64).max() <= 1e-5
        assert np.array_equal(a, engine.build(inp, "fp32", t).profiles())


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
def test_pepsi_kat_twenty_kinds_on_gpu(engine, oracle, precision, tol):
    """The reference's Debye KAT (src/lib.rs:808-854) through the engine: 4LZT
    as 129 residue beads of 20 kinds (poly_amino_acid.yaml), vacuum, against
    PepsiSAXS Iat within 4 % and against the oracle within the mode's
    tolerance."""
    import pripps_amd
    p = pripps_amd.Pripps(os.path.join(REFDATA, "poly_amino_acid.yaml"), engine=engine)
    st = pripps_amd.structure_from_file(os.path.join(REFDATA, "4lzt.xyz"), p.formfactors.kinds())
    assert len(set(st.ids.tolist())) == 20
    scale = None
    for line in open(os.path.join(REFDATA, "pepsi.log")):
        if "Intensity scaling" in line:
            scale = float(line.split()[-1])
    ref = []
    for line in open(os.path.join(REFDATA, "pepsi.out")):
        t = line.strip()
        if t and not t.startswith("#"):
            f = t.split()
            ref.append((float(f[0]), float(f[2]) / scale))
    ref = np.array(ref)
    m = p.debye_model(os.path.join(REFDATA, "4lzt.xyz"), qmin=ref[0, 0], qmax=ref[-1, 0],
                      nq=len(ref), precision=precision)
    atoms = m.intensity(1.0, 0.0).atoms
    assert np.abs((atoms - ref[:, 1]) / ref[:, 1]).max() < 0.04
    used = sorted(set(st.ids.tolist()))
    remap = {k: r for r, k in enumerate(used)}
    kind = np.array([remap[i] for i in st.ids])
    q = m.q_values
    orc = oracle.build_model(st.pos, kind, q, p.formfactors.table(used, q), mode=oracle.COMPENSATED)
    assert np.abs(m.profiles()[0] - orc[0]).max() <= tol * np.abs(orc[0]).max()
    assert np.abs((m.profiles()[0] - orc[0]) / orc[0]).max() <= tol


def test_fit_curve_and_fit_csv(engine, foxs, oracle, tmp_path):
    """write_fit_csv (src/fit.rs:548-560): i_fit is scale x the natural spline
    of the fitted total, computed on the device; checked against the oracle's
    spline of the same combined curve."""
    import pripps_amd
    pos = random_cluster(600, 31)
    ids = np.arange(600) % 4
    m = foxs.debye_model_from_arrays(pos, ids, qmin=0.0, qmax=0.5, nq=120, excluded="foxs",
                                     precision="fp64")
    exp = pripps_amd.read_exp_data(os.path.join(REFDATA, "lyzexp.dat"))
    qe = exp.q[exp.q <= 0.55]
    vs, cd = 1.03, 0.02
    got = m.fit_curve(qe, vs, cd)
    total = m.intensity(vs, cd).total
    ref = oracle.spline_eval(m.q_values, total, qe)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    e2 = pripps_amd.ExpData(qe, 2.5 * ref, np.full(len(qe), 0.5))
    fit = m.fit(e2.q, e2.intensity, e2.sigma, method="grid", c1_points=21, c2_points=11,
                c1_bounds=(0.93, 1.13), c2_bounds=(0.0, 0.1))
    path = tmp_path / "fit.csv"
    pripps_amd.write_fit_csv(path, e2, m, fit)
    lines = path.read_text().splitlines()
    assert lines[0] == "q,i_exp,i_fit,sigma" and len(lines) == len(qe) + 1
    rows = np.array([[float(x) for x in ln.split(",")] for ln in lines[1:]])
    assert np.allclose(rows[:, 0], qe, atol=6e-7, rtol=0)
    i_fit = fit.scale * m.fit_curve(qe, fit.volume_scale, fit.contrast_density)
    assert np.allclose(rows[:, 2], i_fit, atol=6e-7, rtol=1e-15)
    # no hydration layer: contrast_density has no effect and the grid keeps its
    # first (row-major) minimum, c2 = 0
    assert np.abs(fit.volume_scale - vs) < 1e-9 and fit.contrast_density == 0.0
    assert abs(fit.scale - 2.5) < 1e-9 and fit.chi2 < 1e-10
