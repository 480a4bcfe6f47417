"""GPU parity of the grid solvent models: the engine's union structure
(atoms plus dummy points of their own kinds) against the oracle's literal
restatement of the reference's set sums (src/debye.rs:90-128): self sums over
atoms, excluded and hydration sets, and full cross sums between them."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import random_cluster
from test_gpu_parity import TOL, envelope_error

from pripps_amd import solvent as S

pytestmark = pytest.mark.gpu


def set_sums(oracle, sets, q):
    """Six profiles from explicit sets: sets = (pos_A, fA[n][nq]), (pos_E, fE), (pos_H, fH)."""
    (pa, fa), (pe, fe), (ph, fh) = sets
    out = np.zeros((6, len(q)))
    for m, qq in enumerate(q):
        out[0, m] = oracle.self_intensity_at_q(pa, fa[:, m], qq)
        out[1, m] = oracle.self_intensity_at_q(pe, fe[:, m], qq) if len(pe) else 0.0
        out[2, m] = oracle.self_intensity_at_q(ph, fh[:, m], qq) if len(ph) else 0.0
        out[3, m] = oracle.cross_intensity_at_q(pa, fa[:, m], pe, fe[:, m], qq) if len(pe) else 0.0
        out[4, m] = oracle.cross_intensity_at_q(pa, fa[:, m], ph, fh[:, m], qq) if len(ph) else 0.0
        out[5, m] = (oracle.cross_intensity_at_q(pe, fe[:, m], ph, fh[:, m], qq)
                     if len(pe) and len(ph) else 0.0)
    return out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("excluded,hydration", [("grid", "grid"), ("foxs", "grid"),
                                                ("grid", "weights"), ("grid", None)])
def test_grid_solvent_matches_set_sums(engine, oracle, foxs, precision, excluded, hydration):
    rng = np.random.default_rng(12)
    n = 90
    pos = random_cluster(n, 41)
    names = ["C", "N", "O", "S"]
    ids = foxs.names_to_ids([names[k] for k in rng.integers(0, 4, size=n)])
    q = np.linspace(0.0, 0.5, 11)
    w = rng.uniform(0, 1, size=n) if hydration == "weights" else None
    kw = dict(excluded=excluded, hydration=hydration, hydration_weights=w, spacing=2.0,
              thickness=3.0, probe=1.4, excluded_spacing=1.3)
    print(f"grid case excluded={excluded} hydration={hydration} precision={precision}", flush=True)
    m = foxs.debye_model_from_arrays(pos, ids, q_values=q, precision=precision, **kw)
    got = m.profiles()
    # explicit sets as the reference builds them
    used = sorted(set(ids.tolist()))
    fa = np.array([foxs.formfactors.table([k], q)[0] for k in ids])
    radii = S.atom_radii(foxs.formfactors, ids, "test")
    if excluded == "grid":
        pe = S.grid_excluded_dummies(pos, radii, 1.3)
        fe = np.tile(S.grid_excluded_form_factor(q, 0.334, 1.3), (len(pe), 1))
    else:
        pe = pos
        tab = S.excluded_table("foxs", foxs.formfactors, used, q, 0.334)
        fe = np.array([tab[used.index(k)] for k in ids])
    water = S.water_form_factor(foxs.formfactors, q, "test")
    if hydration == "grid":
        ph = S.grid_hydration_dummies(pos, radii, 2.0, 3.0, 1.4)
        fh = np.outer(np.full(len(ph), S.molecules_per_cell(2.0)), water)
    elif hydration == "weights":
        ph, fh = pos, np.outer(w, water)
    else:
        ph, fh = np.zeros((0, 3)), np.zeros((0, len(q)))
    assert len(pe) > 0 and (hydration is None or len(ph) > 0)
    # 1. the union construction against the reference's own set sums; these
    #    sum ~1e6 products sequentially, so they carry ~1e-12 of their own
    ref = set_sums(oracle, ((pos, fa), (pe, fe), (ph, fh)), q)
    err = envelope_error(got, ref)
    assert err.max() <= max(TOL[precision], 1e-10), (err.max(), np.unravel_index(err.argmax(), err.shape))
    # 2. the precision bar against the compensated oracle on the same points
    inp = foxs.prepare_input(pos, ids, q, **kw)
    comp = oracle.build_model(np.stack([inp.x, inp.y, inp.z], 1), inp.kind, q, inp.atom_ff,
                              inp.excluded_ff, inp.water_ff, inp.weights, mode=oracle.COMPENSATED)
    err = envelope_error(got, comp)
    assert err.max() <= TOL[precision], (err.max(), np.unravel_index(err.argmax(), err.shape))
    assert envelope_error(comp, ref).max() <= 1e-10
