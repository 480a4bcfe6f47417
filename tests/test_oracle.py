"""Pins the CPU oracle (oracle/debye_oracle.c) before it is trusted.

Every check is chosen so that a plausible slip in the restatement (a dropped
diagonal, a wrong factor 2, a transposed cross term, a wrong form-factor
convention) fails one of them:

* the reference's own known answers: q-grid endpoints and nq = 1
  (src/q_sampler.rs:173-185); q = 0 closed form (sum f)^2 and two atoms
  2 + 2 sinc(q r) (src/multipole/mod.rs:907-967); Cromer-Mann s^2 convention
  value (src/formfactor.rs:759-774); table interpolation
  (src/formfactor.rs:729-740); united/linear resolution at q = 0
  (src/formfactor.rs:860-891); Fraser F(0) = rho V and linear volume scaling
  (src/solvent/fraser.rs:142-163); and the PepsiSAXS Iat profile on 4LZT within
  4 % (src/lib.rs:808-854, the reference's only direct Debye KAT);
* mathematics: one atom gives f^2, disabled sets give zero, invariance under
  translation / rotation / permutation, the colocated-set identity (the two
  summation modes), and mpmath brute force on tiny inputs.
"""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from conftest import REFDATA, random_cluster


def kinds_tables(q, ff_by_kind):
    return np.array([[f(x) for x in q] for f in ff_by_kind])


def test_linear_q_endpoints_and_single(oracle):
    q = oracle.linear_q_values(0.0, 1.0, 11)
    assert len(q) == 11 and q[0] == 0.0 and q[10] == 1.0
    assert list(oracle.linear_q_values(0.42, 0.42, 1)) == [0.42]
    # the exact expression qmin + (qmax - qmin) * i / (nq - 1)
    q = oracle.linear_q_values(0.0, 0.5, 500)
    assert all(q[i] == 0.0 + (0.5 - 0.0) * i / 499.0 for i in range(500))


def test_sinc(oracle):
    assert oracle.sinc(0.0) == 1.0
    assert oracle.sinc(9.9e-9) == 1.0
    assert oracle.sinc(2.0) == math.sin(2.0) / 2.0


@pytest.mark.parametrize("mode", [0, 1])
def test_q_zero_closed_form(oracle, mode):
    pos = np.array([[0, 0, 0], [2, 0, 0], [0, 3, 0], [0, 0, 1.5]], dtype=float)
    q = np.linspace(0.0, 0.1, 5)
    c = 1.0
    prof = oracle.build_model(pos, [0, 0, 0, 0], q, np.full((1, 5), c), mode=mode)
    assert abs(prof[0, 0] - (4 * c) ** 2) < 1e-10


@pytest.mark.parametrize("mode", [0, 1])
def test_two_atoms(oracle, mode):
    pos = np.array([[0, 0, 0], [4, 0, 0]], dtype=float)
    q = oracle.linear_q_values(0.01, 0.5, 20)
    prof = oracle.build_model(pos, [0, 0], q, np.ones((1, 20)), mode=mode)
    expect = 2.0 + 2.0 * np.sin(q * 4.0) / (q * 4.0)
    assert np.max(np.abs(prof[0] - expect)) < 1e-14


def test_one_atom_and_empty(oracle):
    q = np.linspace(0, 0.5, 7)
    f = np.linspace(3, 4, 7)[None, :]
    for mode in (0, 1):
        prof = oracle.build_model(np.zeros((1, 3)), [0], q, f, f * 0.5, np.full(7, 2.0), [0.3],
                                  mode=mode)
        assert np.allclose(prof[0], f[0] ** 2, rtol=0, atol=1e-15)
        assert np.allclose(prof[1], (0.5 * f[0]) ** 2, rtol=0, atol=1e-15)
        assert np.allclose(prof[2], (0.6) ** 2, rtol=0, atol=1e-15)
        assert np.allclose(prof[3], f[0] * 0.5 * f[0], rtol=0, atol=1e-15)
        empty = oracle.build_model(np.zeros((0, 3)), [], q, f, mode=mode)
        assert np.all(empty == 0.0)


def test_disabled_sets_are_zero(oracle):
    pos = random_cluster(40, 1)
    kind = np.arange(40) % 2
    q = np.linspace(0, 0.5, 9)
    ffa = np.vstack([np.full(9, 6.0), np.full(9, 8.0)])
    for mode in (0, 1):
        p = oracle.build_model(pos, kind, q, ffa, mode=mode)
        assert np.all(p[1:] == 0.0)
        p = oracle.build_model(pos, kind, q, ffa, ffa * 0.3, mode=mode)
        assert np.all(p[2] == 0.0) and np.all(p[4] == 0.0) and np.all(p[5] == 0.0)
        assert np.any(p[1] != 0.0) and np.any(p[3] != 0.0)


def _case(n, seed):
    rng = np.random.default_rng(seed)
    pos = random_cluster(n, seed)
    kind = rng.integers(0, 3, size=n)
    q = np.linspace(0.0, 0.6, 13)
    ffa = np.array([[6 + k - 0.5 * x for x in q] for k in range(3)])
    ffe = np.array([[1.5 * (k + 1) * math.exp(-0.7 * x * x) for x in q] for k in range(3)])
    water = np.array([9 - x for x in q])
    w = np.where(rng.uniform(size=n) < 0.4, rng.uniform(size=n), 0.0)
    return pos, kind, q, ffa, ffe, water, w


def test_modes_agree(oracle):
    pos, kind, q, ffa, ffe, water, w = _case(150, 3)
    a = oracle.build_model(pos, kind, q, ffa, ffe, water, w, mode=0)
    b = oracle.build_model(pos, kind, q, ffa, ffe, water, w, mode=1)
    scale = np.sqrt(np.abs(a[[0, 1, 2, 0, 0, 1]]) * np.abs(a[[0, 1, 2, 1, 2, 2]]))
    assert np.max(np.abs(a - b) / scale) < 1e-12


def test_invariances(oracle):
    pos, kind, q, ffa, ffe, water, w = _case(90, 4)
    base = oracle.build_model(pos, kind, q, ffa, ffe, water, w)
    th = 0.7
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    moved = pos @ R.T + np.array([13.0, -7.5, 2.25])
    perm = np.random.default_rng(9).permutation(len(pos))
    other = oracle.build_model(moved[perm], kind[perm], q, ffa, ffe, water, w[perm])
    scale = np.sqrt(np.abs(base[[0, 1, 2, 0, 0, 1]]) * np.abs(base[[0, 1, 2, 1, 2, 2]]))
    assert np.max(np.abs(base - other) / scale) < 1e-12


def test_mpmath_brute_force(oracle):
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 40
    pos, kind, q, ffa, ffe, water, w = _case(12, 5)
    prof = oracle.build_model(pos, kind, q, ffa, ffe, water, w, mode=1)
    n = len(pos)
    for m in (0, 4, 12):
        fa = [mpmath.mpf(ffa[kind[i], m]) for i in range(n)]
        fe = [mpmath.mpf(ffe[kind[i], m]) for i in range(n)]
        fh = [mpmath.mpf(w[i]) * mpmath.mpf(water[m]) for i in range(n)]
        sets = (fa, fe, fh)
        pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
        for p, (x, y) in enumerate(pairs):
            tot = mpmath.mpf(0)
            for i in range(n):
                for j in range(n):
                    d = [mpmath.mpf(pos[i, k]) - mpmath.mpf(pos[j, k]) for k in range(3)]
                    r = mpmath.sqrt(sum(v * v for v in d))
                    xq = mpmath.mpf(q[m]) * r
                    s = mpmath.mpf(1) if abs(xq) < mpmath.mpf("1e-8") else mpmath.sin(xq) / xq
                    tot += sets[x][i] * sets[y][j] * s
            assert abs(float(tot) - prof[p, m]) <= 1e-13 * max(1.0, abs(float(tot)))


def test_combine_order(oracle):
    t = np.array([[5.0], [2.0], [3.0], [1.5], [0.5], [0.25]])
    vs, cd = 1.7, -0.4
    expect = 5.0 + vs * vs * 2.0 + cd * cd * 3.0 - 2.0 * vs * 1.5 + 2.0 * cd * 0.5 - 2.0 * vs * cd * 0.25
    assert oracle.combine(t, vs, cd)[0] == expect
    assert oracle.combine(t, 0.0, 0.0)[0] == 5.0  # intensity(0, 0) == atoms


def test_gaussian_s2_convention(oracle):
    s2 = 0.25 / (16.0 * math.pi * math.pi)
    assert abs(oracle.eval_gaussian(0.5, [(6.0, 20.0)], 0.0) - 6.0 * math.exp(-20.0 * s2)) < 1e-10
    assert oracle.eval_gaussian(0.0, [(6.0, 20.0)], 0.0) == 6.0


def test_single_gaussian_q2(oracle):
    a, b = 5.486, 0.42
    assert oracle.eval_single_gaussian(a, b, 0.0) == a
    assert abs(oracle.eval_single_gaussian(a, b, 0.5) - a * math.exp(-b * 0.25)) < 1e-15


def test_table_interpolation(oracle):
    xs, ys = [0.0, 0.5, 1.0, 1.5], [1.0, 2.0, 3.0, 6.0]
    for x, y in ((0.0, 1.0), (0.5, 2.0), (1.0, 3.0), (1.5, 6.0), (0.25, 1.5), (0.75, 2.5)):
        assert abs(oracle.interpolate(xs, ys, x) - y) < 1e-15


def test_polynomial(oracle):
    assert oracle.eval_polynomial([1, 2, 3, 4, 5], 0.0) == 1.0
    assert oracle.eval_polynomial([1, 2, 3, 4, 5], 1.0) == 15.0


def test_united_and_fraser(oracle):
    # UNITED: centre O (Constant 1.0) + 2 H (Constant 0.2) at 1 A: F(0) = 1.4
    qs, table = oracle.united_table(0.0, 2.0, lambda q: 1.0, [(lambda q: 0.2, 1.0, 2.0)])
    assert table[0] == 1.4 and qs[0] == 0.0 and qs[-1] == 2.0 and len(qs) == 101
    a, b = oracle.fraser_params(16.43, 0.334, 4.0 * math.pi)
    assert a == 0.334 * 16.43
    assert oracle.eval_single_gaussian(a, b, 0.0) == 0.334 * 16.43
    a2, _ = oracle.fraser_params(2.0 * 16.43, 0.334, 4.0 * math.pi)
    assert a2 == 2.0 * a


def test_pepsi_iat_kat(oracle):
    """Debye vacuum on 4lzt.xyz with poly_amino_acid.yaml vs PepsiSAXS Iat
    divided by the log's intensity scaling, max rel < 0.04 (src/lib.rs:808-854)."""
    import pripps_amd
    p = pripps_amd.Pripps(os.path.join(REFDATA, "poly_amino_acid.yaml"))
    from pripps_amd.structure import structure_from_file
    st = structure_from_file(os.path.join(REFDATA, "4lzt.xyz"), p.formfactors.kinds())
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
    q = oracle.linear_q_values(ref[0, 0], ref[-1, 0], len(ref))
    assert np.max(np.abs(q - ref[:, 0])) < 1e-12
    used = sorted(set(st.ids.tolist()))
    remap = {k: r for r, k in enumerate(used)}
    kind = np.array([remap[i] for i in st.ids])
    ffa = p.formfactors.table(used, q)
    prof = oracle.build_model(st.pos, kind, q, ffa, mode=0)
    rel = np.abs((prof[0] - ref[:, 1]) / ref[:, 1])
    assert rel.max() < 0.04, rel.max()
