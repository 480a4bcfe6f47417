"""CPU tests of the Direct Fourier transform's host side and oracle: the
reference's q-grid tests (src/q_grid.rs:327-485), its grouping and phase-table
tests (src/explicit.rs:468-515), mathematical pins of the oracle, and the
argument guards of ``Pripps.intensity`` (src/lib.rs:354-403), which fail before
any GPU work."""
from __future__ import annotations

import math
import os
import time

import numpy as np
import pytest

from conftest import REFDATA

from pripps_amd import errors as E
from pripps_amd.direct import QGrid, average_duplicates, group_by_qnorm, miller_indices, vector_norms

L405 = 405.0


# ---------------------------------------------------------------------------
# q grids (src/q_grid.rs tests)
# ---------------------------------------------------------------------------
def test_first_bin_holds_the_unit_and_face_diagonal_vectors():
    grid = QGrid.lattice([L405] * 3, 0.023, 2**62)
    assert grid.vectors_per_bin() == [9]
    expected = 2.0 * math.pi / L405 * (3.0 + 6.0 * math.sqrt(2.0)) / 9.0
    assert abs(grid.q()[0] - expected) <= 1e-12 * expected


def test_bin_counts_match_a_brute_force_count():
    grid = QGrid.lattice([L405] * 3, 0.5, 2**62)
    assert grid.vectors_per_bin()[:8] == [9, 31, 49, 105, 175, 225, 301, 381]


def test_bin_counts_match_an_independent_enumeration():
    # every integer triple in the upper half space, binned by round(|q| / w)
    L = 60.0
    grid = QGrid.lattice([L] * 3, 0.8, 2**62)
    w = 2.0 * math.pi / L
    nmax = int(0.8 * L / (2 * math.pi)) + 1
    counts = {}
    for h in range(-nmax, nmax + 1):
        for k in range(-nmax, nmax + 1):
            for l in range(-nmax, nmax + 1):
                if not (h > 0 or (h == 0 and (k > 0 or (k == 0 and l > 0)))):
                    continue
                q = 2 * math.pi * math.sqrt(h * h + k * k + l * l) / L
                if q <= 0.8:
                    b = math.floor(q / w + 0.5)
                    counts[b] = counts.get(b, 0) + 1
    assert grid.vectors_per_bin() == [counts[b] for b in sorted(counts)]


def test_the_cap_thins_large_bins_only_and_keeps_their_q():
    full = QGrid.lattice([L405] * 3, 0.5, 2**62)
    capped = QGrid.lattice([L405] * 3, 0.5, 100)
    for a, k in zip(full.vectors_per_bin(), capped.vectors_per_bin()):
        assert k == min(a, 100)
    assert np.array_equal(full.q(), capped.q())


def _assert_isotropic(grid, b, eps):
    idx = grid.indices()[np.asarray(grid._bins) == b].astype(float)
    t = np.zeros((3, 3))
    for v in idx:
        t += np.outer(v, v) / (v @ v) / len(idx)
    assert np.allclose(t, np.eye(3) / 3.0, atol=eps), t


def test_a_capped_bin_still_spans_the_sphere():
    grid = QGrid.lattice([L405] * 3, 0.5, 50)
    _assert_isotropic(grid, len(grid.q()) - 1, 0.03)


def test_a_sampled_bin_spans_the_sphere():
    grid = QGrid.lattice([6900.0] * 3, 0.5, 200)
    _assert_isotropic(grid, len(grid.q()) - 2, 0.03)


@pytest.mark.parametrize("box,qmax", [([L405] * 3, 0.01), ([L405, 0.0, L405], 0.5),
                                      ([L405, float("nan"), L405], 0.5), ([L405] * 3, float("inf")),
                                      ([L405] * 3, float("nan")), ([L405] * 3, -0.5)])
def test_impossible_grids_are_errors(box, qmax):
    with pytest.raises(E.UsageError):
        QGrid.lattice(box, qmax, 10)


def test_a_narrow_last_bin_survives_any_cap():
    q_max = 0.054_190_196_758_008_29
    small = QGrid.lattice([6900.0] * 3, q_max, 16)
    large = QGrid.lattice([6900.0] * 3, q_max, 200)
    assert np.array_equal(small.q(), large.q())


def test_huge_boxes_are_sampled_not_enumerated():
    L = 6900.0
    t0 = time.time()
    small = QGrid.lattice([L] * 3, 0.5, 16)
    large = QGrid.lattice([L] * 3, 0.5, 200)
    assert time.time() - t0 < 60.0
    assert np.array_equal(small.q(), large.q())
    assert np.all(np.diff(small.q()) > 0)
    counts = small.vectors_per_bin()
    assert all(n <= 16 for n in counts)
    assert sum(1 for n in counts if n < 12) < 3, counts
    width = 2.0 * math.pi / L
    qm = small.q()[len(small.q()) // 2]
    k = round(qm / width)
    expected = 0.75 * ((k + 0.5) ** 4 - (k - 0.5) ** 4) / ((k + 0.5) ** 3 - (k - 0.5) ** 3) * width
    assert abs(qm - expected) <= 1e-6 * expected


def test_miller_grid_and_grouping():
    idx = miller_indices(2)
    assert idx.shape == (26, 3)
    assert idx[:2].tolist() == [[1, 0, 0], [2, 0, 0]]
    assert idx[-1].tolist() == [2, 2, -2]
    qn = vector_norms(idx, (200.0, 200.0, 200.0))
    group, gq = group_by_qnorm(qn)
    # cubic box: |n|^2 in {1, 2, 3, 4, 8, 12} -> six groups of 3, 6, 4, 3, 6, 4 vectors
    assert np.bincount(group).tolist() == [3, 6, 4, 3, 6, 4]
    assert np.all(np.diff(gq) > 0)
    assert gq[0] == qn[0]


# ---------------------------------------------------------------------------
# grouping and phase table (src/explicit.rs tests)
# ---------------------------------------------------------------------------
def test_average_duplicates():
    v = [(1.0, 1.0), (3.0, 0.1), (2.0, 3.0), (2.0, 4.0), (1.0, 2.0)]
    assert average_duplicates(v) == [(1.0, 1.5), (2.0, 3.5), (3.0, 0.1)]


def _phase_table_structure():
    L = 37.0
    pos = np.array([[math.fmod(7.3 * t, L) - 5.0, math.fmod(3.1 * t * t, L),
                     math.fmod(11.7 * t + 2.0, L)] for t in map(float, range(12))])
    kind = np.array([i % 2 for i in range(12)], dtype=np.int32)
    return L, pos, kind


def test_phase_table_matches_sin_cos(oracle):
    L, pos, kind = _phase_table_structure()
    grid = QGrid.lattice([L] * 3, 0.9, 40)
    idx = grid.indices()
    ffa = np.array([np.full(len(idx), 1.0), np.full(len(idx), 0.5)])
    slow = oracle.direct_rows(pos, kind, [L] * 3, idx, ffa, num_molecules=4,
                              mode=oracle.DIRECT_REFERENCE)
    fast = oracle.direct_rows(pos, kind, [L] * 3, idx, ffa, num_molecules=4,
                              mode=oracle.DIRECT_TABLE)
    ext = oracle.direct_rows(pos, kind, [L] * 3, idx, ffa, num_molecules=4,
                             mode=oracle.DIRECT_EXTENDED)
    for a, b in ((fast, slow), (ext, slow)):
        ia, ib = a[:, 0] ** 2 + a[:, 1] ** 2, b[:, 0] ** 2 + b[:, 1] ** 2
        for x, y in ((ia, ib), (a[:, 6], b[:, 6])):
            assert np.all(np.abs(x - y) <= 1e-10 * np.maximum(np.abs(x), np.abs(y)) + 1e-10)


def test_oracle_one_atom_and_self_terms(oracle):
    idx = miller_indices(5)
    ff = np.array([np.linspace(2.0, 1.0, len(idx))])
    r = oracle.direct_rows([[3.3, -4.1, 7.9]], [0], [50.0] * 3, idx, ff)
    assert np.allclose(r[:, 0] ** 2 + r[:, 1] ** 2, ff[0] ** 2, rtol=1e-15)
    # whole box: self term = |A|^2
    rng = np.random.default_rng(3)
    pos = rng.uniform(-20, 20, size=(30, 3))
    ffa = rng.uniform(1, 3, size=(2, len(idx)))
    kind = rng.integers(0, 2, size=30).astype(np.int32)
    r = oracle.direct_rows(pos, kind, [50.0] * 3, idx, ffa)
    assert np.allclose(r[:, 6], r[:, 0] ** 2 + r[:, 1] ** 2, rtol=1e-12)
    # N identical molecules at the same place: A = N F, self = N |F|^2 = |A|^2 / N
    mol = pos[:5]
    stack = np.vstack([mol] * 4)
    r = oracle.direct_rows(stack, np.tile(kind[:5], 4), [50.0] * 3, idx, ffa, num_molecules=4)
    assert np.allclose(r[:, 6], (r[:, 0] ** 2 + r[:, 1] ** 2) / 4.0, rtol=1e-12)


def test_oracle_lattice_translation_and_inversion(oracle):
    rng = np.random.default_rng(5)
    L = np.array([40.0, 50.0, 60.0])
    pos = rng.uniform(-30, 30, size=(25, 3))
    kind = rng.integers(0, 3, size=25).astype(np.int32)
    idx = QGrid.lattice(L, 0.6, 30).indices()
    ffa = rng.uniform(1, 3, size=(3, len(idx)))
    base = oracle.direct_rows(pos, kind, L, idx, ffa)
    shifted = pos + rng.integers(-3, 4, size=pos.shape) * L  # each atom by a lattice vector
    r = oracle.direct_rows(shifted, kind, L, idx, ffa)
    scale = np.abs(base[:, :2]).max()
    assert np.max(np.abs(r[:, :2] - base[:, :2])) <= 1e-12 * scale
    inv = oracle.direct_rows(-pos, kind, L, idx, ffa)
    assert np.max(np.abs(inv[:, 0] - base[:, 0])) <= 1e-12 * scale
    assert np.max(np.abs(inv[:, 1] + base[:, 1])) <= 1e-12 * scale


# ---------------------------------------------------------------------------
# Pripps.intensity guards (no GPU work happens before they fail)
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def xyz_pair(tmp_path_factory):
    d = tmp_path_factory.mktemp("direct")
    yaml = d / "x.yaml"
    yaml.write_text("X: { formfactor: !Constant 1.0 }\nY: { formfactor: !Constant 0.5 }\n")
    xyz = d / "six.xyz"
    xyz.write_text("6\nb\nX 0 0 0\nX 0 0 4\nX 30 10 0\nX 30 10 4\nX 70 50 0\nX 70 50 4\n")
    mixed = d / "mixed.xyz"
    mixed.write_text("4\nb\nX 0 0 0\nY 0 0 4\nY 30 10 0\nX 30 10 4\n")
    return str(yaml), str(xyz), str(mixed)


def test_intensity_guards(xyz_pair):
    import pripps_amd
    yaml, xyz, mixed = xyz_pair
    p = pripps_amd.Pripps(yaml)
    with pytest.raises(E.UsageError, match="multipole"):
        p.intensity(xyz)
    with pytest.raises(ValueError, match="unknown model"):
        p.intensity(xyz, model="fourier")
    with pytest.raises(TypeError):
        p.intensity(xyz, model=3)
    with pytest.raises(E.UsageError, match="requires the `direct` scheme"):
        p.intensity(xyz, model="debye", num_molecules=3)
    with pytest.raises(E.UsageError, match="at least 1"):
        p.intensity(xyz, model="direct", num_molecules=0, box_sides=100.0)
    with pytest.raises(E.UsageError, match="vacuum runs only"):
        p.intensity(xyz, model="direct", num_molecules=3, box_sides=100.0, excluded="fraser")
    with pytest.raises(E.ValidationError, match="not a multiple"):
        p.intensity(xyz, model="direct", num_molecules=4, box_sides=100.0)
    with pytest.raises(E.ValidationError, match="identical"):
        p.intensity(mixed, model="direct", num_molecules=2, box_sides=100.0)
    with pytest.raises(E.UsageError, match="box side lengths"):
        p.intensity(xyz, model="direct")
    with pytest.raises(E.UsageError, match="XTC"):
        p.intensity(xyz, model="direct", box_sides=100.0, xtc="traj.xtc")
    with pytest.raises(ValueError, match="box_sides"):
        p.intensity(xyz, model="direct", box_sides=(1.0, 2.0))
