"""CPU tests of the grid solvent generators (host side), restating the
reference's own tests of src/solvent/grid.rs and src/solvent/grid_excluded.rs,
plus a brute-force check of the k-d tree nearest-surface search."""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import random_cluster

from pripps_amd import solvent as S


def brute_surface_distance(points, centers, radii):
    out = np.full(len(points), np.inf)
    for c, r in zip(centers, radii):
        d = points - c
        out = np.minimum(out, np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]) - r)
    return out


def test_nearest_surface_distance_matches_brute_force():
    rng = np.random.default_rng(4)
    centers = random_cluster(200, 9)
    radii = rng.uniform(1.2, 2.1, size=200)
    pts = S._grid_points(centers, 4.0, 0.9)
    brute = brute_surface_distance(pts, centers, radii)
    for cutoff in (radii.max(), radii.max() + 1.8 + 3.0):
        fast = S.nearest_surface_distance(pts, centers, radii, cutoff)
        inside = brute <= cutoff - radii.max()
        assert np.array_equal(fast[inside], brute[inside])
        assert np.all(fast[~inside] >= cutoff - radii.max())
    # the generators select exactly the brute-force point sets
    exv = S.grid_excluded_dummies(centers, radii, 0.9)
    pts_e = S._grid_points(centers, radii.max() + 0.9, 0.9)
    assert np.array_equal(exv, pts_e[brute_surface_distance(pts_e, centers, radii) <= 0.0])
    hyd = S.grid_hydration_dummies(centers, radii, 1.5, 3.0, 1.4)
    pts_h = S._grid_points(centers, radii.max() + 1.4 + 3.0 + 1.5, 1.5)
    d = brute_surface_distance(pts_h, centers, radii)
    assert np.array_equal(hyd, pts_h[(d > 1.4) & (d < 4.4)])


def test_single_atom_shell_is_spherical():
    """src/solvent/grid.rs:121-143: radius 1.7, spacing 1.5, thickness 3, probe 1.4."""
    shell = S.grid_hydration_dummies(np.zeros((1, 3)), np.array([1.7]), 1.5, 3.0, 1.4)
    assert len(shell) > 50
    r = np.linalg.norm(shell, axis=1)
    assert np.all(r > 3.0) and np.all(r < 6.2)


def test_no_shell_points_inside_protein():
    """src/solvent/grid.rs:145-163."""
    pos = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    radii = np.array([1.7, 1.7])
    shell = S.grid_hydration_dummies(pos, radii, 1.5, 3.0, 1.4)
    assert len(shell) > 0
    assert np.all(brute_surface_distance(shell, pos, radii) > 1.39)


def test_single_atom_fills_sphere():
    """src/solvent/grid_excluded.rs:91-108."""
    d = S.grid_excluded_dummies(np.zeros((1, 3)), np.array([1.7]), 0.5)
    assert len(d) > 10
    assert np.all(np.linalg.norm(d, axis=1) - 1.7 <= 0.0)


def test_grid_volume_approximates_sphere():
    """src/solvent/grid_excluded.rs:110-130: within 15 % at spacing 0.3."""
    d = S.grid_excluded_dummies(np.zeros((1, 3)), np.array([1.7]), 0.3)
    ratio = len(d) * 0.3 ** 3 / (4.0 / 3.0 * math.pi * 1.7 ** 3)
    assert 0.85 < ratio < 1.15


def test_grid_excluded_form_factor_at_zero_and_empty():
    """src/solvent/grid_excluded.rs:132-166: F(0) = rho V_cell; no atoms, no points."""
    f = S.grid_excluded_form_factor(np.array([0.0, 0.3]), 0.334, 1.0)
    assert abs(f[0] - 0.334) < 1e-10
    assert abs(f[1] - 0.334 * math.exp(-0.09 / (4 * math.pi))) < 1e-15
    assert len(S.grid_excluded_dummies(np.zeros((0, 3)), np.zeros(0), 1.0)) == 0
    assert len(S.grid_hydration_dummies(np.zeros((0, 3)), np.zeros(0), 1.0, 3.0, 1.4)) == 0
    assert S.molecules_per_cell(3.10516) == 1.0


def test_union_input_layout(foxs):
    """Grid dummies join the structure as extra kinds: excluded cells carry
    only the cell form factor, hydration cells only the water weight."""
    from pripps_amd.errors import ValidationError
    pos = random_cluster(60, 5)
    ids = np.zeros(60, dtype=np.int64)
    q = np.linspace(0, 0.4, 7)
    inp = foxs.prepare_input(pos, ids, q, excluded="grid", hydration="grid", spacing=2.5,
                             excluded_spacing=1.2)
    kinds = inp.kind
    assert inp.atom_ff.shape == (3, 7) and inp.excluded_ff.shape == (3, 7)
    assert np.all(inp.atom_ff[1:] == 0) and np.all(inp.excluded_ff[0] == 0)
    assert np.all(inp.excluded_ff[2] == 0)
    assert np.allclose(inp.excluded_ff[1], S.grid_excluded_form_factor(q, 0.334, 1.2))
    n_e, n_h = int((kinds == 1).sum()), int((kinds == 2).sum())
    assert n_e > 0 and n_h > 0 and len(kinds) == 60 + n_e + n_h
    assert np.all(inp.weights[kinds != 2] == 0)
    assert np.allclose(inp.weights[kinds == 2], (2.5 / 3.10516) ** 3)
    # grid excluded + explicit per-atom hydration weights
    w = np.linspace(0, 1, 60)
    inp2 = foxs.prepare_input(pos, ids, q, excluded="grid", hydration="weights",
                              hydration_weights=w)
    assert np.array_equal(inp2.weights[:60], w) and np.all(inp2.weights[60:] == 0)

def test_grid_models_need_radius_and_water(tmp_path):
    import pripps_amd
    from pripps_amd.errors import ValidationError
    y = tmp_path / "noradius.yaml"
    y.write_text("C:\n  formfactor: !Constant 6.0\nH2O:\n  formfactor: !Constant 10.0\n")
    p = pripps_amd.Pripps(str(y))
    pos = random_cluster(20, 2)
    ids = np.full(20, p.formfactors.find_name("C"), dtype=np.int64)
    q = np.linspace(0, 0.3, 4)
    with pytest.raises(ValidationError, match="has no `radius`; required by Hydration::Grid"):
        p.prepare_input(pos, ids, q, hydration="grid")
    with pytest.raises(ValidationError, match="required by ExcludedVolume::Grid"):
        p.prepare_input(pos, ids, q, excluded="grid")
    y.write_text("C:\n  formfactor: !Constant 6.0\n  radius: 1.7\n")
    p = pripps_amd.Pripps(str(y))
    with pytest.raises(ValidationError, match="requires an `H2O` entry"):
        p.prepare_input(pos, np.zeros(20, dtype=np.int64), q, hydration="grid")
    inp = p.prepare_input(pos, np.zeros(20, dtype=np.int64), q, excluded="grid")
    assert inp.water_ff is None and inp.weights is None and len(inp.kind) > 20
