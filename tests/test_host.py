"""CPU tests of the host side: form-factor YAML resolution, XYZ loading,
solvent tables, the synthetic workload generator, error mapping, and the
C-ABI library's exported symbols (loaded, not called: no GPU here)."""
from __future__ import annotations

import ctypes
import math
import os
import re

import numpy as np
import pytest

from conftest import REFDATA, REPO

import pripps_amd
from pripps_amd import errors as E
from pripps_amd.formfactor import read_formfactors, read_formfactors_merged
from pripps_amd.structure import parse_xyz_text, read_structure, resolve_ids, write_xyz


def test_library_exports_every_declared_symbol():
    declared = set()
    for name in ("pripps_debye.h", "pripps_direct.h"):
        header = open(os.path.join(REPO, "include", name)).read()
        header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
        declared |= set(re.findall(r"\b(pripps_[a-z_0-9]+)\s*\(", header))
    assert len(declared) >= 29
    lib = ctypes.CDLL(os.path.join(REPO, "pripps_amd", "libpripps_debye.so"))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    from pripps_amd import _native
    assert set(_native.EXPORTS) == declared


def test_foxs_yaml_resolution(oracle):
    ff = read_formfactors(os.path.join(REFDATA, "foxs_formfactors.yaml"))
    names = [k.name for k in ff.kinds()]
    assert names == sorted(names) and "H2O" in names and "C" in names
    c = ff.entries[ff.find_name("C")]
    ab = [(2.31, 20.8439), (1.02, 10.2075), (1.5886, 0.5687), (0.865, 51.6512)]
    for q in (0.0, 0.13, 0.5, 1.0):
        assert abs(c.formfactor.eval(q) - oracle.eval_gaussian(q, ab, 0.2156)) < 1e-14
    assert c.kind.displaced_volume == 16.44 and abs(c.kind.sigma - 3.16) < 1e-15
    s = ff.entries[ff.find_name("S")]
    assert abs(s.kind.displaced_volume - 4.0 / 3.0 * math.pi * 1.68 ** 3) < 1e-12
    water = ff.form_factor_by_name("H2O")
    assert water.qrange() == (0.0, 6.0) and len(water.points) == 101
    # H2O = F_O + 2 F_H sinc(0.9584 q) on the 101-point grid, linearly interpolated
    o_ab = [(3.0485, 13.2771), (2.2868, 5.7011), (1.5463, 0.3239), (0.867, 32.9089)]
    h_ab = [(0.493002, 10.5109), (0.322912, 26.1257), (0.140191, 3.14236), (0.04081, 57.7997)]
    qs, tab = oracle.united_table(0.0, 6.0, lambda x: oracle.eval_gaussian(x, o_ab, 0.2508),
                                  [(lambda x: oracle.eval_gaussian(x, h_ab, 0.003038), 0.9584, 2.0)])
    assert np.max(np.abs(np.array([p[1] for p in water.points]) - tab)) < 1e-12
    for q in (0.0, 0.031, 0.5, 1.0):
        assert abs(water.eval(q) - oracle.interpolate(qs, tab, q)) < 1e-12
    with pytest.raises(E.QOutOfRangeError):
        c.formfactor.eval(6.5)


def test_symbolic_excluded_solvent(tmp_path):
    """src/formfactor.rs:860-891: alias / linear / united resolution."""
    y = tmp_path / "sym.yaml"
    y.write_text("""
H:
  formfactor: !Constant 1.0
  excluded_solvent: !Constant 0.2
O:
  formfactor: !Constant 8.0
  excluded_solvent: !Constant 1.0
ALIAS:
  formfactor: !Constant 2.0
  excluded_solvent: !Alias H
LINEAR:
  formfactor: !Constant 3.0
  excluded_solvent: !Linear [[H, 2.0], [O, 3.0]]
UNITED:
  formfactor: !Constant 4.0
  excluded_solvent: !United { center: O, orbit: [[H, 1.0, 2]] }
""")
    ff = read_formfactors(y)
    ex = lambda n: ff.entries[ff.find_name(n)].excluded_solvent.eval(0.0)
    assert abs(ex("ALIAS") - 0.2) < 1e-15
    assert abs(ex("LINEAR") - 3.4) < 1e-15
    assert abs(ex("UNITED") - 1.4) < 1e-15


def test_merged_latest_wins_and_cross_file_alias(tmp_path):
    a = tmp_path / "base.yaml"
    b = tmp_path / "over.yaml"
    a.write_text("C:\n  formfactor: !Constant 6.0\nN:\n  formfactor: !Constant 7.0\n")
    b.write_text("C:\n  formfactor: !Constant 60.0\nCA:\n  formfactor: !Alias N\n")
    ff = read_formfactors_merged([a, b])
    assert len(ff) == 3
    ev = lambda n: ff.form_factor_by_name(n).eval(0.0)
    assert ev("C") == 60.0 and ev("N") == 7.0 and ev("CA") == 7.0


def test_polynomial_range_and_table(tmp_path):
    y = tmp_path / "p.yaml"
    y.write_text("P:\n  formfactor: !Polynomial { qrange: [0.0, 1.0], coeff: [1.0, 2.0, 3.0, 4.0, 5.0] }\n"
                 "T:\n  formfactor: !Table [[0.0, 1.0], [0.5, 2.0], [1.0, 3.0], [1.5, 6.0]]\n")
    ff = read_formfactors(y)
    p = ff.form_factor_by_name("P")
    assert p.eval(0.0) == 1.0 and p.eval(1.0) == 15.0
    with pytest.raises(E.QOutOfRangeError):
        p.eval(2.0)
    t = ff.form_factor_by_name("T")
    assert t.eval(0.25) == 1.5 and t.eval(0.75) == 2.5 and t.eval(1.5) == 6.0


def test_xyz_parsing(tmp_path):
    names, pos = parse_xyz_text("3\ncomment line\n; skip\n# skip\n\nC 1 2 3\nN 4.5 -1 0 extra\nbad 1 x 3\nO 0 0 0\n")
    assert names == ["C", "N", "O"]
    assert pos.tolist() == [[1, 2, 3], [4.5, -1, 0], [0, 0, 0]]
    p = tmp_path / "empty.xyz"
    p.write_text("; nothing\n")
    with pytest.raises(E.ValidationError, match="no atoms found"):
        read_structure(p)
    with pytest.raises(E.FileError):
        read_structure(tmp_path / "missing.xyz")
    ff = read_formfactors(os.path.join(REFDATA, "foxs_formfactors.yaml"))
    with pytest.raises(E.UnknownAtomError, match="unknown atom name: XX"):
        resolve_ids(["C", "XX"], ff.kinds())
    n2, p2 = parse_xyz_text(open(os.path.join(REFDATA, "4lzt.xyz")).read())
    assert len(n2) == 129


def test_solvent_tables():
    from pripps_amd.solvent import FOXS_DENOMINATOR, excluded_table, water_table
    ff = read_formfactors(os.path.join(REFDATA, "foxs_formfactors.yaml"))
    used = [ff.find_name(n) for n in ("C", "N", "O", "S")]
    q = np.linspace(0, 0.5, 11)
    t = excluded_table("foxs", ff, used, q, 0.334)
    for r, kid in enumerate(used):
        v = ff.entries[kid].kind.displaced_volume
        assert t[r, 0] == 0.334 * v  # F(0) = rho V
        b = v ** (2.0 / 3.0) / FOXS_DENOMINATOR
        assert abs(t[r, 7] - 0.334 * v * math.exp(-b * q[7] ** 2)) < 1e-14
    assert excluded_table("none", ff, used, q, 0.334) is None
    with pytest.raises(E.UsageError, match="voronota"):
        excluded_table("voronoi", ff, used, q, 0.334)
    assert np.all(excluded_table("grid", ff, used, q, 0.334) == 0.0)  # cells carry it
    with pytest.raises(ValueError, match="unknown excluded-volume model"):
        excluded_table("bogus", ff, used, q, 0.334)
    wf, w = water_table("weights", ff, q, np.ones(3), 3)
    assert len(wf) == 11 and w.tolist() == [1, 1, 1]
    with pytest.raises(E.UsageError):
        water_table("sasa", ff, q, None, 3)
    with pytest.raises(E.ValidationError):
        water_table("weights", ff, q, np.ones(2), 3)


def test_missing_displaced_volume(tmp_path):
    from pripps_amd.solvent import excluded_table
    y = tmp_path / "nv.yaml"
    y.write_text("C:\n  formfactor: !Constant 6.0\n")
    ff = read_formfactors(y)
    with pytest.raises(E.ValidationError, match="atom kind C"):
        excluded_table("fraser", ff, [0], np.array([0.0]), 0.334)


def test_error_mapping():
    assert isinstance(E.from_status(1, "x"), IOError)
    for code in (3, 4, 5, 6, 7, 8):
        assert isinstance(E.from_status(code, "x"), ValueError)
    assert isinstance(E.from_status(9, "x"), RuntimeError)
    assert isinstance(E.from_status(11, "x"), E.OutOfDeviceMemoryError)
    assert isinstance(E.from_status(12, "x"), TimeoutError)
    assert str(E.from_status(7, "bad")) == "bad"


def test_synthetic_globule():
    from scipy.spatial import cKDTree
    from pripps_amd.synth import globule, helix60k
    names, pos, w = globule(2000, 21.4, 7)
    assert len(names) == 2000 and pos.shape == (2000, 3)
    assert np.all(np.linalg.norm(pos, axis=1) <= 21.4 + 1e-12)
    d, _ = cKDTree(pos).query(pos, k=2)
    assert d[:, 1].min() >= 1.4
    counts = {e: names.count(e) for e in "CNOS"}
    assert counts == {"C": 1260, "N": 340, "O": 380, "S": 20}
    assert np.all((w >= 0) & (w <= 1)) and 0.1 < np.mean(w > 0) < 0.6
    r = np.linalg.norm(pos, axis=1)
    assert r[w > 0].mean() > r[w == 0].mean()
    names2, pos2, w2 = globule(2000, 21.4, 7)
    assert np.array_equal(pos, pos2) and names == names2  # seeded


def test_debye_model_requires_library_or_raises(tmp_path, monkeypatch):
    """No silent CPU fallback: a missing library raises ImportError."""
    from pripps_amd import _native
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError, match="no CPU fallback"):
        _native.load()


def test_read_weights_semantics(tmp_path):
    """Weights files as the reference's read_weights (src/trajectory.rs:117-135):
    one number per line, no normalisation, finite non-zero total."""
    p = tmp_path / "w.txt"
    p.write_text("1.0\n0\n2.5e-1\n")
    assert pripps_amd.read_weights(p).tolist() == [1.0, 0.0, 0.25]
    p.write_text("1\r\n-1e0\r\n3\r\n")  # CRLF, signed weights
    assert pripps_amd.read_weights(p).tolist() == [1.0, -1.0, 3.0]
    p.write_text("1.0\n 2.0\n")  # the reference's parser does not trim
    with pytest.raises(E.ValidationError, match="weight parse error"):
        pripps_amd.read_weights(p)
    p.write_text("1.0\n\n2.0\n")
    with pytest.raises(E.ValidationError, match="weight parse error"):
        pripps_amd.read_weights(p)
    p.write_text("1.0\n-1.0\n")
    with pytest.raises(E.ValidationError, match="zero or infinite"):
        pripps_amd.read_weights(p)
    p.write_text("inf\n1\n")
    with pytest.raises(E.ValidationError, match="zero or infinite"):
        pripps_amd.read_weights(p)
    with pytest.raises(E.FileError):
        pripps_amd.read_weights(tmp_path / "missing.txt")


def test_tuning_entry_and_stats_layout():
    """The ctypes mirrors of pripps_tuning_entry / pripps_build_stats have the
    C layout (sizes and offsets of an LP64 compile of the header)."""
    from pripps_amd import _native as N
    assert ctypes.sizeof(N.TuningEntry) == 8 * 4 + 8
    assert N.TuningEntry.pair_ms.offset == 32
    assert ctypes.sizeof(N.BuildStats) == 3 * 8 + 3 * 4 + 4 + 8 + 4 * 4
    assert N.BuildStats.work_items.offset == 40 and N.BuildStats.frames.offset == 48


def test_read_exp_data_lyzexp():
    """The reference's load_lyzexp (src/fit.rs:577-584) on its own data file."""
    exp = pripps_amd.read_exp_data(os.path.join(REFDATA, "lyzexp.dat"))
    assert len(exp.q) > 100
    assert exp.q[0] > 0 and exp.intensity[0] > 0 and exp.sigma[0] > 0


def test_read_exp_data_rules(tmp_path):
    """ExpData::from_file's line rules (src/fit.rs:28-84)."""
    p = tmp_path / "d.dat"
    p.write_text("# comment\nq I sigma\n\n0.1, 2.0, 0.5\n 0.2 3.0\n0.3\t4.0 abc\nabc 1 2\n"
                 "0.4 x\n1_0 2\n+0.5 .25 1e-1\n0.6,7\r\n0.7\n")
    exp = pripps_amd.read_exp_data(p)
    assert exp.q.tolist() == [0.1, 0.2, 0.3, 0.5, 0.6]
    assert exp.intensity.tolist() == [2.0, 3.0, 4.0, 0.25, 7.0]
    assert exp.sigma.tolist() == [0.5, 1.0, 1.0, 0.1, 1.0]
    p.write_text("# nothing\nq I\n")
    with pytest.raises(E.ValidationError, match="no data points"):
        pripps_amd.read_exp_data(p)
    with pytest.raises(E.FileError):
        pripps_amd.read_exp_data(tmp_path / "missing.dat")
    a = pripps_amd.ExpData.from_arrays([0.1, 0.2], [1.0, 2.0])
    assert a.sigma.tolist() == [1.0, 1.0]
    with pytest.raises(E.ValidationError, match="length mismatch"):
        pripps_amd.ExpData.from_arrays([0.1, 0.2], [1.0, 2.0], [1.0])
    with pytest.raises(E.ValidationError, match="empty"):
        pripps_amd.ExpData.from_arrays([], [])


def test_write_intensity_csv(tmp_path):
    """The CLI's CSV (src/cli.rs:510-551): header, then `{:.6}` fields."""
    q = np.array([0.0, 0.25, 0.5])
    full = pripps_amd.Intensity(q, np.array([1.0, -0.0, float("nan")]), np.array([2.0, 3.0, 4.0]),
                                np.array([0.0, 1e-7, 7.5e-7]), np.array([float("inf"), -1.5, 2.25]))
    path = tmp_path / "i.csv"
    pripps_amd.write_intensity_csv(path, full)
    assert path.read_text() == ("q,total,atoms,excluded,hydration\n"
                                "0.000000,1.000000,2.000000,0.000000,inf\n"
                                "0.250000,-0.000000,3.000000,0.000000,-1.500000\n"
                                "0.500000,NaN,4.000000,0.000001,2.250000\n")
    bare = pripps_amd.Intensity(q[:1], np.array([12.5]))
    pripps_amd.write_intensity_csv(path, bare)
    assert path.read_text() == "q,total\n0.000000,12.500000\n"


def test_linear_q_grid_is_the_oracle_grid_bitwise(oracle):
    """The product's q grid (Pripps.debye_model's qmin/qmax/nq) against the
    oracle's restatement of src/q_sampler.rs:24-29, bit for bit, and the
    reference's endpoint tests (src/q_sampler.rs:173-185)."""
    from pripps_amd.synth import linear_q
    for qmin, qmax, nq in ((0.0, 0.5, 500), (0.0, 0.3, 20), (0.0, 0.2, 8), (0.005, 1.0, 1000),
                           (0.013, 0.7391, 777), (0.42, 0.42, 1), (0.1, 0.2, 2), (0.0, 1.0, 0)):
        want = oracle.linear_q_values(qmin, qmax, nq)
        got = pripps_amd.linear_q_values(qmin, qmax, nq)
        assert got.dtype == np.float64 and got.tobytes() == want.tobytes(), (qmin, qmax, nq)
        assert linear_q(qmin, qmax, nq).tobytes() == want.tobytes()
    q = pripps_amd.linear_q_values(0.0, 1.0, 11)
    assert len(q) == 11 and q[0] == 0.0 and q[10] == 1.0
    assert pripps_amd.linear_q_values(0.42, 0.42, 1).tolist() == [0.42]
    with pytest.raises(E.ValidationError):
        pripps_amd.linear_q_values(0.0, 0.5, -1)


def _yaml_tagged(path):
    """The form-factor YAML read with a plain safe loader that keeps each
    tagged node as {"tag": ..., **mapping} -- independent of the package's
    own resolver."""
    import yaml

    class L(yaml.SafeLoader):
        pass

    def tagged(loader, suffix, node):
        return {"tag": suffix, **loader.construct_mapping(node, deep=True)}

    L.add_multi_constructor("!", tagged)
    with open(path) as fh:
        return yaml.load(fh, Loader=L)


def test_per_kind_excluded_table(oracle):
    """ExcludedVolume::PerKind (src/solvent/per_kind_excluded.rs:16-39): each
    kind's `excluded_solvent` entry of the same YAML, evaluated at q (the
    polynomial of src/formfactor.rs:379-385 through the oracle)."""
    from pripps_amd.solvent import excluded_table
    path = os.path.join(REFDATA, "poly_amino_acid.yaml")
    ff = read_formfactors(path)
    raw = _yaml_tagged(path)
    st = pripps_amd.structure_from_file(os.path.join(REFDATA, "4lzt.xyz"), ff.kinds())
    used = sorted(set(st.ids.tolist()))
    q = pripps_amd.linear_q_values(0.0, 0.5, 51)
    tab = excluded_table("per-kind", ff, used, q, 0.334)
    assert tab.shape == (len(used), len(q)) and len(used) == 20
    for r, kid in enumerate(used):
        spec = raw[ff.entries[kid].kind.name]["excluded_solvent"]
        assert spec["tag"] == "Polynomial"
        want = [oracle.eval_polynomial(spec["coeff"], x) for x in q]
        assert np.array_equal(tab[r], want), ff.entries[kid].kind.name


def test_per_kind_excluded_reference_cases(tmp_path):
    """src/solvent/per_kind_excluded.rs:58-90: a kind whose excluded_solvent is
    the constant 3.5 gives 3.5; a kind without one is a validation error
    naming `excluded_solvent`."""
    from pripps_amd.solvent import excluded_table
    (tmp_path / "a.yaml").write_text(
        "ALA: { formfactor: !Polynomial { qrange: [0.0, 1.0], coeff: [10.0] },\n"
        "       excluded_solvent: !Polynomial { qrange: [0.0, 1.0], coeff: [3.5] } }\n")
    (tmp_path / "b.yaml").write_text(
        "ALA: { formfactor: !Polynomial { qrange: [0.0, 1.0], coeff: [10.0] } }\n")
    ff = read_formfactors(str(tmp_path / "a.yaml"))
    assert excluded_table("per-kind", ff, [0], np.array([0.2]), 0.334).tolist() == [[3.5]]
    ff = read_formfactors(str(tmp_path / "b.yaml"))
    with pytest.raises(E.ValidationError, match="excluded_solvent"):
        excluded_table("per-kind", ff, [0], np.array([0.2]), 0.334)
