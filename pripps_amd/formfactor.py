"""Form-factor tables for the Debye engine (host side).

Restates the reference's form-factor layer so that a Python caller can build
the per-(kind, q) tables the engine consumes:

* YAML schema and symbolic resolution -- ``src/formfactor.rs:310-321``
  (``AtomSpec``), ``446-501`` (``read_formfactors(_merged)``), resolution order
  alias -> linear -> united (``560-564``), ``convert_linear`` (``566-616``),
  ``convert_alias`` (``618-640``), ``convert_united`` (``642-698``).
* Runtime evaluation -- ``FormFactor::eval`` (``188-201``), ``eval_gaussian``
  with the Cromer-Mann s^2 = q^2/(16 pi^2) convention (``369-375``),
  ``eval_polynomium`` (``379-385``), ``interpolate`` (``389-408``).
* ``AtomKind`` derivation: sigma = 2 r, displaced volume explicit or
  (4/3) pi r^3 (``src/formfactor.rs:350-362``, ``src/atomkind.rs:25-36``).
* ``formfactor_by_index`` semantics (``528-544``): only kinds referenced by the
  structure are evaluated, so unused kinds with narrow q ranges never raise.

Species are kept in name order, the order of the reference's ``BTreeMap``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import yaml

from .errors import QOutOfRangeError, UndefinedFormFactorError, ValidationError, YamlError


# ---------------------------------------------------------------------------
# runtime form factors
# ---------------------------------------------------------------------------
class FormFactor:
    """Runtime F(q).  Subclasses mirror the reference's ``FormFactor`` enum."""

    def qrange(self) -> Tuple[float, float]:
        return (0.0, math.inf)

    def _value(self, q: float) -> float:
        raise NotImplementedError

    def eval(self, q: float) -> float:
        lo, hi = self.qrange()
        if not (lo <= q <= hi):
            raise QOutOfRangeError(q, lo, hi)
        return self._value(q)


@dataclass
class Constant(FormFactor):
    c: float

    def _value(self, q):
        return self.c


@dataclass
class Polynomial(FormFactor):
    qlim: Tuple[float, float]
    coeff: List[float]

    def qrange(self):
        return self.qlim

    def _value(self, q):
        total = 0.0
        for i, c in enumerate(self.coeff):
            total += c * _powi(q, i)
        return total


@dataclass
class Table(FormFactor):
    points: List[Tuple[float, float]]

    def qrange(self):
        return (self.points[0][0], self.points[-1][0])

    def _value(self, q):
        return _interpolate(self.points, q)


@dataclass
class Gaussian(FormFactor):
    qlim: Tuple[float, float]
    ab: List[Tuple[float, float]]
    c: float

    def qrange(self):
        return self.qlim

    def _value(self, q):
        s2 = q * q / (16.0 * math.pi * math.pi)
        total = 0.0
        for a, b in self.ab:
            total += a * math.exp(-b * s2)
        return total + self.c


@dataclass
class SingleGaussian(FormFactor):
    """F(q) = a exp(-b q^2), the q^2 convention used by Fraser dummies."""
    a: float
    b: float

    def _value(self, q):
        return self.a * math.exp(-self.b * q * q)


def _powi(x: float, n: int) -> float:
    # binary exponentiation, as Rust's f64::powi lowers to
    r = 1.0
    while True:
        if n & 1:
            r *= x
        n //= 2
        if n == 0:
            break
        x *= x
    return r


def _interpolate(table: Sequence[Tuple[float, float]], x: float) -> float:
    low, high = 0, len(table) - 1
    while low < high:
        mid = (low + high) // 2
        xm = table[mid][0]
        if x < xm:
            high = mid
        elif x > xm:
            low = mid + 1
        else:
            return table[mid][1]
    x0, y0 = table[low - 1]
    x1, y1 = table[low]
    return y0 + (x - x0) * (y1 - y0) / (x1 - x0)


def _sinc(x: float) -> float:
    return 1.0 if abs(x) < 1e-8 else math.sin(x) / x


# ---------------------------------------------------------------------------
# atom kinds and the resolved map
# ---------------------------------------------------------------------------
@dataclass
class AtomKind:
    name: str
    sigma: Optional[float] = None
    displaced_volume: Optional[float] = None

    @property
    def radius(self) -> Optional[float]:
        return None if self.sigma is None else 0.5 * self.sigma


@dataclass
class FormFactorEntry:
    kind: AtomKind
    formfactor: FormFactor
    excluded_solvent: Optional[FormFactor] = None


@dataclass
class FormFactorMap:
    entries: List[FormFactorEntry] = field(default_factory=list)

    def __len__(self):
        return len(self.entries)

    def kinds(self) -> List[AtomKind]:
        return [e.kind for e in self.entries]

    def find_name(self, name: str) -> Optional[int]:
        for i, e in enumerate(self.entries):
            if e.kind.name == name:
                return i
        return None

    def form_factor_by_name(self, name: str) -> Optional[FormFactor]:
        i = self.find_name(name)
        return None if i is None else self.entries[i].formfactor

    def table(self, ids: Sequence[int], q: np.ndarray) -> np.ndarray:
        """F_k(q_m) for each listed kind id, shape [len(ids)][nq]."""
        out = np.empty((len(ids), len(q)), dtype=np.float64)
        for r, kid in enumerate(ids):
            ff = self.entries[kid].formfactor
            for m, qq in enumerate(q):
                out[r, m] = ff.eval(float(qq))
        return out


# ---------------------------------------------------------------------------
# YAML reading
# ---------------------------------------------------------------------------
class _Loader(yaml.SafeLoader):
    pass


# Names such as N/ON/NO must stay strings (the reference's YAML reader has no
# YAML-1.1 booleans), so the bool resolver is dropped.
_Loader.yaml_implicit_resolvers = {
    k: [(tag, rx) for tag, rx in v if tag != "tag:yaml.org,2002:bool"]
    for k, v in yaml.SafeLoader.yaml_implicit_resolvers.items()
}


def _tagged(tag):
    def construct(loader, node):
        if isinstance(node, yaml.MappingNode):
            value = loader.construct_mapping(node, deep=True)
        elif isinstance(node, yaml.SequenceNode):
            value = loader.construct_sequence(node, deep=True)
        else:
            value = loader.construct_scalar(node)
        return (tag, value)
    return construct


for _t in ("Constant", "Polynomial", "Table", "Gaussian", "Alias", "United", "Linear"):
    _Loader.add_constructor("!" + _t, _tagged(_t))


def _f(x) -> float:
    try:
        return float(x)
    except (TypeError, ValueError) as exc:
        raise YamlError(f"expected a number, got {x!r}") from exc


def _parse_spec(value):
    """Tagged YAML value -> ('Tag', payload) with numbers converted."""
    if not (isinstance(value, tuple) and len(value) == 2):
        raise YamlError(f"form factor must be a tagged value (!Gaussian, ...), got {value!r}")
    tag, v = value
    if tag == "Constant":
        return ("Constant", _f(v))
    if tag == "Polynomial":
        return ("Polynomial", (tuple(_f(x) for x in v["qrange"]), [_f(c) for c in v["coeff"]]))
    if tag == "Table":
        return ("Table", [(_f(a), _f(b)) for a, b in v])
    if tag == "Gaussian":
        return ("Gaussian", (tuple(_f(x) for x in v["qrange"]),
                             [(_f(a), _f(b)) for a, b in v["ab"]], _f(v["c"])))
    if tag == "Alias":
        return ("Alias", str(v))
    if tag == "United":
        orbit = []
        for o in v["orbit"]:
            w = _f(o[2]) if len(o) > 2 and o[2] is not None else None
            orbit.append((str(o[0]), _f(o[1]), w))
        return ("United", (str(v["center"]), orbit))
    if tag == "Linear":
        return ("Linear", [(str(n), _f(w)) for n, w in v])
    raise YamlError(f"unknown form-factor tag !{tag}")


def _parse_file(path) -> Dict[str, dict]:
    try:
        with open(path, "r", encoding="utf-8") as fh:
            doc = yaml.load(fh, Loader=_Loader)  # noqa: S506 -- SafeLoader subclass
    except OSError:
        raise
    except yaml.YAMLError as exc:
        raise YamlError(f"YAML error: {exc}") from exc
    if not isinstance(doc, dict):
        raise YamlError(f"YAML error: {path}: top level must be a mapping")
    specs = {}
    for name, spec in doc.items():
        if not isinstance(spec, dict):
            raise YamlError(f"YAML error: entry {name!r} must be a mapping")
        ff = spec.get("formfactor", spec.get("ff"))
        if ff is None:
            raise YamlError(f"YAML error: entry {name!r}: missing field `formfactor`")
        specs[str(name)] = {
            "radius": None if spec.get("radius") is None else _f(spec["radius"]),
            "displaced_volume": None if spec.get("displaced_volume") is None
            else _f(spec["displaced_volume"]),
            "ff": _parse_spec(ff),
            "excluded_solvent": None if spec.get("excluded_solvent") is None
            else _parse_spec(spec["excluded_solvent"]),
        }
    return specs


def _spec_qrange(spec):
    tag, v = spec
    if tag == "Gaussian":
        return v[0]
    if tag == "Polynomial":
        return v[0]
    if tag == "Table":
        return (v[0][0], v[-1][0])
    return (0.0, math.inf)


def _to_runtime(spec) -> FormFactor:
    tag, v = spec
    if tag == "Constant":
        return Constant(v)
    if tag == "Polynomial":
        return Polynomial(tuple(v[0]), list(v[1]))
    if tag == "Table":
        return Table(list(v))
    if tag == "Gaussian":
        return Gaussian(tuple(v[0]), list(v[1]), v[2])
    raise UndefinedFormFactorError(f"unresolved {tag} form factor")


def _resolve(specs: Dict[str, dict], role: str) -> None:
    # alias
    plain = {n: s[role] for n, s in specs.items() if s[role] is not None and s[role][0] != "Alias"}
    for n, s in specs.items():
        sp = s[role]
        if sp is not None and sp[0] == "Alias":
            if sp[1] not in plain:
                raise UndefinedFormFactorError(sp[1])
            s[role] = plain[sp[1]]
    # linear: 100-point table over the intersection of component ranges (<= 2)
    plain = {n: s[role] for n, s in specs.items() if s[role] is not None and s[role][0] != "Linear"}
    for n, s in specs.items():
        sp = s[role]
        if sp is None or sp[0] != "Linear":
            continue
        comps = sp[1]
        qmin, qmax = 0.0, 2.0
        for cname, _w in comps:
            if cname not in plain:
                raise UndefinedFormFactorError(cname)
            lo, hi = _spec_qrange(plain[cname])
            qmin = max(lo, qmin)
            qmax = min(hi, qmax)
        table = []
        denom = float(max(100 - 1, 1))
        for i in range(100):
            q = qmin + (qmax - qmin) * i / denom
            f = 0.0
            for cname, w in comps:
                f += w * _to_runtime(plain[cname]).eval(q)
            table.append((q, f))
        s[role] = ("Table", table)
    # united: 101-point table over the centre's range, qmax capped at 2 if infinite
    plain = {n: s[role] for n, s in specs.items() if s[role] is not None and s[role][0] != "United"}
    resolved = {}
    for n, s in specs.items():
        sp = s[role]
        if sp is None or sp[0] != "United":
            continue
        center, orbit = sp[1]
        if center not in plain:
            raise UndefinedFormFactorError(center)
        cff = _to_runtime(plain[center])
        qmin, qmax = _spec_qrange(plain[center])
        if math.isinf(qmax):
            qmax = 2.0
        table = []
        for i in range(101):
            q = qmin + (qmax - qmin) * i / 100.0
            table.append([q, cff.eval(q)])
        for oname, offset, weight in orbit:
            if oname not in plain:
                raise UndefinedFormFactorError(oname)
            off = _to_runtime(plain[oname])
            w = 1.0 if weight is None else weight
            for row in table:
                row[1] += w * off.eval(row[0]) * _sinc(row[0] * offset)
        resolved[n] = ("Table", [(a, b) for a, b in table])
    for n, t in resolved.items():
        specs[n][role] = t


def _finalize(specs: Dict[str, dict]) -> FormFactorMap:
    _resolve(specs, "ff")
    _resolve(specs, "excluded_solvent")
    entries = []
    for name in sorted(specs):
        s = specs[name]
        r = s["radius"]
        dv = s["displaced_volume"]
        if dv is None and r is not None:
            dv = (4.0 / 3.0) * math.pi * _powi(r, 3)
        kind = AtomKind(name, None if r is None else 2.0 * r, dv)
        ex = None if s["excluded_solvent"] is None else _to_runtime(s["excluded_solvent"])
        entries.append(FormFactorEntry(kind, _to_runtime(s["ff"]), ex))
    return FormFactorMap(entries)


def read_formfactors(path) -> FormFactorMap:
    """``read_formfactors`` (src/formfactor.rs:446-455)."""
    return _finalize(_parse_file(path))


def read_formfactors_merged(paths: Sequence) -> FormFactorMap:
    """Merge several files before resolution, later files win
    (src/formfactor.rs:469-501)."""
    paths = list(paths)
    if not paths:
        raise ValidationError("no form-factor files supplied")
    merged = _parse_file(paths[0])
    for p in paths[1:]:
        merged.update(_parse_file(p))
    return _finalize(merged)


def formfactor_by_index(ffmap: FormFactorMap, q: float, atom_ids: Sequence[int]) -> np.ndarray:
    """N-vector of atom form factors at q, evaluating each used kind once
    (src/formfactor.rs:528-544)."""
    cache: Dict[int, float] = {}
    out = np.empty(len(atom_ids))
    for i, kid in enumerate(atom_ids):
        if kid not in cache:
            cache[kid] = ffmap.entries[kid].formfactor.eval(q)
        out[i] = cache[kid]
    return out
