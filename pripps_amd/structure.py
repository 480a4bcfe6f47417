"""Structure loading for the Debye engine (host side).

Restates the reference's XYZ path and name -> kind resolution:

* ``read_structure`` (``src/input.rs:32-53``): extension decides the format;
  an empty result is a ``Validation`` error "no atoms found in <path>".
* ``parse_xyz`` (``src/input.rs:143-168``): skip blank lines and lines
  starting with ``;`` or ``#``; keep lines that tokenise as
  ``name x y z`` (extra tokens ignored); skip everything else (the count and
  comment lines of strict XYZ).
* ``resolve_atom_kind_id`` (``src/structure.rs:139-163``) for XYZ input:
  exact-name lookup, ``UnknownAtom`` otherwise.
* ``Structure::from_file`` (``src/structure.rs:172-193``); Debye does not
  centre the structure (``src/lib.rs:287-303``).

PDB input (residue-scoped and united-atom mapping, ``src/input.rs:88-137``,
``src/structure.rs:41-137``) is outside this engine's scope; such files raise
``UsageError``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import List, Sequence

import numpy as np

from .errors import FileError, UnknownAtomError, UsageError, ValidationError
from .formfactor import AtomKind


@dataclass
class Structure:
    pos: np.ndarray   # [N, 3] float64, Angstrom
    ids: np.ndarray   # [N] int64 kind ids into the form-factor map


def parse_xyz_text(text: str):
    names: List[str] = []
    coords: List[List[float]] = []
    for line in text.splitlines():
        t = line.strip()
        if not t or t.startswith(";") or t.startswith("#"):
            continue
        parts = t.split()
        if len(parts) < 4:
            continue
        try:
            x, y, z = float(parts[1]), float(parts[2]), float(parts[3])
        except ValueError:
            continue
        names.append(parts[0])
        coords.append([x, y, z])
    return names, np.asarray(coords, dtype=np.float64).reshape(-1, 3)


def read_structure(path):
    ext = os.path.splitext(str(path))[1].lower().lstrip(".")
    try:
        with open(path, "r", encoding="utf-8", errors="replace") as fh:
            text = fh.read()
    except OSError as exc:
        raise FileError(f"{path}: {exc.strerror or exc}") from exc
    if ext in ("pdb", "ent", "pdb1"):
        raise UsageError(f"{path}: PDB input is not supported by the Debye engine; "
                         "convert to XYZ with bare element names")
    names, pos = parse_xyz_text(text)
    if not names:
        raise ValidationError(f"no atoms found in {path}")
    return names, pos


def resolve_ids(names: Sequence[str], kinds: Sequence[AtomKind]) -> np.ndarray:
    index = {}
    for i, k in enumerate(kinds):
        index.setdefault(k.name, i)
    ids = np.empty(len(names), dtype=np.int64)
    for a, name in enumerate(names):
        kid = index.get(name)
        if kid is None:
            raise UnknownAtomError(f"{name}; add an entry for this atom type to the form-factor YAML")
        ids[a] = kid
    return ids


def structure_from_file(path, kinds: Sequence[AtomKind]) -> Structure:
    names, pos = read_structure(path)
    return Structure(pos=pos, ids=resolve_ids(names, kinds))


def write_xyz(path, names: Sequence[str], pos: np.ndarray, comment: str = "") -> None:
    """Strict XYZ (count, comment, ``name x y z`` lines) at full f64 precision."""
    with open(path, "w", encoding="utf-8") as fh:
        fh.write(f"{len(names)}\n{comment}\n")
        for n, (x, y, z) in zip(names, pos):
            fh.write(f"{n} {float(x)!r} {float(y)!r} {float(z)!r}\n")
