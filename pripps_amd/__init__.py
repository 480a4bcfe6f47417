"""pripps_amd -- an MI355X (gfx950) engine for the exact Debye sum.

Drop-in for the reference's Python ``Pripps.debye_model`` path
(``pripps-py/src/lib.rs:397-439``) and the cached model's ``intensity`` /
``fit`` / ``q_values`` (``pripps-py/src/lib.rs:597-659``)::

    import pripps_amd as pripps
    p = pripps.Pripps("assets/foxs_formfactors.yaml")
    model = p.debye_model("protein.xyz", qmin=0.0, qmax=0.5, nq=500,
                          excluded="foxs", hydration="weights",
                          hydration_weights=w)
    I = model.intensity(volume_scale=1.0, contrast_density=0.03)
    fit = model.fit(q_exp, I_exp, sigma_exp, method="grid")

The pair sums run in hand-written HIP kernels through the C ABI of
``include/pripps_debye.h``; there is no CPU fallback.  Extensions over the
reference API, all keyword-only: ``q_values`` (an explicit q grid, as the
reference's crate-private ``QSampler::new``), ``hydration="weights"`` with
``hydration_weights`` (explicit per-atom hydration weights in place of the
SASA tessellation), and ``precision`` ("fp32" compensated, "fp64", or
"fp64_plain").
"""
from __future__ import annotations

import enum
import os
import re
from dataclasses import dataclass
from typing import Optional, Sequence, Union

import numpy as np

from . import _native
from .engine import Engine, NativeModel, default_engine
from .errors import (DeviceError, FileError, InternalError, OutOfDeviceMemoryError,
                     PrippsError, QOutOfRangeError, UndefinedFormFactorError, UnknownAtomError,
                     UsageError, ValidationError, YamlError)
from .formats import ExpData, read_exp_data, write_fit_csv, write_intensity_csv
from .formfactor import FormFactorMap, read_formfactors, read_formfactors_merged
from .solvent import (atom_radii, excluded_name, excluded_table, grid_excluded_dummies,
                      grid_excluded_form_factor, grid_hydration_dummies, hydration_name,
                      molecules_per_cell, water_table)
from .structure import Structure, resolve_ids, structure_from_file

__version__ = "0.1.0"

DEFAULT_C1_BOUNDS = (0.5, 2.0)   # src/fit.rs:253
DEFAULT_C2_BOUNDS = (0.0, 0.1)   # src/fit.rs:259


class Excluded(enum.Enum):
    """Excluded-volume model (pripps-py/src/lib.rs:20-29)."""
    Disabled = "none"
    Fraser = "fraser"
    Foxs = "foxs"
    Grid = "grid"
    Voronoi = "voronoi"
    PerKind = "per-kind"


class Hydration(enum.Enum):
    """Hydration-shell model (pripps-py/src/lib.rs:33-40); ``Weights`` is the
    engine's explicit-weights extension."""
    Disabled = "none"
    Sasa = "sasa"
    Grid = "grid"
    Voronoi = "voronoi"
    Weights = "weights"


class FitMethod(enum.Enum):
    NelderMead = "nelder-mead"
    Grid = "grid"


@dataclass
class Intensity:
    """I(q) and the unscaled self terms (reference ``Intensity`` /
    ``Contributions``, src/lib.rs:89-121).  ``p`` and ``s_eff`` are the
    single-molecule form factor P(q) and the effective structure factor
    S_eff(q) of a Direct run with ``num_molecules`` (src/lib.rs:511-526;
    pripps-py ``Intensity.p`` / ``.s_eff``)."""
    q: np.ndarray
    total: np.ndarray
    atoms: Optional[np.ndarray] = None
    excluded: Optional[np.ndarray] = None
    hydration: Optional[np.ndarray] = None
    p: Optional[np.ndarray] = None
    s_eff: Optional[np.ndarray] = None

    def __len__(self):
        return len(self.q)


@dataclass
class FitResult:
    """Reference ``FitResult`` (src/fit.rs:238-247)."""
    volume_scale: float
    contrast_density: float
    chi2: float
    scale: float


def linear_q_values(qmin: float, qmax: float, nq: int) -> np.ndarray:
    """The linear q grid of ``QSampler::linear`` (src/q_sampler.rs:24-29):
    q_i = qmin + (qmax - qmin) * i / max(nq - 1, 1) in that operation order,
    in FP64 (not qmin + i * dq, which rounds differently)."""
    nq = int(nq)
    if nq < 0:
        raise ValidationError("nq must be >= 0")
    denom = float(max(nq - 1, 1))
    return np.array([qmin + (qmax - qmin) * i / denom for i in range(nq)], dtype=np.float64)


def _fit_method(method) -> int:
    if method is None:
        return _native.FIT_NELDER_MEAD
    name = getattr(method, "value", method)
    if not isinstance(name, str):
        raise TypeError("`method` must be a str or pripps_amd.FitMethod")
    name = name.lower()
    if name in ("nelder-mead", "neldermead", "nelder_mead", "nm"):
        return _native.FIT_NELDER_MEAD
    if name == "grid":
        return _native.FIT_GRID
    raise ValueError(f"unknown fit method: {name} (expected `nelder-mead` or `grid`)")


class DebyeModel:
    """A cached Debye model: six partial profiles resident on the GPU
    (reference ``DebyeModel``, src/debye.rs:139-174)."""

    def __init__(self, native: NativeModel, q: np.ndarray):
        self._native = native
        self._q = np.array(q, dtype=np.float64)

    @property
    def q_values(self) -> np.ndarray:
        return self._native.q_values()

    def intensity(self, volume_scale: float = 1.0, contrast_density: float = 0.03) -> Intensity:
        total, atoms, excluded, hydration = self._native.intensity(volume_scale, contrast_density)
        return Intensity(self._q.copy(), total, atoms, excluded, hydration)

    def profiles(self) -> np.ndarray:
        """The six cached profiles [6][nq] in DebyeTerms order."""
        return self._native.profiles()

    def fit(self, q, intensity, sigma=None, *, method=None, c1_points: int = 151,
            c2_points: int = 101, c1_bounds=None, c2_bounds=None) -> FitResult:
        r = self._native.fit(q, intensity, sigma, _fit_method(method), c1_points, c2_points,
                             c1_bounds, c2_bounds)
        return FitResult(r.volume_scale, r.contrast_density, r.chi2, r.scale)

    def fit_curve(self, q, volume_scale: float, contrast_density: float) -> np.ndarray:
        """The theory curve at experimental q as the reference's fit CSV
        evaluates it: natural spline of I(q; vs, cd) (src/fit.rs:548-560)."""
        return self._native.fit_curve(q, volume_scale, contrast_density)

    def stats(self):
        return self._native.stats()


class Scheme(enum.Enum):
    """Intensity scheme (pripps-py/src/lib.rs:45-50).  Multipole is the
    reference's default and is not part of this engine."""
    Multipole = "multipole"
    Debye = "debye"
    Direct = "direct"


def _scheme_name(model) -> str:
    if model is None:
        return "multipole"
    name = getattr(model, "value", model)
    if not isinstance(name, str):
        raise TypeError("`model` must be a str or pripps.Scheme")
    name = name.lower()
    if name not in ("multipole", "debye", "direct"):
        raise ValueError(f"unknown model: {name} (expected `multipole`, `debye`, or `direct`)")
    return name


class DirectModel:
    """A cached Direct model: per q-vector amplitudes of atoms, excluded volume
    and hydration resident on the GPU (reference ``DirectModel``,
    src/explicit.rs:328-366)."""

    def __init__(self, native, engine: Engine, n_vectors: int):
        self._native = native
        self._engine = engine
        self._n_vectors = n_vectors
        self._profile_model = None

    @property
    def q_values(self) -> np.ndarray:
        return self._native.q_values()

    def intensity(self, volume_scale: float = 1.0, contrast_density: float = 0.03) -> Intensity:
        total, atoms, excluded, hydration = self._native.intensity(volume_scale, contrast_density)
        return Intensity(self.q_values, total, atoms, excluded, hydration)

    def profiles(self) -> np.ndarray:
        """Group means [6][n_groups] of |A|^2, |E|^2, |H|^2, Re A E*, Re A H*,
        Re E H*: the Debye combine of these is ``intensity(vs, cd).total``."""
        return self._native.profiles()

    def amplitudes(self) -> np.ndarray:
        """Per q-vector [n_vectors][7]: Re A, Im A, Re E, Im E, Re H, Im H, self."""
        return self._native.amplitudes(self._n_vectors)

    def self_term(self) -> np.ndarray:
        return self._native.self_term()

    def _fit_model(self) -> NativeModel:
        if self._profile_model is None:
            self._profile_model = self._engine.from_profiles(self.q_values, self.profiles())
        return self._profile_model

    def fit(self, q, intensity, sigma=None, *, method=None, c1_points: int = 151,
            c2_points: int = 101, c1_bounds=None, c2_bounds=None) -> FitResult:
        """Bounded (volume_scale, contrast_density) fit, as the reference fits a
        DirectModel through its ``intensity`` closure (pripps-py/src/lib.rs:625-650)."""
        r = self._fit_model().fit(q, intensity, sigma, _fit_method(method), c1_points, c2_points,
                                  c1_bounds, c2_bounds)
        return FitResult(r.volume_scale, r.contrast_density, r.chi2, r.scale)

    def stats(self):
        return self._native.stats()


class Pripps:
    """Form-factor set plus the Debye and Direct entry points (reference
    ``Pripps``, src/lib.rs:226-528; Python class pripps-py/src/lib.rs)."""

    def __init__(self, atomfile: Union[str, Sequence[str]], *, engine: Optional[Engine] = None):
        if isinstance(atomfile, (list, tuple)):
            self.formfactors = read_formfactors_merged(atomfile)
        else:
            try:
                self.formfactors = read_formfactors(atomfile)
            except OSError as exc:
                raise FileError(f"{atomfile}: {exc.strerror or exc}") from exc
        self._engine = engine

    @property
    def engine(self) -> Engine:
        if self._engine is None:
            self._engine = default_engine()
        return self._engine

    def prepare_input(self, pos: np.ndarray, ids: np.ndarray, q: np.ndarray, *,
                      excluded=None, hydration=None, hydration_weights=None,
                      bulk_electron_density: float = 0.334, probe: float = 1.8,
                      spacing: float = 3.10516, thickness: float = 3.0,
                      excluded_spacing: float = 1.0) -> _native.PreparedInput:
        """Host tables for one structure: used kinds only (as
        formfactor_by_index evaluates only referenced kinds).  Grid solvent
        models append their dummies as points of extra kinds (see
        ``pripps_amd.solvent``): excluded-grid cells carry only the cell's
        excluded form factor, hydration-grid cells only a water weight."""
        ex = excluded_name(excluded)
        hy = hydration_name(hydration)
        ids = np.asarray(ids, dtype=np.int64)
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        if len(ids):
            used_arr, inverse = np.unique(ids, return_inverse=True)
            used = [int(k) for k in used_arr]
            kind = inverse.astype(np.int32).reshape(-1)
        else:
            used, kind = [0], np.zeros(0, dtype=np.int32)
        q = np.ascontiguousarray(q, dtype=np.float64)
        atom_ff = self.formfactors.table(used, q)
        excl = excluded_table(ex, self.formfactors, used, q, bulk_electron_density)
        water, w = water_table(hy, self.formfactors, q, hydration_weights, len(ids))
        if ex != "grid" and hy != "grid":
            return _native.PreparedInput(pos, kind, q, atom_ff, excl, water, w)
        points, kinds, weights = [pos], [kind], [w] if water is not None else None
        tables = [atom_ff]
        etables = [excl] if excl is not None else None
        nq = len(q)
        if ex == "grid":
            radii = atom_radii(self.formfactors, ids, "ExcludedVolume::Grid")
            E = grid_excluded_dummies(pos, radii, float(excluded_spacing))
            points.append(E)
            kinds.append(np.full(len(E), len(tables[0]) + len(tables) - 1, dtype=np.int32))
            tables.append(np.zeros((1, nq)))
            etables.append(grid_excluded_form_factor(q, bulk_electron_density,
                                                     float(excluded_spacing))[None, :])
            if weights is not None:
                weights.append(np.zeros(len(E)))
        if hy == "grid":
            radii = atom_radii(self.formfactors, ids, "Hydration::Grid")
            H = grid_hydration_dummies(pos, radii, float(spacing), float(thickness), float(probe))
            points.append(H)
            kinds.append(np.full(len(H), len(tables[0]) + len(tables) - 1, dtype=np.int32))
            tables.append(np.zeros((1, nq)))
            if etables is not None:
                etables.append(np.zeros((1, nq)))
            weights.append(np.full(len(H), molecules_per_cell(float(spacing))))
        return _native.PreparedInput(
            np.vstack(points), np.concatenate(kinds), q, np.vstack(tables),
            None if etables is None else np.vstack(etables), water,
            None if weights is None else np.concatenate(weights))

    def debye_model(self, structure_file, *, qmin: float = 0.0, qmax: float = 0.5, nq: int = 100,
                    excluded=None, hydration=None, probe: float = 1.8, spacing: float = 3.10516,
                    thickness: float = 3.0, excluded_spacing: float = 1.0,
                    bulk_electron_density: float = 0.334, q_values=None,
                    hydration_weights=None, precision: str = "fp32",
                    tuning: Optional[_native.Tuning] = None) -> DebyeModel:
        """Build a cached Debye model on the GPU (reference
        ``Pripps::debye_model``, src/lib.rs:287-303).  ``spacing``,
        ``thickness`` and ``probe`` parameterise ``hydration="grid"``
        (pripps-py/src/lib.rs:731-758), ``excluded_spacing`` parameterises
        ``excluded="grid"``."""
        st = structure_from_file(structure_file, self.formfactors.kinds())
        return self.debye_model_from_arrays(st.pos, st.ids, qmin=qmin, qmax=qmax, nq=nq,
                                            excluded=excluded, hydration=hydration,
                                            bulk_electron_density=bulk_electron_density,
                                            q_values=q_values,
                                            hydration_weights=hydration_weights,
                                            precision=precision, tuning=tuning, probe=probe,
                                            spacing=spacing, thickness=thickness,
                                            excluded_spacing=excluded_spacing)

    def debye_model_from_arrays(self, pos, ids, *, qmin=0.0, qmax=0.5, nq=100, excluded=None,
                                hydration=None, bulk_electron_density=0.334, q_values=None,
                                hydration_weights=None, precision="fp32",
                                tuning=None, probe=1.8, spacing=3.10516, thickness=3.0,
                                excluded_spacing=1.0) -> DebyeModel:
        if q_values is None:
            q = linear_q_values(qmin, qmax, nq)
        else:
            q = np.ascontiguousarray(q_values, dtype=np.float64)
        inp = self.prepare_input(pos, ids, q, excluded=excluded, hydration=hydration,
                                 hydration_weights=hydration_weights,
                                 bulk_electron_density=bulk_electron_density, probe=probe,
                                 spacing=spacing, thickness=thickness,
                                 excluded_spacing=excluded_spacing)
        native = self.engine.build(inp, precision, tuning)
        return DebyeModel(native, q)

    def trajectory_average(self, structure_file, frames, *, qmin: float = 0.0, qmax: float = 0.5,
                           nq: int = 100, excluded=None, hydration=None, weights=None,
                           stride: int = 1, units: str = "angstrom",
                           bulk_electron_density: float = 0.334, q_values=None,
                           hydration_weights=None, precision: str = "fp32",
                           tuning: Optional[_native.Tuning] = None,
                           max_frames_per_launch: int = 0) -> DebyeModel:
        """Weighted average of the Debye profiles over trajectory frames: the
        reference's ``trajectory_average`` (src/trajectory.rs:167-251) for the
        Debye scheme, with ``XtcIterator``'s stride and per-frame weights
        (src/trajectory.rs:29-135).  ``structure_file`` is the template (atom
        names and count); ``frames`` is an array [n_frames][n_atoms][3] in
        ``units`` ("angstrom" or "nm", the XTC unit the reference converts
        with a factor 10).  ``weights`` is an array with one weight per frame
        or the path of a weights file (one number per line), indexed by the
        original frame number.  The returned model's ``intensity(vs, cd)`` is
        the weighted average of the per-frame intensities."""
        st = structure_from_file(structure_file, self.formfactors.kinds())
        return self.trajectory_average_from_arrays(
            st.ids, frames, qmin=qmin, qmax=qmax, nq=nq, excluded=excluded, hydration=hydration,
            weights=weights, stride=stride, units=units,
            bulk_electron_density=bulk_electron_density, q_values=q_values,
            hydration_weights=hydration_weights, precision=precision, tuning=tuning,
            max_frames_per_launch=max_frames_per_launch)

    def trajectory_average_from_arrays(self, ids, frames, *, qmin=0.0, qmax=0.5, nq=100,
                                       excluded=None, hydration=None, weights=None, stride=1,
                                       units="angstrom", bulk_electron_density=0.334,
                                       q_values=None, hydration_weights=None, precision="fp32",
                                       tuning=None, max_frames_per_launch=0) -> DebyeModel:
        if units not in ("angstrom", "nm"):
            raise ValueError(f"unknown units: {units} (expected `angstrom` or `nm`)")
        ids = np.asarray(ids, dtype=np.int64)
        f = np.asarray(frames, dtype=np.float64)
        if f.ndim != 3 or f.shape[1:] != (len(ids), 3):
            raise ValidationError(f"number of atoms in structure ({len(ids)}) does not match "
                                  f"trajectory frames of shape {list(f.shape)}")
        if units == "nm":
            f = f * 10.0
        if isinstance(weights, (str, os.PathLike)):
            weights = read_weights(weights)
        if q_values is None:
            q = linear_q_values(qmin, qmax, nq)
        else:
            q = np.ascontiguousarray(q_values, dtype=np.float64)
        pos0 = f[0] if len(f) else np.zeros((len(ids), 3))
        inp = self.prepare_input(pos0, ids, q, excluded=excluded, hydration=hydration,
                                 hydration_weights=hydration_weights,
                                 bulk_electron_density=bulk_electron_density)
        native = self.engine.frames_average(inp, f, weights=weights, stride=stride,
                                            precision=precision, tuning=tuning,
                                            max_frames_per_launch=max_frames_per_launch)
        return DebyeModel(native, q)

    # -- Direct Fourier transform -------------------------------------------
    def direct_model(self, structure_file, *, pmax: int = 6, side_length: float = 200.0,
                     excluded=None, hydration=None, probe: float = 1.8,
                     spacing: float = 3.10516, thickness: float = 3.0,
                     excluded_spacing: float = 1.0, bulk_electron_density: float = 0.334,
                     box_sides=None, grid=None, hydration_weights=None,
                     precision: str = "fp32") -> DirectModel:
        """Cached Direct model in a periodic box (reference
        ``Pripps::direct_model``, src/lib.rs:305-331; pripps-py/src/lib.rs:441-483):
        Miller q-vectors p = 1..pmax in a cubic box of ``side_length`` A.
        Extensions: ``box_sides`` (an orthorhombic box), ``grid`` (a
        ``pripps_amd.direct.QGrid``, e.g. a lattice grid), ``hydration_weights``,
        ``precision`` ("fp32" or "fp64")."""
        st = structure_from_file(structure_file, self.formfactors.kinds())
        return self.direct_model_from_arrays(
            st.pos, st.ids, pmax=pmax, box_sides=side_length if box_sides is None else box_sides,
            excluded=excluded, hydration=hydration, probe=probe, spacing=spacing,
            thickness=thickness, excluded_spacing=excluded_spacing,
            bulk_electron_density=bulk_electron_density, grid=grid,
            hydration_weights=hydration_weights, precision=precision)

    def direct_model_from_arrays(self, pos, ids, *, pmax: int = 6, box_sides=200.0,
                                 excluded=None, hydration=None, probe=1.8, spacing=3.10516,
                                 thickness=3.0, excluded_spacing=1.0, bulk_electron_density=0.334,
                                 grid=None, hydration_weights=None, num_molecules: int = 0,
                                 precision="fp32") -> DirectModel:
        from .direct import QGrid, build_native, check_box, parse_box
        box = parse_box(box_sides)
        check_box(box)
        grid = QGrid.miller(pmax) if grid is None else grid
        idx, qn, group, group_q = grid.groups(box)
        inp = self.prepare_input(pos, ids, qn, excluded=excluded, hydration=hydration,
                                 hydration_weights=hydration_weights,
                                 bulk_electron_density=bulk_electron_density, probe=probe,
                                 spacing=spacing, thickness=thickness,
                                 excluded_spacing=excluded_spacing)
        native = build_native(self.engine, inp, box, idx, group, group_q,
                              num_molecules=num_molecules, precision=precision)
        return DirectModel(native, self.engine, len(idx))

    def direct_trajectory_average(self, structure_file, frames, *, box_sides, pmax: int = 6,
                                  grid=None, num_molecules=None, weights=None, stride: int = 1,
                                  units: str = "angstrom", excluded=None, hydration=None,
                                  probe: float = 1.8, spacing: float = 3.10516,
                                  thickness: float = 3.0, excluded_spacing: float = 1.0,
                                  bulk_electron_density: float = 0.334,
                                  volume_scale: float = 1.0, contrast_density: float = 0.03,
                                  hydration_weights=None, precision: str = "fp32") -> Intensity:
        """Weighted average of the Direct I(q) over trajectory frames: the
        reference's ``trajectory_average`` (src/trajectory.rs:167-251) with the
        Direct scheme, a fixed box and ``XtcIterator``'s stride / weights
        (src/trajectory.rs:29-135).  ``frames`` is [n_frames][n_atoms][3] in
        ``units`` ("angstrom" or "nm", the XTC factor 10); ``weights`` an array
        or a weights file indexed by the original frame.  With
        ``num_molecules`` the result carries P(q) and S_eff(q) formed from the
        averaged self term (src/lib.rs:511-526)."""
        from .direct import QGrid, check_box, frames_average, parse_box
        if units not in ("angstrom", "nm"):
            raise ValueError(f"unknown units: {units} (expected `angstrom` or `nm`)")
        if int(stride) < 1:
            raise ValueError("`stride` must be >= 1")
        st = structure_from_file(structure_file, self.formfactors.kinds())
        ids = st.ids
        f = np.asarray(frames, dtype=np.float64)
        if f.ndim != 3 or f.shape[1:] != (len(ids), 3):
            raise ValidationError(f"number of atoms in structure ({len(ids)}) does not match "
                                  f"trajectory frames of shape {list(f.shape)}")
        if units == "nm":
            f = f * 10.0
        if isinstance(weights, (str, os.PathLike)):
            weights = read_weights(weights)
        ex, hy = excluded_name(excluded), hydration_name(hydration)
        if num_molecules is not None:
            n = int(num_molecules)
            if n == 0:
                raise UsageError("--num-molecules must be at least 1")
            if ex != "none" or hy != "none":
                raise UsageError("--num-molecules supports vacuum runs only (no "
                                 "excluded-volume/hydration model)")
            if len(ids) % n != 0:
                raise ValidationError(f"box atom count ({len(ids)}) is not a multiple of "
                                      f"--num-molecules ({n})")
        box = parse_box(box_sides)
        check_box(box)
        grid = QGrid.miller(pmax) if grid is None else grid
        idx, qn, group, gq = grid.groups(box)
        pos0 = f[0] if len(f) else np.zeros((len(ids), 3))
        inp = self.prepare_input(pos0, ids, qn, excluded=excluded, hydration=hydration,
                                 hydration_weights=hydration_weights,
                                 bulk_electron_density=bulk_electron_density, probe=probe,
                                 spacing=spacing, thickness=thickness,
                                 excluded_spacing=excluded_spacing)
        if len(inp.x) != len(ids):
            raise UsageError("grid solvent models add dummies per frame; trajectories take "
                             "colocated solvent models only")
        nmol = 0 if num_molecules is None else int(num_molecules)
        vs, cd = (volume_scale, contrast_density) if (ex != "none" or hy != "none") else (1.0, 0.0)
        total, self_term = frames_average(self.engine, inp, box, idx, group, gq, f,
                                          weights=weights, stride=stride, num_molecules=nmol,
                                          precision=precision, volume_scale=vs,
                                          contrast_density=cd)
        res = Intensity(gq.copy(), total)
        if self_term is not None:
            res.p = self_term / float(nmol)
            with np.errstate(divide="ignore", invalid="ignore"):
                res.s_eff = np.where(self_term > 0.0,
                                     total / np.where(self_term > 0, self_term, 1.0), np.nan)
        return res

    def intensity(self, structure_file, *, model=None, xtc=None, stride: int = 1, weights=None,
                  box_sides=None, excluded=None, hydration=None, qmin: float = 0.0,
                  qmax: float = 0.5, nq=None, l_max=None, pmax: int = 6, probe: float = 1.8,
                  spacing: float = 3.10516, thickness: float = 3.0,
                  excluded_spacing: float = 1.0, bulk_electron_density: float = 0.334,
                  volume_scale: float = 1.0, contrast_density: float = 0.03,
                  num_molecules=None, hydration_weights=None, grid=None,
                  precision: str = "fp32") -> Intensity:
        """One-shot I(q) (reference ``Pripps::intensity``, src/lib.rs:333-528;
        pripps-py/src/lib.rs:485-595) for ``model="debye"`` or ``"direct"``.
        The Multipole scheme and XTC input are not part of this engine."""
        from .direct import parse_box
        scheme = _scheme_name(model)
        if scheme == "multipole":
            raise UsageError("the multipole scheme is not part of this engine; pass "
                             "model='debye' or model='direct'")
        if xtc is not None:
            raise UsageError("XTC trajectories are not read by this engine; pass decoded frames "
                             "to Pripps.trajectory_average")
        if int(stride) < 1:
            raise ValueError("`stride` must be >= 1")
        st = structure_from_file(structure_file, self.formfactors.kinds())
        ex, hy = excluded_name(excluded), hydration_name(hydration)
        solvent_active = ex != "none" or hy != "none"
        if num_molecules is not None:
            n = int(num_molecules)
            if scheme != "direct":
                raise UsageError("--num-molecules requires the `direct` scheme (the periodic-box "
                                 "method)")
            if n == 0:
                raise UsageError("--num-molecules must be at least 1")
            if solvent_active:
                raise UsageError("--num-molecules supports vacuum runs only (no "
                                 "excluded-volume/hydration model)")
            if len(st.ids) % n != 0:
                raise ValidationError(f"box atom count ({len(st.ids)}) is not a multiple of "
                                      f"--num-molecules ({n})")
            m = len(st.ids) // n
            if m > 0 and any(not np.array_equal(st.ids[b * m:(b + 1) * m], st.ids[:m])
                             for b in range(n)):
                raise ValidationError(f"--num-molecules {n} does not partition the box into {n} "
                                      "identical contiguous molecules: the atom composition "
                                      "differs between blocks")
        if scheme == "debye":
            dm = self.debye_model_from_arrays(
                st.pos, st.ids, qmin=qmin, qmax=qmax, nq=100 if nq is None else nq,
                excluded=excluded, hydration=hydration,
                bulk_electron_density=bulk_electron_density,
                hydration_weights=hydration_weights, precision=precision, probe=probe,
                spacing=spacing, thickness=thickness, excluded_spacing=excluded_spacing)
            res = dm.intensity(volume_scale, contrast_density)
            if not solvent_active:
                res.atoms = res.excluded = res.hydration = None
            return res
        if box_sides is None:
            raise UsageError("direct Fourier transform requires box side lengths on the "
                             "structure")
        box = parse_box(box_sides)
        dmod = self.direct_model_from_arrays(
            st.pos, st.ids, pmax=pmax, box_sides=box, excluded=excluded, hydration=hydration,
            probe=probe, spacing=spacing, thickness=thickness, excluded_spacing=excluded_spacing,
            bulk_electron_density=bulk_electron_density, grid=grid,
            hydration_weights=hydration_weights,
            num_molecules=0 if num_molecules is None else int(num_molecules), precision=precision)
        if solvent_active:
            return dmod.intensity(volume_scale, contrast_density)
        total = dmod.intensity(1.0, 0.0).total
        res = Intensity(dmod.q_values, total)
        if num_molecules is not None:
            self_term = dmod.self_term()
            res.p = self_term / float(int(num_molecules))
            with np.errstate(divide="ignore", invalid="ignore"):
                res.s_eff = np.where(self_term > 0.0, total / np.where(self_term > 0, self_term, 1.0),
                                     np.nan)
        return res

    def names_to_ids(self, names: Sequence[str]) -> np.ndarray:
        return resolve_ids(names, self.formfactors.kinds())

    def prepare_workload(self, wl, bulk_electron_density: float = 0.334) -> _native.PreparedInput:
        """Host tables of a synthetic workload (``pripps_amd.synth.Workload``)."""
        ids = self.names_to_ids(wl.names)
        hydration = "weights" if wl.weights is not None else None
        return self.prepare_input(wl.pos, ids, wl.q, excluded=wl.excluded, hydration=hydration,
                                  hydration_weights=wl.weights,
                                  bulk_electron_density=bulk_electron_density)


_RUST_F64 = re.compile(r"[+-]?(inf|infinity|nan|(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)",
                       re.IGNORECASE)


def read_weights(path) -> np.ndarray:
    """Per-frame weights, one number per line (reference ``read_weights``,
    src/trajectory.rs:117-135): the whole line must be a number; the weights
    are returned as they are (not normalised) and must have a finite non-zero
    sum."""
    try:
        with open(path, "r", encoding="utf-8") as fh:
            text = fh.read()
    except OSError as exc:
        raise FileError(f"{path}: {exc.strerror or exc}") from exc
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    out = []
    for line in lines:
        if line.endswith("\r"):
            line = line[:-1]
        if not _RUST_F64.fullmatch(line):
            raise ValidationError(f"weight parse error: invalid float literal ({line!r})")
        out.append(float(line))
    w = np.array(out, dtype=np.float64)
    total = float(np.sum(w)) if len(w) else 0.0
    if not (np.isfinite(total) and abs(total) >= np.finfo(np.float64).tiny):
        raise ValidationError("sum of weights is zero or infinite")
    return w


__all__ = [
    "Pripps", "DebyeModel", "Intensity", "FitResult", "Excluded", "Hydration", "FitMethod",
    "Engine", "read_weights", "linear_q_values", "ExpData", "read_exp_data", "write_intensity_csv", "write_fit_csv",
    "DEFAULT_C1_BOUNDS", "DEFAULT_C2_BOUNDS", "PrippsError", "FileError", "YamlError",
    "UnknownAtomError", "UndefinedFormFactorError", "QOutOfRangeError", "ValidationError",
    "UsageError", "InternalError", "DeviceError", "OutOfDeviceMemoryError",
]
