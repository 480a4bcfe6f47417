"""Wire formats around the Debye path, as the reference reads and writes them.

* ``read_exp_data``: ``ExpData::from_file`` (src/fit.rs:28-84): CSV or
  whitespace columns q, I [, sigma]; comment, blank and header lines skipped.
* ``write_intensity_csv``: the CLI's ``write_csv`` (src/cli.rs:510-551):
  ``q,total[,atoms,excluded,hydration]`` with six decimals.
* ``write_fit_csv``: ``write_fit_csv`` (src/fit.rs:548-560):
  ``q,i_exp,i_fit,sigma`` with i_fit = scale x spline of the fitted total.
"""
from __future__ import annotations

import math
import os
import re
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .errors import FileError, ValidationError

# Rust's str::parse::<f64> grammar (no surrounding whitespace, no '_')
_F64 = re.compile(r"[+-]?(inf|infinity|nan|(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", re.IGNORECASE)


def _parse_f64(text: str) -> Optional[float]:
    return float(text) if _F64.fullmatch(text) else None


@dataclass
class ExpData:
    """Experimental curve (reference ``ExpData``, src/fit.rs:17-25)."""
    q: np.ndarray
    intensity: np.ndarray
    sigma: np.ndarray

    @classmethod
    def from_file(cls, path) -> "ExpData":
        return read_exp_data(path)

    @classmethod
    def from_arrays(cls, q, intensity, sigma=None) -> "ExpData":
        """``ExpData::from_arrays`` (src/fit.rs:89-118): equal lengths, sigma
        empty (sigma = 1) or one per point."""
        q = np.asarray(q, dtype=np.float64)
        i = np.asarray(intensity, dtype=np.float64)
        s = np.asarray([] if sigma is None else sigma, dtype=np.float64)
        if len(q) == 0:
            raise ValidationError("experimental data is empty")
        if len(i) != len(q):
            raise ValidationError(f"q ({len(q)}) and intensity ({len(i)}) length mismatch")
        if len(s) == 0:
            s = np.ones(len(q))
        elif len(s) != len(q):
            raise ValidationError(f"q ({len(q)}) and sigma ({len(s)}) length mismatch")
        return cls(q, i, s)


def read_exp_data(path) -> ExpData:
    """Load q, I [, sigma] from a text file.  A line is skipped when, after
    trimming, it is empty, starts with '#' or with an alphabetic character,
    has fewer than two fields, or its first two fields are not numbers.  A
    comma anywhere on the line selects comma separation (fields trimmed),
    otherwise whitespace.  A missing or unparsable third field gives
    sigma = 1."""
    try:
        with open(path, "r", encoding="utf-8") as fh:
            lines = fh.read().split("\n")
    except OSError as exc:
        raise FileError(f"{path}: {exc.strerror or exc}") from exc
    q, inten, sig = [], [], []
    for line in lines:
        t = line.strip()
        if not t or t.startswith("#") or t[0].isalpha():
            continue
        parts = [p.strip() for p in t.split(",")] if "," in t else t.split()
        if len(parts) < 2:
            continue
        qi, ii = _parse_f64(parts[0]), _parse_f64(parts[1])
        if qi is None or ii is None:
            continue
        si = _parse_f64(parts[2]) if len(parts) > 2 else None
        q.append(qi)
        inten.append(ii)
        sig.append(1.0 if si is None else si)
    if not q:
        raise ValidationError(
            f"no data points found in {os.fspath(path)} (expected 2–3 columns: q, I [, σ]; "
            "whitespace or comma separated)")
    return ExpData(np.array(q), np.array(inten), np.array(sig))


def _fmt(x: float) -> str:
    """Rust's `{:.6}` for f64."""
    x = float(x)
    if math.isnan(x):
        return "NaN"
    if math.isinf(x):
        return "inf" if x > 0 else "-inf"
    return f"{x:.6f}"


def write_intensity_csv(path, curve) -> None:
    """``q,total`` plus ``atoms,excluded,hydration`` when the curve carries
    the three contributions (a cached model's intensity always does)."""
    has_c = curve.atoms is not None and curve.excluded is not None and curve.hydration is not None
    try:
        with open(path, "w", encoding="utf-8", newline="\n") as out:
            out.write("q,total" + (",atoms,excluded,hydration" if has_c else "") + "\n")
            for k in range(len(curve.q)):
                row = [_fmt(curve.q[k]), _fmt(curve.total[k])]
                if has_c:
                    row += [_fmt(curve.atoms[k]), _fmt(curve.excluded[k]), _fmt(curve.hydration[k])]
                out.write(",".join(row) + "\n")
    except OSError as exc:
        raise FileError(f"{path}: {exc.strerror or exc}") from exc


def write_fit_csv(path, exp: ExpData, model, fit) -> None:
    """``q,i_exp,i_fit,sigma`` for a fit of ``model`` (a ``DebyeModel``) to
    ``exp``: i_fit = fit.scale x spline of I(q; fit.volume_scale,
    fit.contrast_density) at the experimental q, computed on the device."""
    theory = model.fit_curve(exp.q, fit.volume_scale, fit.contrast_density)
    try:
        with open(path, "w", encoding="utf-8", newline="\n") as out:
            out.write("q,i_exp,i_fit,sigma\n")
            for k in range(len(exp.q)):
                out.write(",".join([_fmt(exp.q[k]), _fmt(exp.intensity[k]),
                                    _fmt(fit.scale * theory[k]), _fmt(exp.sigma[k])]) + "\n")
    except OSError as exc:
        raise FileError(f"{path}: {exc.strerror or exc}") from exc
