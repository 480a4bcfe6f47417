"""GPU test of pripps_debye_build_noalloc, the build whose device scratch is
one fixed array per GPU (no device allocation), through the C ABI: two calls
in a row, the second on a smaller structure, each against
pripps_debye_model_build with the same plan (bit for bit) and with the
default plan (FP32 1e-5, FP64 1e-12 envelope error); a structure whose plan
cannot fit is refused with an out-of-memory status."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from test_gpu_parity import envelope_error, make_case

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-5, "fp64": 1e-12}


def _noalloc(engine, inp, precision, tune=None):
    from pripps_amd import _native as N
    out = np.full((6, len(inp.q)), np.nan)
    rc = N.load().pripps_debye_build_noalloc(
        engine.handle, ctypes.byref(inp.c), N.PRECISIONS[precision],
        None if tune is None else ctypes.byref(tune), out.ctypes.data_as(N._D))
    return rc, out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_noalloc_matches_model_build(engine, precision):
    from pripps_amd._native import PreparedInput, tuning
    from pripps_amd.errors import from_status
    big = PreparedInput(*make_case(3000, 71, nq=300))
    small = PreparedInput(*make_case(700, 72, nq=90))
    for inp in (big, small):  # second call: a smaller input on the same array
        print(f"noalloc {precision} n={len(inp.x)} nq={len(inp.q)}", flush=True)
        t = tuning(q_per_lane=64, block=64, span=256)
        rc, got = _noalloc(engine, inp, precision, t)
        if rc:
            raise from_status(rc, "noalloc build failed")
        want = engine.build(inp, precision, t).profiles()
        assert np.array_equal(got, want)
        rc, got = _noalloc(engine, inp, precision)
        assert rc == 0
        m = engine.build(inp, precision)
        assert envelope_error(got, m.profiles()).max() <= TOL[precision]


def test_noalloc_wider_span_and_refusal(engine, foxs):
    """The 10,000-atom solvent globule at nq = 500.  Its default (Chebyshev)
    plan fits the fixed scratch and meets the FP32 bound against the ordinary
    build.  The 128-q recurrence plan's item partials (about 77 MB) do not:
    without a span the call takes longer j-spans until they fit and still
    meets the bound; with an explicit short span it is refused."""
    from pripps_amd._native import tuning
    from pripps_amd.errors import OutOfDeviceMemoryError, from_status
    from pripps_amd.synth import solvent10k
    wl = solvent10k()
    inp = foxs.prepare_workload(wl)
    ref = engine.build(inp, "fp32").profiles()
    rc, got = _noalloc(engine, inp, "fp32")
    assert rc == 0
    assert envelope_error(got, ref).max() <= 1e-5
    rc, got = _noalloc(engine, inp, "fp32", tuning(q_per_lane=128))
    assert rc == 0
    assert envelope_error(got, ref).max() <= 1e-5
    rc, _ = _noalloc(engine, inp, "fp32", tuning(q_per_lane=128, span=64))
    with pytest.raises(OutOfDeviceMemoryError, match="fixed scratch"):
        from pripps_amd._native import load
        raise from_status(rc, load().pripps_last_error().decode())
    # the array is still usable after a refusal
    small = make_case(500, 73, nq=64)
    from pripps_amd._native import PreparedInput
    s_inp = PreparedInput(*small)
    rc, got = _noalloc(engine, s_inp, "fp32", tuning(q_per_lane=64, block=64, span=64))
    assert rc == 0
    assert np.array_equal(got, engine.build(s_inp, "fp32", tuning(q_per_lane=64, block=64,
                                                                   span=64)).profiles())
