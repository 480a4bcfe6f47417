#!/usr/bin/env python3
"""Benchmark of the exact Debye engine on one MI355X.

Workload (the metric's configuration): a seeded 10,000-atom globule
(R = 37 A, C/N/O/S 63/17/19/1 %), FoXS excluded volume plus surface
hydration weights, uniform q on [0, 0.5] 1/A with nq = 500, all six partial
profiles, FP32-compensated.  One step is one full pass of the pair kernel and
the finalize kernel over that structure, with inputs already in HBM.

Reported beside the throughput:
  * precision: max envelope error of the FP32 profiles against the GPU FP64
    mode (all q) and against the CPU oracle (sampled q);
  * roofline of the pair kernel: its executed FP32 FLOP rate (counter figures
    of the same plan over the live HIP-event time) against the FP32 vector
    peak, its VALU issue fraction, and the baseline's 14 FLOP per pair-q
    term over the same time for reference;
  * the plain FP64 GPU path on the same structure (its pair kernel time);
  * cpu_baseline: the oracle's reference-order FP64 loop (3 self + 3 cross
    sums per q, libm sin, OpenMP over q) on k evenly spaced q points;
  * extra.batch512: the batched configuration (512 structures in one
    persistent launch) against the plain FP64 path on a sample of them;
  * served: Poisson request streams at three offered rates through the
    serving loop, p50/p99 latency.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Debye pair·q terms/s at ≤1e-5 rel err vs FP64; served p50/p99 latency"
UNIT = "pair·q terms/s"
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
FLOP_PER_TERM_SOLVENT = 14.0
# Counter figures of the pair kernel on the default solvent10k plan, from one
# rocprofv3 run each (profiles/debye_cheb_f32_solvent10k.md):
#   bytes: HBM bytes per launch, FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (scale
#          factors from a plain-copy calibration);
#   executed_flop_per_term: FP32 FLOP per dispatch, (2 FMA + ADD + MUL + TRANS)
#          x 64 lanes from SQ_INSTS_VALU_{FMA,ADD,MUL,TRANS}_F32, over the pass's
#          pair-q terms;
#   valu_wave_instr: SQ_INSTS_VALU per dispatch; with the dispatch's
#          GRBM_GUI_ACTIVE cycles per XCD it gives the issue rate.
# Keyed by the library's version string (pripps_version(), changed whenever
# the kernel's instruction stream or its work plan changes) and the plan, so
# that a library or plan the counters were not measured on finds no entry.
COUNTERS = {
    ("pripps_amd 0.3.1 (gfx950)", "debye_cheb_kernel<26>", 256, 48, 0, 3072): {
        "bytes": (1245.7 * 2.0 + 7772.1 * 1.0) * 1024,
        "executed_flop_per_term": 64.0 * (2 * 9.9451e7 + 2.0069e7 + 3.1723e7 + 2.5944e6) / 2.49975e10,
        "valu_wave_instr": 1.7990e8, "cycles_per_xcd": 4.8853e6 / 8,
        "profile": "profiles/debye_cheb_f32_solvent10k.md"},
}
# The same figures for the FP64 pair kernel on the default FP64 plan of
# solvent10k (profiles/debye_cheb64_f64_solvent10k.md); executed FLOP from
# SQ_INSTS_VALU_{FMA,ADD,MUL,TRANS}_F64.  FP64 instructions do not dual-issue:
# one per quad-cycle per SIMD at most.
COUNTERS_F64 = {
    ("pripps_amd 0.3.1 (gfx950)", "debye_cheb64_kernel<36>", 256, 2048): {
        "bytes": (2244.3 * 2.0 + 7096.1 * 1.0) * 1024,
        "executed_flop_per_term": 64.0 * (2 * 1.9034e8 + 4.1057e7 + 5.9005e7 + 5.1888e6) / 2.49975e10,
        "valu_wave_instr": 3.4242e8, "cycles_per_xcd": 1.2583e7 / 8,
        "profile": "profiles/debye_cheb64_f64_solvent10k.md"},
}
VALU_ISSUE_PEAK_F64 = 0.25
# VALU wave instructions a SIMD can issue per cycle (two per quad-cycle with
# dual issue), and the engine clock
VALU_ISSUE_PEAK = 0.5
CLOCK_HZ = 2.4e9
PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def envelope_error(got, ref):
    import numpy as np
    worst = 0.0
    for p, (x, y) in enumerate(PAIRS):
        env = np.sqrt(np.abs(ref[x]) * np.abs(ref[y]))
        ok = env > 0
        if np.any(ok):
            worst = max(worst, float(np.max(np.abs(got[p][ok] - ref[p][ok]) / env[ok])))
    return worst


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # the GPU's clocks take a few hundred milliseconds of load to settle:
    # 300 untimed passes (~0.1 s) before the 200 timed ones (~75 ms)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--nq", type=int, default=500)
    ap.add_argument("--q-per-lane", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--span", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--fp64-q-per-lane", type=int, default=0,
                    help="q-points per strip of the FP64 leg (0: the default plan; 32 or 64: "
                         "the tiled FP64 recurrence kernel)")
    ap.add_argument("--cpu-q", type=int, default=8, help="q points of the CPU baseline sample")
    ap.add_argument("--cpu-q1", type=int, default=1,
                    help="q points of the one-core CPU baseline sample (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--workload", default="solvent10k",
                    choices=("solvent10k", "vacuum1k", "helix60k", "batch512", "served", "direct",
                             "frames"))
    ap.add_argument("--frames", type=int, default=64, help="frames: trajectory frames")
    ap.add_argument("--cpu-vectors", type=int, default=256,
                    help="direct: q-vectors of the oracle / CPU baseline sample")
    ap.add_argument("--count", type=int, default=512, help="batch512: structures")
    ap.add_argument("--plain-sample", type=int, default=512,
                    help="batch512: structures timed on the sequential plain FP64 path "
                         "(all of them by default)")
    ap.add_argument("--rates", default="10,50,100,200,500,1000,2000")
    ap.add_argument("--requests", type=int, default=100, help="served: requests per rate")
    ap.add_argument("--total-requests", type=int, default=0,
                    help="served: requests over the whole sweep, in proportion to the rate "
                         "(0: --requests per rate)")
    ap.add_argument("--served-rates", default="100,400,1000",
                    help="default run: offered rates of its served streams ('' skips them)")
    ap.add_argument("--served-seconds", type=float, default=1.5,
                    help="default run: length of each served stream (requests = rate x seconds)")
    ap.add_argument("--served-requests", type=int, default=None,
                    help="0 skips the default run's served streams (same as --served-rates '')")
    ap.add_argument("--batch-count", type=int, default=512,
                    help="default run: structures of its batched launch (0: skip)")
    ap.add_argument("--batch-plain-sample", type=int, default=24,
                    help="default run: batch structures timed on the sequential plain FP64 path")
    args = ap.parse_args()
    if args.served_requests == 0:
        args.served_rates = ""
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the engine runs on one GPU and has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pripps_amd
    from pripps_amd import _native
    from pripps_amd.synth import solvent10k

    if args.workload != "solvent10k":
        WORKLOADS[args.workload](args)
        return

    t_setup = time.time()
    wl = solvent10k(n=args.n, nq=args.nq)
    yaml_path = os.path.join(HERE, "tests", "golden", "reference_data", "foxs_formfactors.yaml")
    eng = pripps_amd.Engine(0)
    p = pripps_amd.Pripps(yaml_path, engine=eng)
    ids = p.names_to_ids(wl.names)
    inp = p.prepare_input(wl.pos, ids, wl.q, excluded="foxs", hydration="weights",
                          hydration_weights=wl.weights)
    terms = wl.pair_q_terms
    log(f"workload {wl.name}: N={wl.n_atoms} nq={len(wl.q)} terms={terms:.4e} "
        f"(setup {time.time() - t_setup:.1f}s)")
    tune = _native.tuning(args.q_per_lane, args.block, args.span, args.variant)

    # FP64 references on the GPU (also timed: the plain path is the baseline of
    # the batched/compensated target)
    fp64_prof = None
    extra = {}
    if not args.no_fp64:
        m64 = eng.build(inp, "fp64_plain")
        plain_prof = m64.profiles()
        s = m64.stats()
        extra["fp64_plain_kernel_ms"] = s.kernel_ms
        log(f"fp64_plain build: kernel {s.kernel_ms:.2f} ms")
        m64t = eng.build(inp, "fp64", _native.tuning(args.fp64_q_per_lane) if args.fp64_q_per_lane
                         else None)
        fp64_prof = m64t.profiles()
        s64 = m64t.stats()
        # warm passes first: the clocks take a few tens of passes to settle
        m64t.recompute(30)
        pm, _ = m64t.recompute(20)
        k64 = (f"debye_cheb64_kernel<{s64.cheb_order}>" if s64.q_per_lane == 256 else
               "debye_tiled_f64_kernel<%d>" % s64.q_per_lane)
        log(f"fp64 build: Q={s64.q_per_lane} P={s64.cheb_order} units={s64.work_items} "
            f"kernel {k64}")
        extra["fp64_pair_kernel_ms"] = pm
        extra["fp64_kernel"] = k64
        extra["fp64_vs_plain_pair_kernel"] = s.kernel_ms / pm
        ctr64 = COUNTERS_F64.get((_native.load().pripps_version().decode(), k64, s64.q_per_lane,
                                  s64.work_items))
        r64 = {"bound": "valu", "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "kernel": k64,
               "kernel_ms": pm, "achieved": None, "frac": None, "issue_frac": None,
               "achieved_algorithmic": FLOP_PER_TERM_SOLVENT * terms / (pm * 1e-3) / 1e12,
               "flop_per_term_algorithmic": FLOP_PER_TERM_SOLVENT}
        if ctr64 is not None:
            tf64 = ctr64["executed_flop_per_term"] * terms / (pm * 1e-3) / 1e12
            r64.update({
                "achieved": tf64, "frac": tf64 / FP64_PEAK_TFLOPS,
                "executed_flop_per_term": ctr64["executed_flop_per_term"],
                "traffic": ctr64["bytes"],
                "issue_frac": ctr64["valu_wave_instr"] / (1024 * pm * 1e-3 * CLOCK_HZ) / VALU_ISSUE_PEAK_F64,
                "issue_frac_profiled": ctr64["valu_wave_instr"] / 1024 / ctr64["cycles_per_xcd"]
                / VALU_ISSUE_PEAK_F64,
                "source": ctr64["profile"]})
        extra["fp64_roofline"] = r64
        extra["fp64_vs_plain_maxerr"] = envelope_error(fp64_prof, plain_prof)
        log(f"fp64 pair kernel {pm:.3f} ms; against the plain path {extra['fp64_vs_plain_maxerr']:.2e}")
        del m64, m64t

    model = eng.build(inp, "fp32", tune)
    st = model.stats()
    prof32 = model.profiles()
    log(f"fp32 build: Q={st.q_per_lane} block={st.block} span={st.span} items={st.work_items} "
        f"kernel {st.kernel_ms:.2f} ms")
    # the timed passes replay the build's kernel launches as one captured HIP
    # graph per pass; the pair kernel's own time comes from eager passes with
    # HIP events around it, right after
    if args.warmup > 0:
        model.recompute(args.warmup, graph=True)
    t0 = time.perf_counter()
    model.recompute(args.steps, graph=True)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / max(args.steps, 1)
    value = terms / (ms_per_step * 1e-3)
    pair_ms, _ = model.recompute(max(20, min(args.steps, 100)))
    again = model.profiles()
    assert np.array_equal(again, prof32), "recompute changed the profiles"
    # PCIe-inclusive: host arrays in, profiles back on the host (upload, plan,
    # kernels, download), median of fresh builds; reported beside `value`
    walls = []
    for _ in range(5):
        tb = time.perf_counter()
        mb = eng.build(inp, "fp32", tune)
        mb.profiles()
        walls.append((time.perf_counter() - tb) * 1e3)
        del mb
    extra["fp32_build_wall_ms_median"] = float(np.median(walls))
    extra["pcie_inclusive_terms_per_s"] = terms / (np.median(walls) * 1e-3)
    log(f"fp32 fresh build, host in/out: median {np.median(walls):.2f} ms "
        f"(min {min(walls):.2f}, max {max(walls):.2f})")
    achieved = FLOP_PER_TERM_SOLVENT * terms / (pair_ms * 1e-3) / 1e12
    log(f"fp32: {ms_per_step:.3f} ms/step wall, pair kernel {pair_ms:.3f} ms (events), "
        f"{value:.4e} terms/s, {achieved:.1f} TFLOP/s algorithmic")

    err_fp64 = envelope_error(prof32, fp64_prof) if fp64_prof is not None else None
    if fp64_prof is not None and "fp64_plain_kernel_ms" in extra:
        extra["speedup_vs_fp64_plain"] = extra["fp64_plain_kernel_ms"] / pair_ms

    cpu = None
    err_oracle = None
    if not args.no_cpu and args.cpu_q > 0:
        sys.path.insert(0, os.path.join(HERE, "oracle"))
        import oracle as O
        k = min(args.cpu_q, len(wl.q))
        sel = np.unique(np.linspace(0, len(wl.q) - 1, k).round().astype(int))
        threads = min(16, os.cpu_count() or 1)
        tc = time.perf_counter()
        ref = O.build_model(wl.pos, inp.kind, inp.q[sel], inp.atom_ff[:, sel],
                            inp.excluded_ff[:, sel], inp.water_ff[sel], inp.weights,
                            mode=O.REFERENCE_ORDER, nthreads=threads)
        tcpu = time.perf_counter() - tc
        cpu_terms = wl.n_atoms * (wl.n_atoms - 1) / 2 * len(sel)
        cpu = {"value": cpu_terms / tcpu, "unit": UNIT, "cores": threads, "kind": "port",
               "sample": f"oracle reference-order FP64 loop (3 self + 3 cross sums per q) on "
                         f"{len(sel)} evenly spaced q of the same structure, {tcpu:.1f} s"}
        err_oracle = envelope_error(prof32[:, sel], ref)
        log(f"cpu baseline: {cpu['value']:.3e} terms/s on {threads} threads ({tcpu:.1f} s); "
            f"fp32 vs oracle {err_oracle:.2e}")
        if args.cpu_q1 > 0:
            # the same loop on one core, over fewer q
            sel1 = sel[len(sel) // 2: len(sel) // 2 + args.cpu_q1]
            tc = time.perf_counter()
            O.build_model(wl.pos, inp.kind, inp.q[sel1], inp.atom_ff[:, sel1],
                          inp.excluded_ff[:, sel1], inp.water_ff[sel1], inp.weights,
                          mode=O.REFERENCE_ORDER, nthreads=1)
            t1c = time.perf_counter() - tc
            extra["cpu_baseline_1core"] = {
                "value": wl.n_atoms * (wl.n_atoms - 1) / 2 * len(sel1) / t1c, "unit": UNIT,
                "cores": 1, "kind": "port",
                "sample": f"same loop, {len(sel1)} q, {t1c:.1f} s"}
            log(f"cpu baseline, 1 core: {extra['cpu_baseline_1core']['value']:.3e} terms/s "
                f"({t1c:.1f} s)")

    kernel = (f"debye_cheb_kernel<{st.cheb_order}>" if st.q_per_lane == 256 else
              "debye_dual_kernel" if st.q_per_lane == 128 else
              "debye_tiled_kernel<float,%d,true,3,true>" % st.q_per_lane)
    lib_version = _native.load().pripps_version().decode()
    ctr = COUNTERS.get((lib_version, kernel, st.q_per_lane, st.block, st.span, st.work_items))
    roofline = {"bound": "valu", "unit": "TFLOP/s", "peak": FP32_PEAK_TFLOPS, "kernel": kernel,
                "kernel_ms": pair_ms, "traffic": None, "achieved": None, "frac": None,
                "issue_frac": None,
                # the baseline's 14 FLOP per pair-q term (angle-addition
                # count) over the same time: this kernel executes far fewer
                "achieved_algorithmic": achieved,
                "flop_per_term_algorithmic": FLOP_PER_TERM_SOLVENT}
    if ctr is not None:
        tf = ctr["executed_flop_per_term"] * terms / (pair_ms * 1e-3) / 1e12
        roofline.update({
            "achieved": tf, "frac": tf / FP32_PEAK_TFLOPS,
            "executed_flop_per_term": ctr["executed_flop_per_term"],
            "traffic": ctr["bytes"],
            # issued VALU wave instructions per SIMD-cycle over the live
            # kernel time at the engine clock, against two per quad-cycle
            "issue_frac": ctr["valu_wave_instr"] / (1024 * pair_ms * 1e-3 * CLOCK_HZ) / VALU_ISSUE_PEAK,
            "issue_frac_profiled": ctr["valu_wave_instr"] / 1024 / ctr["cycles_per_xcd"] / VALU_ISSUE_PEAK,
            "valu_instr_per_term": ctr["valu_wave_instr"] * 64 / terms,
            "source": ctr["profile"]})
    log(f"roofline: {roofline}")

    if args.batch_count > 0:
        extra["batch512"] = _batch_line(eng, p, args)
    served = None
    if args.served_rates:
        rates = [float(r) for r in args.served_rates.split(",")]
        counts = [int(round(r * args.served_seconds)) for r in rates]
        served = {"workload": SERVED_WORKLOAD, "seconds_per_rate": args.served_seconds,
                  "rates": _served_streams(eng, p, rates, counts)}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": UNIT,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "solvent10k: 10,000-atom globule R=37 A, FoXS excluded volume + "
                               "surface hydration weights, q uniform [0,0.5] nq=500, 6 profiles",
                   "n_atoms": wl.n_atoms, "nq": len(wl.q), "pair_q_terms": terms,
                   "precision": "fp32_compensated", "q_per_lane": st.q_per_lane,
                   "cheb_order": st.cheb_order, "block": st.block, "span": st.span,
                   "work_items": st.work_items},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "max_rel_err_vs_fp64_gpu": err_fp64,
        "max_rel_err_vs_oracle_sampled_q": err_oracle,
        "served": served,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------------------
# the other benchmark configurations (one JSON line each)
# ---------------------------------------------------------------------------
def _engine_and_ff():
    import pripps_amd
    eng = pripps_amd.Engine(0)
    yaml_path = os.path.join(HERE, "tests", "golden", "reference_data", "foxs_formfactors.yaml")
    return eng, pripps_amd.Pripps(yaml_path, engine=eng)


def _oracle():
    sys.path.insert(0, os.path.join(HERE, "oracle"))
    import oracle as O
    return O


def _per_q_error(got, ref):
    """Envelope error per q (max over the six profiles)."""
    import numpy as np
    err = np.zeros(got.shape[1])
    for p, (x, y) in enumerate(PAIRS):
        env = np.sqrt(np.abs(ref[x]) * np.abs(ref[y]))
        env = np.where(env > 0, env, 1.0)
        err = np.maximum(err, np.abs(got[p] - ref[p]) / env)
    return err


def _timed(model, steps, warmup):
    if warmup:
        model.recompute(warmup)
    t0 = time.perf_counter()
    pair_ms, _ = model.recompute(steps)
    return (time.perf_counter() - t0) * 1e3 / max(steps, 1), pair_ms


def bench_vacuum1k(args):
    """1,000-atom globule in vacuum, nq = 500: FP64 modes against the oracle."""
    import numpy as np
    from pripps_amd.synth import vacuum1k
    O = _oracle()
    eng, ff = _engine_and_ff()
    wl = vacuum1k()
    inp = ff.prepare_workload(wl)
    terms = wl.pair_q_terms
    threads = min(16, os.cpu_count() or 1)
    tc = time.perf_counter()
    ref_order = O.build_model(wl.pos, inp.kind, inp.q, inp.atom_ff, mode=O.REFERENCE_ORDER,
                              nthreads=threads)
    tcpu = time.perf_counter() - tc
    ref = O.build_model(wl.pos, inp.kind, inp.q, inp.atom_ff, mode=O.COMPENSATED, nthreads=threads)
    out = {"workload": "vacuum1k: 1,000-atom globule R=17 A, vacuum, q uniform [0,0.5] nq=500",
           "pair_q_terms": terms,
           "cpu_reference_order": {"value": terms / tcpu, "unit": UNIT, "cores": threads,
                                   "kind": "port", "seconds": tcpu,
                                   "max_err_vs_compensated": envelope_error(ref_order, ref)}}
    for prec in ("fp64_plain", "fp64", "fp32"):
        m = eng.build(inp, prec)
        prof = m.profiles()
        rec = {"max_err_vs_oracle": envelope_error(prof, ref), "build_kernel_ms": m.stats().kernel_ms}
        if prec != "fp64_plain":
            ms, pair_ms = _timed(m, args.steps, args.warmup)
            rec.update({"ms_per_step": ms, "pair_kernel_ms": pair_ms, "terms_per_s": terms / (ms * 1e-3)})
        else:
            rec["terms_per_s"] = terms / (rec["build_kernel_ms"] * 1e-3)
        out[prec] = rec
        log(f"vacuum1k {prec}: {rec}")
    print(json.dumps(out), flush=True)


def bench_helix60k(args):
    """Six 10k-atom globules on a helix, log-spaced q (no recurrence)."""
    import numpy as np
    from pripps_amd.synth import helix60k
    O = _oracle()
    eng, ff = _engine_and_ff()
    wl = helix60k()
    inp = ff.prepare_workload(wl)
    terms = wl.pair_q_terms
    out = {"workload": "helix60k: 6 x 10,000-atom globules on a helix (pitch 60 A, 3/4 turn per "
                       "globule, radius 45 A), solvent, q log-spaced [0.005,1.0] nq=1000",
           "pair_q_terms": terms}
    profs = {}
    for prec in ("fp64", "fp32"):
        m = eng.build(inp, prec)
        profs[prec] = m.profiles()
        ms, pair_ms = _timed(m, max(1, args.steps // 10), 0)
        out[prec] = {"ms_per_step": ms, "pair_kernel_ms": pair_ms, "terms_per_s": terms / (ms * 1e-3),
                     "build_kernel_ms": m.stats().kernel_ms}
        log(f"helix60k {prec}: {out[prec]}")
    perq = _per_q_error(profs["fp32"], profs["fp64"])
    worst = np.argsort(perq)[::-1][:5]
    out["fp32_vs_fp64_per_q"] = {"max": float(perq.max()), "median": float(np.median(perq)),
                                 "p90": float(np.percentile(perq, 90)),
                                 "p99": float(np.percentile(perq, 99)),
                                 "q_at_max": float(wl.q[int(perq.argmax())]),
                                 "worst": [[float(wl.q[k]), float(perq[k])] for k in worst],
                                 "values": [float(f"{v:.3g}") for v in perq]}
    sel = np.array([150, 500, 900])
    threads = min(16, os.cpu_count() or 1)
    tc = time.perf_counter()
    ref = O.build_model(wl.pos, inp.kind, inp.q[sel], inp.atom_ff[:, sel], inp.excluded_ff[:, sel],
                        inp.water_ff[sel], inp.weights, mode=O.COMPENSATED, nthreads=threads)
    out["oracle_sampled_q"] = {"q": [float(x) for x in wl.q[sel]], "seconds": time.perf_counter() - tc,
                               "fp64_max_err": envelope_error(profs["fp64"][:, sel], ref),
                               "fp32_max_err": envelope_error(profs["fp32"][:, sel], ref)}
    log(f"helix60k oracle: {out['oracle_sampled_q']}")
    print(json.dumps(out), flush=True)


def _make_batch(count, seed=4):
    from concurrent.futures import ProcessPoolExecutor
    from pripps_amd import synth
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(500), math.log(20000), size=count)).astype(int)
    nqs = rng.integers(100, 1001, size=count)
    qmaxs = rng.uniform(0.3, 0.8, size=count)
    jobs = [(int(ns[s]), synth.radius_for(int(ns[s])), seed * 100000 + s) for s in range(count)]
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        gl = list(ex.map(_globule_job, jobs))
    return [synth.Workload(f"batch{s}", gl[s][0], gl[s][1], synth.linear_q(0.0, float(qmaxs[s]), int(nqs[s])),
                           gl[s][2]) for s in range(count)]


def _globule_job(job):
    from pripps_amd.synth import globule
    n, r, seed = job
    return globule(n, r, seed, solvent=True, jitter=0.5)


def _batch_line(eng, ff, args):
    """The batched configuration inside the default run: the structures in
    one persistent FP32 launch per kernel shape (resident inputs,
    `recompute`), against an evenly spaced sample of them built one at a time
    on the plain FP64 path (the rate extrapolated by pair-q terms)."""
    import numpy as np
    t = time.perf_counter()
    wls = _make_batch(args.batch_count)
    inps = [ff.prepare_workload(w) for w in wls]
    terms = float(sum(w.pair_q_terms for w in wls))
    tb = time.perf_counter()
    models = eng.build_batch(inps, "fp32")
    first_wall = (time.perf_counter() - tb) * 1e3
    ms, pair_ms = _timed(models[0], 4, 1)
    del models
    k = min(args.batch_plain_sample, len(wls))
    sel = np.unique(np.linspace(0, len(wls) - 1, k).round().astype(int))
    plain_ms, plain_terms = 0.0, 0.0
    for s in sel:
        m = eng.build(inps[s], "fp64_plain")
        plain_ms += m.stats().kernel_ms
        plain_terms += wls[s].pair_q_terms
        del m
    plain_rate = plain_terms / (plain_ms * 1e-3)
    rec = {"workload": f"batch{len(wls)}: globules N log-uniform [500,20000], jitter 0.5 A, own q "
                       "grid nq U[100,1000] qmax U[0.3,0.8], solvent; one persistent FP32 launch per kernel shape",
           "structures": len(wls), "pair_q_terms": terms, "ms_per_step": ms, "pair_kernel_ms": pair_ms,
           "terms_per_s": terms / (ms * 1e-3), "first_build_wall_ms": first_wall,
           "fp64_plain_sequential": {"sample": int(len(sel)), "sample_terms": plain_terms,
                                     "sample_kernel_ms": plain_ms, "terms_per_s": plain_rate},
           "speedup_vs_sequential_plain": terms / (ms * 1e-3) / plain_rate,
           "setup_s": tb - t}
    log(f"batch: {rec}")
    return rec


def bench_batch512(args):
    """512 structures of uneven size and q grid in one persistent launch per
    kernel shape, against the plain FP64 path launched once per structure."""
    import numpy as np
    eng, ff = _engine_and_ff()
    t = time.perf_counter()
    wls = _make_batch(args.count)
    inps = [ff.prepare_workload(w) for w in wls]
    terms = float(sum(w.pair_q_terms for w in wls))
    log(f"batch: {len(wls)} structures, {terms:.4e} terms (setup {time.perf_counter() - t:.1f} s)")
    tb = time.perf_counter()
    models = eng.build_batch(inps, "fp32")
    first_wall = (time.perf_counter() - tb) * 1e3  # host arrays in, models on the device
    first_ms = models[0].stats().kernel_ms
    ms, pair_ms = _timed(models[0], max(1, args.steps // 4), 1)
    out = {"workload": f"batch{len(wls)}: globules N log-uniform [500,20000], jitter 0.5 A, own q grid "
                       "nq U[100,1000] qmax U[0.3,0.8], solvent",
           "structures": len(wls), "pair_q_terms": terms,
           "fp32_batch": {"ms_per_step": ms, "pair_kernel_ms": pair_ms, "first_build_kernel_ms": first_ms,
                          "first_build_wall_ms": first_wall,
                          "terms_per_s": terms / (ms * 1e-3)}}
    log(f"fp32 batch: {out['fp32_batch']}")
    m64 = eng.build_batch(inps, "fp64")
    ms64, pair64 = _timed(m64[0], 1, 0)
    out["fp64_batch"] = {"ms_per_step": ms64, "pair_kernel_ms": pair64, "terms_per_s": terms / (ms64 * 1e-3)}
    errs = [envelope_error(a.profiles(), b.profiles()) for a, b in zip(models, m64)]
    out["fp32_vs_fp64_max_err"] = float(max(errs))
    del m64
    # sequential plain FP64 launches, one per structure (all of them, or an
    # evenly spaced sample extrapolated by terms)
    k = min(args.plain_sample, len(wls))
    sel = np.unique(np.linspace(0, len(wls) - 1, k).round().astype(int))
    plain_ms, plain_terms = 0.0, 0.0
    tw = time.perf_counter()
    for n_done, s in enumerate(sel):
        m = eng.build(inps[s], "fp64_plain")
        plain_ms += m.stats().kernel_ms
        plain_terms += wls[s].pair_q_terms
        del m
        if (n_done + 1) % 128 == 0:
            log(f"plain sequential: {n_done + 1} of {len(sel)} structures")
    plain_wall = (time.perf_counter() - tw) * 1e3
    plain_rate = plain_terms / (plain_ms * 1e-3)
    out["fp64_plain_sequential"] = {"sample": int(len(sel)), "sample_terms": plain_terms,
                                    "sample_kernel_ms": plain_ms, "sample_wall_ms": plain_wall,
                                    "terms_per_s": plain_rate,
                                    "extrapolated_ms": terms / plain_rate * 1e3}
    out["speedup_vs_sequential_plain"] = out["fp32_batch"]["terms_per_s"] / plain_rate
    log(f"plain sequential sample: {out['fp64_plain_sequential']}; speedup {out['speedup_vs_sequential_plain']:.1f}x")
    print(json.dumps(out), flush=True)


def bench_frames(args):
    """Trajectory frame averaging (the reference's trajectory_average for the
    Debye scheme): a solvated globule topology with Gaussian-jittered frames
    and per-frame weights, through the batched, graph-captured frames path,
    against building the frames one at a time.  Two topologies: the
    solvent10k structure (each frame fills the GPU by itself) and a
    2,000-atom one (many frames per launch)."""
    import numpy as np
    from pripps_amd._native import PreparedInput
    from pripps_amd.synth import solvent10k
    eng, ff = _engine_and_ff()
    cases = []
    for n_atoms, n_frames in ((10000, args.frames), (2000, 8 * args.frames)):
        wl = solvent10k(n=n_atoms)
        inp = ff.prepare_workload(wl)
        rng = np.random.default_rng(7)
        coords = wl.pos[None, :, :] + rng.normal(0.0, 0.5, size=(n_frames, wl.n_atoms, 3))
        weights = rng.uniform(0.5, 1.5, size=n_frames)
        terms = float(wl.pair_q_terms) * n_frames
        case = {"topology": f"{n_atoms}-atom globule, nq 500, FoXS solvent", "frames": n_frames,
                "pair_q_terms": terms}
        m = eng.frames_average(inp, coords, weights, precision="fp32")  # warm-up
        walls = []
        for _ in range(max(1, args.steps // 5)):
            t0 = time.perf_counter()
            m = eng.frames_average(inp, coords, weights, precision="fp32")
            avg = m.profiles()
            walls.append((time.perf_counter() - t0) * 1e3)
        st = m.stats()
        wall = float(np.median(walls))
        case["fp32_frames_path"] = {"wall_ms": wall, "terms_per_s": terms / (wall * 1e-3),
                                    "frames_per_s": n_frames / (wall * 1e-3),
                                    "kernel_ms": st.kernel_ms,
                                    "frames_per_launch": st.frames_per_launch,
                                    "graph_launches": st.graph_launches}
        log(f"frames {n_atoms} atoms x {n_frames}: {case['fp32_frames_path']}")
        # the same frames one build at a time (host arrays in, profiles out)
        ref = np.zeros_like(avg)
        t0 = time.perf_counter()
        for f in range(n_frames):
            one = PreparedInput(coords[f], inp.kind, inp.q, inp.atom_ff, inp.excluded_ff,
                                inp.water_ff, inp.weights)
            ref += weights[f] * eng.build(one, "fp32").profiles()
        seq = (time.perf_counter() - t0) * 1e3
        ref /= weights.sum()
        case["fp32_sequential_builds"] = {"wall_ms": seq, "terms_per_s": terms / (seq * 1e-3)}
        case["speedup_vs_sequential"] = seq / wall
        case["max_rel_err_vs_sequential"] = envelope_error(avg, ref)
        log(f"  sequential: {case['fp32_sequential_builds']}; speedup "
            f"{case['speedup_vs_sequential']:.2f}x, max envelope difference "
            f"{case['max_rel_err_vs_sequential']:.2e}")
        cases.append(case)
    print(json.dumps({"workload": "frames: Gaussian-jittered frames (N(0, 0.5 A)) of a solvated "
                                  "globule, weights U[0.5, 1.5], stride 1, frames as host arrays "
                                  "(upload included)", "cases": cases}), flush=True)


def _request_job(job):
    """One request's globule; elements come back as indices into synth.ELEMENTS."""
    import numpy as np
    from pripps_amd.synth import ELEMENTS, globule, radius_for
    n, seed = job
    names, pos, w = globule(n, radius_for(n), seed, solvent=True)
    idx = np.array([ELEMENTS.index(e) for e in names], dtype=np.int8)
    return idx, pos, w


def _served_streams(eng, ff, rates, counts):
    """Poisson request streams through the serving loop, one stream per rate
    (counts[r] requests at rates[r]).

    Each request is a fresh seeded globule (N log-uniform in [300, 30000], nq
    uniform in [200, 800], q on [0, 0.5], solvent on) followed by five
    intensity() calls at random knobs.  Latency runs from arrival to the last
    of those calls; one client thread per in-flight request does the waiting."""
    import numpy as np
    from concurrent.futures import ProcessPoolExecutor, ThreadPoolExecutor
    from pripps_amd.server import Server
    from pripps_amd.synth import ELEMENTS, linear_q, request_stream
    total = int(sum(counts))
    ns, nqs, knobs, seeds = request_stream(total)
    t = time.perf_counter()
    lut = ff.names_to_ids(list(ELEMENTS))
    inps = []
    jobs = [(int(ns[i]), int(seeds[i])) for i in range(total)]
    with ProcessPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        for i, (idx, pos, w) in enumerate(ex.map(_request_job, jobs, chunksize=8)):
            inps.append(ff.prepare_input(pos, lut[idx], linear_q(0.0, 0.5, int(nqs[i])),
                                         excluded="foxs", hydration="weights",
                                         hydration_weights=w))
            if (i + 1) % 1000 == 0:
                log(f"served: prepared {i + 1}/{total} requests ({time.perf_counter() - t:.0f} s)")
    terms = np.array([ns[i] * (ns[i] - 1) / 2 * nqs[i] for i in range(total)], dtype=float)
    log(f"served: {total} requests prepared in {time.perf_counter() - t:.1f} s, "
        f"mean {terms.mean():.3e} terms/request")
    rng = np.random.default_rng(11)
    results = []

    def finish(srv, i, rid, ts):
        m, build_ms = srv.wait(rid)
        for vs, cd in knobs[i]:
            m.intensity(vs, cd)
        done = time.perf_counter()
        m.close()
        return (done - ts) * 1e3, build_ms, done

    with Server(eng, "fp32") as srv, ThreadPoolExecutor(max_workers=64) as clients:
        for i in range(2):  # warm both lanes
            srv.wait(srv.submit(inps[i]))
        start = 0
        for rate, count in zip(rates, counts):
            idx = list(range(start, start + int(count)))
            start += int(count)
            if not idx:
                continue
            arrivals = np.cumsum(rng.exponential(1.0 / rate, size=len(idx)))
            t0 = time.perf_counter()
            futs = []
            for k, i in enumerate(idx):
                delay = t0 + arrivals[k] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ts = time.perf_counter()
                futs.append(clients.submit(finish, srv, i, srv.submit(inps[i]), ts))
            res = [f.result() for f in futs]
            t1 = max(r[2] for r in res)
            lat = np.array([r[0] for r in res])
            build = np.array([r[1] for r in res])
            rec = {"offered_rps": rate, "requests": len(idx), "wall_s": t1 - t0,
                   "completed_rps": len(idx) / (t1 - t0), "terms_per_s": float(terms[idx].sum() / (t1 - t0)),
                   "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                   "build_p50_ms": float(np.percentile(build, 50)),
                   "build_p99_ms": float(np.percentile(build, 99))}
            results.append(rec)
            log(f"served rate {rate}: {rec}")
            for i in idx:
                inps[i] = None  # release host copies as the sweep goes
    return results


SERVED_WORKLOAD = ("served: Poisson request streams, globules N log-U[300,30000], nq U[200,800], "
                   "q [0,0.5], solvent, +5 intensity() per request; latency = arrival to the last "
                   "intensity()")


def bench_served(args):
    """Poisson request streams at several rates through the serving loop."""
    import numpy as np
    eng, ff = _engine_and_ff()
    rates = [float(r) for r in args.rates.split(",")]
    if args.total_requests > 0:
        # requests in proportion to the rate, so every stream lasts about as long
        share = np.array(rates) / sum(rates) * args.total_requests
        counts = np.floor(share).astype(int)
        counts[np.argsort(counts - share)[: args.total_requests - counts.sum()]] += 1
    else:
        counts = [args.requests] * len(rates)
    results = _served_streams(eng, ff, rates, [int(c) for c in counts])
    print(json.dumps({"workload": SERVED_WORKLOAD, "rates": results}), flush=True)


def direct_box_workload(n_mol=256, n_atoms=1000, box=400.0, seed=6):
    """n_mol copies of one seeded globule (same element sequence), each rotated
    at random and placed on a jittered cubic lattice in a periodic box."""
    import numpy as np
    from scipy.spatial.transform import Rotation
    from pripps_amd.synth import globule, radius_for
    names, pos, _ = globule(n_atoms, radius_for(n_atoms), seed, solvent=False)
    pos = pos - pos.mean(axis=0)
    rng = np.random.default_rng(seed)
    side = int(math.ceil(n_mol ** (1.0 / 3.0)))
    cell = box / side
    sites = [(i, j, k) for i in range(side) for j in range(side) for k in range(side)][:n_mol]
    rot = Rotation.random(n_mol, random_state=seed).as_matrix()
    parts = []
    for m, (i, j, k) in enumerate(sites):
        centre = (np.array([i, j, k]) + 0.5) * cell + rng.uniform(-0.1, 0.1, 3) * cell
        parts.append(pos @ rot[m].T + centre)
    return names * n_mol, np.vstack(parts)


def bench_direct(args):
    """Direct Fourier transform of a periodic box of identical molecules on a
    lattice q-grid, with the per-molecule self term (effective structure
    factor): FP32 and FP64 against the oracle, and the oracle's reference-order
    loop (libm cos/sin, FP64) on a sample of q-vectors as the CPU baseline."""
    import numpy as np
    from pripps_amd.direct import QGrid
    O = _oracle()
    eng, ff = _engine_and_ff()
    n_mol, box = 256, 400.0
    t = time.perf_counter()
    names, pos = direct_box_workload(n_mol=n_mol, box=box)
    ids = ff.names_to_ids(names)
    grid = QGrid.lattice([box] * 3, 0.5, 500)
    idx, qn, group, gq = grid.groups((box,) * 3)
    terms = float(len(pos)) * len(idx)
    log(f"direct: {len(pos)} atoms, {len(idx)} q-vectors in {len(gq)} bins, {terms:.3e} atom-vector "
        f"terms (setup {time.perf_counter() - t:.1f} s)")
    out = {"workload": f"direct: {n_mol} x 1,000-atom globules (rotated, jittered lattice) in a "
                       f"{box:.0f} A periodic box, lattice q-grid q_max 0.5 max 500 vectors per bin, "
                       "per-molecule self term",
           "n_atoms": len(pos), "n_vectors": len(idx), "n_bins": len(gq), "atom_vector_terms": terms}
    res = {}
    for prec in ("fp64", "fp32"):
        m = ff.direct_model_from_arrays(pos, ids, box_sides=box, grid=grid, num_molecules=n_mol,
                                        precision=prec)
        m._native.recompute(1)
        t0 = time.perf_counter()
        kms = m._native.recompute(args.steps)
        wall = (time.perf_counter() - t0) * 1e3 / max(args.steps, 1)
        res[prec] = (m.amplitudes(), m.intensity(1.0, 0.0).total, m.self_term())
        out[prec] = {"kernel_ms": kms, "ms_per_step": wall, "terms_per_s": terms / (kms * 1e-3),
                     "work_items": int(m.stats().work_items)}
        log(f"direct {prec}: {out[prec]}")
        del m
    a32, a64 = res["fp32"][0], res["fp64"][0]
    out["fp32_vs_fp64"] = {
        "total_max_rel": float(np.max(np.abs(res["fp32"][1] - res["fp64"][1]) / res["fp64"][1])),
        "self_max_rel": float(np.max(np.abs(res["fp32"][2] - res["fp64"][2]) / res["fp64"][2])),
        "amplitude_max_abs_over_max": float(np.max(np.abs(a32[:, :2] - a64[:, :2])) / np.max(np.abs(a64[:, :2])))}
    # oracle (extended) and CPU baseline (reference order) on evenly spaced vectors
    k = min(args.cpu_vectors, len(idx))
    sel = np.unique(np.linspace(0, len(idx) - 1, k).round().astype(int))
    inp = ff.prepare_input(pos, ids, qn)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    tc = time.perf_counter()
    ref_order = O.direct_rows(pos, inp.kind, (box,) * 3, idx[sel], inp.atom_ff[:, sel],
                              num_molecules=n_mol, mode=O.DIRECT_REFERENCE)
    tcpu = time.perf_counter() - tc
    ext = O.direct_rows(pos, inp.kind, (box,) * 3, idx[sel], inp.atom_ff[:, sel],
                        num_molecules=n_mol, mode=O.DIRECT_EXTENDED)
    i_ext = ext[:, 0] ** 2 + ext[:, 1] ** 2
    for prec in ("fp32", "fp64"):
        a = res[prec][0][sel]
        out[prec]["vs_oracle_sampled"] = {
            "intensity_max_rel": float(np.max(np.abs(a[:, 0] ** 2 + a[:, 1] ** 2 - i_ext) / i_ext)),
            "self_max_rel": float(np.max(np.abs(a[:, 6] - ext[:, 6]) / ext[:, 6]))}
    out["cpu_baseline"] = {"value": len(pos) * len(sel) / tcpu, "unit": "atom-vector terms/s",
                           "cores": threads, "kind": "port",
                           "sample": f"oracle reference-order FP64 loop (libm cos/sin) over all "
                                     f"{len(pos)} atoms for {len(sel)} evenly spaced q-vectors, "
                                     f"{tcpu:.1f} s"}
    out["speedup_fp32_vs_cpu"] = out["fp32"]["terms_per_s"] / out["cpu_baseline"]["value"]
    log(f"direct oracle/cpu: {out['cpu_baseline']}; fp32 {out['fp32']['vs_oracle_sampled']}")
    print(json.dumps(out), flush=True)


WORKLOADS = {"vacuum1k": bench_vacuum1k, "helix60k": bench_helix60k, "batch512": bench_batch512,
             "frames": bench_frames,
             "served": bench_served, "direct": bench_direct}


if __name__ == "__main__":
    main()
