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
  * roofline of the pair kernel: algorithmic FLOPs (14 per pair-q term with
    solvent) over its HIP-event time, against the FP32 vector peak;
  * the plain FP64 GPU path on the same structure (its pair kernel time);
  * cpu_baseline: the oracle's reference-order FP64 loop (3 self + 3 cross
    sums per q, libm sin, OpenMP over q) on k evenly spaced q points.
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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=10000)
    ap.add_argument("--nq", type=int, default=500)
    ap.add_argument("--q-per-lane", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--span", type=int, default=0)
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--cpu-q", type=int, default=8, help="q points of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this build "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import pripps_amd
    from pripps_amd import _native
    from pripps_amd.synth import solvent10k

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
        m64t = eng.build(inp, "fp64")
        fp64_prof = m64t.profiles()
        pm, _ = m64t.recompute(2)
        extra["fp64_tiled_pair_kernel_ms"] = pm
        extra["fp64_tiled_vs_plain_maxerr"] = envelope_error(fp64_prof, plain_prof)
        log(f"fp64 tiled pair kernel {pm:.2f} ms; tiled vs plain {extra['fp64_tiled_vs_plain_maxerr']:.2e}")
        del m64, m64t

    model = eng.build(inp, "fp32", tune)
    st = model.stats()
    prof32 = model.profiles()
    log(f"fp32 build: Q={st.q_per_lane} block={st.block} span={st.span} items={st.work_items} "
        f"kernel {st.kernel_ms:.2f} ms")
    if args.warmup > 0:
        model.recompute(args.warmup)
    t0 = time.perf_counter()
    pair_ms, dev_ms = model.recompute(args.steps)
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / max(args.steps, 1)
    value = terms / (ms_per_step * 1e-3)
    again = model.profiles()
    assert np.array_equal(again, prof32), "recompute changed the profiles"
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
                   "block": st.block, "span": st.span, "work_items": st.work_items},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "debye_tiled_kernel<float,%d,true>" % st.q_per_lane,
                     "kernel_ms": pair_ms, "flop_per_term": FLOP_PER_TERM_SOLVENT},
        "cpu_baseline": cpu,
        "max_rel_err_vs_fp64_gpu": err_fp64,
        "max_rel_err_vs_oracle_sampled_q": err_oracle,
        "extra": extra,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
