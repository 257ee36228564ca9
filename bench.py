#!/usr/bin/env python3
"""Benchmark of the MI355X DEAC engine: genome evaluations per second.

Workload (one GPU): plain FP64 DEAC, single instance, standard model, the
synthetic three-peak ISF with 65 timeslices (relative error 1e-4), genome 4096
per sign, population 8, --normalize, bfull, seed 1.  A step is one generation
(every member evaluated once: trial vector, model ISF over 65 timeslices and
8192 frequencies, chi^2, selection), so genome-evals/step = P = 8.

Prints one JSON line.  `value` is genome-evals/s over exactly --steps
generations, from device events around the generation loop; the inputs are
resident in HBM when it starts (deac_run uploads them before the loop).
`roofline` prices the fused generation kernel against MI355X FP64 peak,
`cpu_baseline` times the reference CPU executable built from source
(oracle/_ref, one core) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "deac-cpp_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "genome-evals/sec at G=4096,T=65: batched mixed-precision vs plain FP64"
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 (vector and matrix), dense
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(isf="isf_c2_T65.bin", temperature=1.35, G=4096, omega_max=512.0, P=8, T=65, seed=1)


def load_isf(name):
    import numpy as np
    return np.fromfile(os.path.join(ROOT, "tests", "golden", name)).reshape(3, -1)


def cpu_baseline(gens):
    """Reference CPU build on the workload for `gens` generations (1 core)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "deac_standard.e")
    w = WORKLOAD
    with tempfile.TemporaryDirectory() as d:
        if os.path.exists(exe):
            cmd = [exe, "-T", repr(w["temperature"]), "-N", str(gens + 1), "-P", str(w["P"]), "-M", str(w["G"]),
                   "--omega_max", repr(w["omega_max"]), "--normalize", "--spectra_type", "bfull",
                   "--stop_minimum_fitness", "0", "--seed", str(w["seed"]), "--save_directory", d,
                   os.path.join(ROOT, "tests", "golden", w["isf"])]
            t0 = time.perf_counter()
            subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
            dt = time.perf_counter() - t0
            kind = "reference"
        else:
            import oracle_binding as ob
            t, f, s = load_isf(w["isf"])
            p = ob.make_problem(t, f, s, ob.frequency_grid(w["G"], w["omega_max"]), w["temperature"],
                                number_of_generations=gens + 1, population_size=w["P"], normalize=True,
                                stop_minimum_fitness=0.0)
            t0 = time.perf_counter()
            ob.run(p, w["seed"])
            dt = time.perf_counter() - t0
            kind = "port"
    return {"value": w["P"] * gens / dt, "unit": "genome-evals/s", "cores": 1, "kind": kind,
            "sample": f"{gens} generations of the same workload (G=4096, P=8, T=65), one process, "
                      f"wall time incl. setup {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--cpu-gens", type=int, default=2000, help="generations of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the engine runs one instance per device and nothing here spans "
              "devices (see DESIGN.md for how the work shards); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import deac_amd as da
    w = WORKLOAD
    t, f, s = load_isf(w["isf"])
    freq = da.frequency_grid(w["G"], w["omega_max"])
    prob = da.make_problem(t, f, s, freq, w["temperature"], population_size=w["P"], normalize=True,
                           spectra_type="bfull")
    with da.Engine(prob) as eng:
        if args.warmup > 0:
            eng.run(w["seed"], args.warmup + 1, 0.0, track_stats=False)
        t0 = time.perf_counter()
        out = eng.run(w["seed"], args.steps + 1, 0.0, track_stats=False)  # exactly --steps generations
        wall = time.perf_counter() - t0
    prof = out["profile"]
    assert prof["generations"] == args.steps
    loop_s = prof["loop_ms"] / 1e3
    value = w["P"] * args.steps / loop_s

    # dominant kernel: the fused generation kernel.  Algorithmic work per
    # genome evaluation: 4*G*T flops (2 signs x G x T multiply-adds) and 2*G*8
    # bytes (the member's FP64 genome pair read once per generation).
    flops_per_gen = w["P"] * 4 * w["G"] * w["T"]
    evo_s = prof["evolve_ms"] / 1e3
    achieved_tflops = flops_per_gen * args.steps / evo_s / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "genome-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": prof["loop_ms"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain FP64 single instance (configs[1]): standard model, bfull, --normalize, "
                               "G=4096 per sign, P=8, T=65, synthetic three-peak ISF, rel. error 1e-4, seed 1",
                   "generations_per_s": args.steps / loop_s, "precision": "fp64 exact (bit-identical to "
                   "the reference CPU build)", "wall_s_including_setup": wall},
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "evolve_kernel", "evolve_ms_per_generation": prof["evolve_ms"] / args.steps,
                     "walk_ms_per_generation": prof["walk_ms"] / args.steps,
                     "mask_ms_per_generation": prof["mask_ms"] / args.steps},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_gens)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
