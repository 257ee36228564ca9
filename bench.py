#!/usr/bin/env python3
"""Benchmark of the MI355X DEAC engine: genome evaluations per second.

Headline workload (one GPU): the seed batch -- 256 independent
reconstructions (seeds 1..256) of the synthetic three-peak ISF with 65
timeslices (relative error 1e-4), genome 4096 per sign, population 32,
omega_max 512, standard model, bfull, --normalize, per-instance stop at
chi^2 <= 1.0, all in one persistent launch (deac_batch_run).  The batch
computes in plain FP64 in the reference's order: every instance is
bit-identical to the reference CPU build for its seed.

A step is one generation of every instance (256 x 32 genome evaluations: trial
vector, model ISF over 65 timeslices and 8192 frequencies, chi^2, selection).
`value` is genome-evals/s over the timed generations, from HIP events on the
stream the batch kernel runs on, with the problems and initial populations
already resident in HBM.  The plain single-instance FP64 path (G=4096, P=8,
T=65, seed 1) is timed beside it, and its ratio reported.

`roofline` prices the batch kernel against MI355X FP64 matrix peak
(4*G*T flops per genome evaluation), with the HBM traffic and the matrix
pipe's busy fraction from the committed rocprofv3 summaries under profiles/;
`cpu_baseline` times the reference CPU
executable built from source (oracle/_ref) on a bounded sample of the seed
batch, one process per seed on the host cores at once, as the reference is
run for many seeds.

`config.uneven_batch` times the uneven batch (tests/golden/make_uneven.py:
64 wavevectors x 16 seeds with mixed T, G, omega_max, model and
zero-temperature kernels, population 16) for a bounded number of generations
in one deac_batch_run, as instance-generations/s and sustained FP64 TFLOP/s.

`config.precision_trajectories` runs the T=257, G=16384, P=64 hyperbolic
workload for a bounded number of generations (--trajectory-gens) with K.S in exact
FP64 (the fused kernel) and in FP32 and split-BF16x3 (the split loop, chi^2
and selection in FP64), and reports each mixed run's final |dchi^2|/chi^2 and
DSF L2 error against the FP64 run.

`config.precision_sweep` evaluates K.S with deac_model_isf at FP64, FP32 and
split-BF16x3 (chi^2 in FP64): |dchi^2|/chi^2 against FP64 for 64 genomes
near the solution of the T=257, G=16384 hyperbolic workload (the true
spectrum on the genome grid, 0.1 % multiplicative noise), and the
contraction's throughput for a stack of 8192 genomes (256 seeds x 32
members) at G=4096, T=65.
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
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

METRIC = "genome-evals/sec at G=4096,T=65: batched mixed-precision vs plain FP64"
FP64_PEAK_TFLOPS = 78.6          # MI355X FP64 matrix (and vector), dense
HBM_PEAK_GBS = 8000.0

ISF = "isf_c2_T65.bin"
TEMPERATURE, G, OMEGA_MAX, T = 1.35, 4096, 512.0, 65
BATCH = dict(seeds=256, P=32, stop=1.0)
PLAIN = dict(P=8, seed=1)


def load_isf():
    import numpy as np
    return np.fromfile(os.path.join(ROOT, "tests", "golden", ISF)).reshape(3, -1)


def cpu_baseline(gens):
    """Reference CPU build on the seed-batch workload, used as the reference
    is used for many seeds: one process per seed (seeds 1..n), n = the
    host cores available (at most 16, the GPU's share of the box), all at
    once, `gens` generations each; aggregate genome-evals/s over the wall
    time.  Without the reference build, the oracle (one seed, one core)."""
    P = BATCH["P"]
    exe = os.path.join(ROOT, "oracle", "_ref", "deac_standard.e")
    isf = os.path.join(ROOT, "tests", "golden", ISF)
    if os.path.exists(exe):
        procs = min(BATCH["seeds"], len(os.sched_getaffinity(0)), 16)
        with tempfile.TemporaryDirectory() as d:
            cmds = [[exe, "-T", repr(TEMPERATURE), "-N", str(gens + 1), "-P", str(P), "-M", str(G),
                     "--omega_max", repr(OMEGA_MAX), "--normalize", "--spectra_type", "bfull",
                     "--stop_minimum_fitness", repr(BATCH["stop"]), "--seed", str(k), "--save_directory", d, isf]
                    for k in range(1, procs + 1)]
            t0 = time.perf_counter()
            running = [subprocess.Popen(c, stdout=subprocess.DEVNULL) for c in cmds]
            codes = [r.wait() for r in running]
            dt = time.perf_counter() - t0
        if any(codes):
            raise RuntimeError(f"reference CPU build failed: exit codes {codes}")
        return {"value": procs * P * gens / dt, "unit": "genome-evals/s", "cores": procs, "kind": "reference",
                "sample": f"seeds 1..{procs} of the seed-batch workload (G=4096, P=32, T=65), one reference "
                          f"process per seed on {procs} cores at once, {gens} generations each, wall time incl. "
                          f"setup {dt:.2f} s"}
    import oracle_binding as ob
    t, f, s = load_isf()
    p = ob.make_problem(t, f, s, ob.frequency_grid(G, OMEGA_MAX), TEMPERATURE, number_of_generations=gens + 1,
                        population_size=P, normalize=True, stop_minimum_fitness=BATCH["stop"])
    t0 = time.perf_counter()
    ob.run(p, 1)
    dt = time.perf_counter() - t0
    return {"value": P * gens / dt, "unit": "genome-evals/s", "cores": 1, "kind": "port",
            "sample": f"seed 1 of the seed-batch workload (G=4096, P=32, T=65), {gens} generations, one core, "
                      f"wall time incl. setup {dt:.2f} s"}


def traffic_per_generation():
    """HBM bytes per generation of the seed batch, from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a plain copy)."""
    import re
    path = os.path.join(ROOT, "profiles", "batch_kernel_seed256_hbm_counters.md")
    try:
        m = re.search(r"traffic_bytes_per_generation:\s*([0-9.]+)", open(path).read())
    except OSError:
        return None
    return float(m.group(1)) if m else None


def mfma_busy_profiled():
    """Fraction of cycles the FP64 matrix pipe was busy in the seed batch's
    kernel, from the committed rocprofv3 SQ-counter summary."""
    import re
    path = os.path.join(ROOT, "profiles", "batch_kernel_seed256_sq_counters.md")
    try:
        m = re.search(r"MFMA busy: .*= ([0-9.]+) %", open(path).read())
    except OSError:
        return None
    return float(m.group(1)) / 100.0 if m else None


def uneven_batch(da, gens):
    """The uneven batch for `gens` timed generations (after one warm-up run)."""
    import make_uneven as mu
    insts = mu.instances()
    probs = [da.make_problem(x["tau"], x["isf"], x["err"], da.frequency_grid(x["G"], x["omega_max"]), mu.TEMPERATURE,
                             model=x["model"], zero_temperature=x["zero_temperature"], population_size=x["P"],
                             normalize=True, spectra_type="bfull") for x in insts]
    seeds = [x["seed"] for x in insts]
    flop_per_gen = sum(x["P"] * 2 * (1 if x["zero_temperature"] else 2) * x["G"] * x["T"] for x in insts)
    with da.Batch(probs) as b:
        b.run(seeds, 3, 0.0, track_stats=False)
        t0 = time.perf_counter()
        b.run(seeds, gens + 1, 0.0, track_stats=False)
        wall = time.perf_counter() - t0
        prof = b.profile
    loop_s = prof["loop_ms"] / 1e3
    inst_gens = prof["generations"]
    return {"workload": f"{len(insts)} instances (64 wavevectors x 16 seeds), T in {{33,65,129,257}}, "
                        "G in {1024..8192}, omega_max in [128,512], standard/normalization/zero-T, P=16, "
                        f"{gens} generations each, one deac_batch_run",
            "instance_generations_per_s": inst_gens / loop_s,
            "genome_evals_per_s": inst_gens * mu.POPULATION / loop_s,
            "fp64_tflops": flop_per_gen * gens / loop_s / 1e12,
            "ms": prof["loop_ms"], "wall_s_including_host_setup": wall}


def precision_sweep(da):
    import numpy as np
    import make_inputs as mi
    out = {}
    # near-solution genomes of the T=257, G=16384 hyperbolic workload:
    # genome = 2 S(w) e^{beta w / 2} on both signs reproduces the ISF
    tau, f, s = np.fromfile(os.path.join(ROOT, "tests", "golden", "isf_c3_T257.bin")).reshape(3, -1)
    G3, beta = 16384, 1.0 / TEMPERATURE
    w = da.frequency_grid(G3, OMEGA_MAX)
    grid = np.linspace(0.0, OMEGA_MAX, 65536)
    norm = np.trapezoid(sum(wt * mi.gaussian(grid, mu, sd) for mu, sd, wt in mi.CONFIG2_PEAKS), grid)
    S = sum(wt * mi.gaussian(w, mu, sd) for mu, sd, wt in mi.CONFIG2_PEAKS) / norm
    g0 = np.concatenate([2.0 * S * np.exp(0.5 * beta * w)] * 2)
    pop = g0 * (1.0 + 1e-3 * np.random.default_rng(3).standard_normal((64, 2 * G3)))
    prob = da.make_problem(tau, f, s, w, TEMPERATURE, model="hyperbolic", spectra_type="bfull", population_size=64,
                           normalize=True)
    E64, c64, _ = da.model_isf(prob, pop, "fp64")
    near = {"chi2_fp64_median": float(np.median(c64))}
    for prec in ("fp32", "bf16x3"):
        E, c, _ = da.model_isf(prob, pop, prec)
        rel = np.abs(c - c64) / c64
        near[prec] = {"dchi2_rel_median": float(np.median(rel)), "dchi2_rel_max": float(np.max(rel)),
                      "model_isf_rel_max": float(np.max(np.abs(E - E64) / np.abs(E64)))}
    out["near_solution_T257_G16384_hyperbolic"] = near
    # stacked contraction at G=4096, T=65: 8192 genomes
    t2, f2, s2 = load_isf()
    p2 = da.make_problem(t2, f2, s2, da.frequency_grid(G, OMEGA_MAX), TEMPERATURE, population_size=32, normalize=True,
                         spectra_type="bfull")
    genomes = np.random.default_rng(4).uniform(0.0, 1.0, (8192, 2 * G)) / G
    flop = 2.0 * T * (2 * G) * genomes.shape[0]
    thr = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        da.model_isf(p2, genomes, prec)
        _, _, ms = da.model_isf(p2, genomes, prec)
        thr[prec] = {"ms": ms, "tflops": flop / (ms / 1e3) / 1e12}
    out["stacked_8192_genomes_G4096_T65"] = thr
    return out


def precision_trajectories(da, gens):
    """Precision trajectories: the T=257, G=16384, P=64 hyperbolic workload run
    for `gens` generations (seed 1) with K.S in exact FP64 (the fused kernel,
    the reference's trajectory) and in FP32 / BF16x3 (the split loop, chi^2
    and selection in FP64).  Each mixed run reports its final chi^2, that of
    its DSF re-evaluated in FP64, |dchi^2|/chi^2 and the DSF L2 error against
    the FP64 run, the first generation at which its best-fitness history
    leaves the FP64 run's, and its time per generation."""
    import numpy as np
    tau, f, s = np.fromfile(os.path.join(ROOT, "tests", "golden", "isf_c3_T257.bin")).reshape(3, -1)
    G3, beta = 16384, 1.0 / TEMPERATURE
    w = da.frequency_grid(G3, OMEGA_MAX)
    prob = da.make_problem(tau, f, s, w, TEMPERATURE, model="hyperbolic", spectra_type="bfull", population_size=64)
    out = {"workload": f"T=257, G=16384, P=64, omega_max 512, hyperbolic, bfull, seed 1, {gens} generations",
           "runs": {}}
    runs = {}
    with da.Engine(prob) as e:
        for prec in ("fp64", "fp32", "bf16x3"):
            r = e.run(1, gens + 1, 0.0, track_stats=True, precision=prec)
            runs[prec] = r
            out["runs"][prec] = {"chi2": r["minimum_fitness"],
                                 "ms_per_generation": r["profile"]["loop_ms"] / r["loop_generations"],
                                 "path": "fused kernel, exact" if prec == "fp64" else "split loop"}
    ref = runs["fp64"]
    # hyperbolic bfull DSF: 0.5 S+ e^{beta w / 2} at G-1+i, 0.5 S- e^{-beta w / 2} at G-1-i
    for prec in ("fp32", "bf16x3"):
        r = runs[prec]
        dsf = r["dsf"]
        genome = np.concatenate([2.0 * dsf[G3 - 1:] / np.exp(0.5 * beta * w),
                                 2.0 * dsf[G3 - 1::-1] / np.exp(-0.5 * beta * w)])
        _, chi64, _ = da.model_isf(prob, genome, "fp64")
        out["runs"][prec].update(
            chi2_fp64_reevaluated=float(chi64[0]),
            dchi2_rel=abs(r["minimum_fitness"] - ref["minimum_fitness"]) / ref["minimum_fitness"],
            dchi2_rel_fp64_reevaluated=abs(float(chi64[0]) - ref["minimum_fitness"]) / ref["minimum_fitness"],
            dsf_l2_rel=float(np.linalg.norm(dsf - ref["dsf"]) / np.linalg.norm(ref["dsf"])))
        # first generation whose best fitness departs from the FP64 run's by
        # more than 1e-4 relative: 100x the contraction's own error, so a
        # different selection and not the evaluation alone
        dev = np.abs(r["minimum"] - ref["minimum"]) / ref["minimum"]
        far = np.nonzero(dev > 1e-4)[0]
        out["runs"][prec]["first_divergent_generation"] = int(far[0]) if len(far) else None
        out["runs"][prec]["best_fitness_rel_dev_generation_0"] = float(dev[0])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000, help="timed generations of the seed batch")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--plain-steps", type=int, default=5000, help="timed generations of the plain path")
    ap.add_argument("--cpu-gens", type=int, default=150, help="generations per seed of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--uneven-gens", type=int, default=200, help="timed generations of the uneven batch (0: skip)")
    ap.add_argument("--no-precision-sweep", action="store_true")
    ap.add_argument("--trajectory-gens", type=int, default=2000,
                    help="generations of the T=257, G=16384 precision trajectories (0: skip)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device copy of this many bytes (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: instances are independent (across GPUs they would shard by "
              "instance, with no communication per generation) and nothing here spans devices; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_bytes:
        # Reference kernel for the HBM byte counters: reads and writes exactly
        # copy_bytes (three launches).
        import torch
        x = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes, "launches": 3}), flush=True)
        return

    import deac_amd as da
    t, f, s = load_isf()
    freq = da.frequency_grid(G, OMEGA_MAX)
    flop_per_eval = 4 * G * T

    # ---- seed batch (headline)
    nb, Pb = BATCH["seeds"], BATCH["P"]
    prob = da.make_problem(t, f, s, freq, TEMPERATURE, population_size=Pb, normalize=True, spectra_type="bfull")
    seeds = list(range(1, nb + 1))
    with da.Batch([prob] * nb) as b:
        if args.warmup > 0:
            b.run(seeds, args.warmup + 1, BATCH["stop"], track_stats=False)
        t0 = time.perf_counter()
        outs = b.run(seeds, args.steps + 1, BATCH["stop"], track_stats=False)
        wall = time.perf_counter() - t0
        prof = b.profile
    loop_s = prof["loop_ms"] / 1e3
    evals = Pb * prof["generations"]            # genome evaluations actually performed
    value = evals / loop_s
    kern_s = prof["evolve_ms"] / 1e3
    achieved = evals * flop_per_eval / kern_s / 1e12
    stopped = sum(1 for o in outs if o["generation"] < args.steps)
    per_gen = traffic_per_generation()
    gens_per_launch = min(args.steps, 4096)
    traffic = per_gen * gens_per_launch if per_gen else None
    alg_launch = nb * Pb * 2 * G * 8 * gens_per_launch

    # ---- plain FP64 single instance
    pp = da.make_problem(t, f, s, freq, TEMPERATURE, population_size=PLAIN["P"], normalize=True, spectra_type="bfull")
    with da.Engine(pp) as e:
        e.run(PLAIN["seed"], 201, 0.0, track_stats=False)
        pout = e.run(PLAIN["seed"], args.plain_steps + 1, 0.0, track_stats=False)
    pprof = pout["profile"]
    plain_value = PLAIN["P"] * pprof["generations"] / (pprof["loop_ms"] / 1e3)
    # the plain path on the batch's own shape: seeds as sequential
    # single-instance runs (P=32), timed on seed 1
    pq = da.make_problem(t, f, s, freq, TEMPERATURE, population_size=Pb, normalize=True, spectra_type="bfull")
    with da.Engine(pq) as e:
        e.run(1, 101, 0.0, track_stats=False)
        qout = e.run(1, args.plain_steps // 4 + 1, 0.0, track_stats=False)
    qprof = qout["profile"]
    plain32_value = Pb * qprof["generations"] / (qprof["loop_ms"] / 1e3)
    uneven = uneven_batch(da, args.uneven_gens) if args.uneven_gens > 0 else None
    sweep = None if args.no_precision_sweep else precision_sweep(da)
    trajectories = precision_trajectories(da, args.trajectory_gens) if args.trajectory_gens > 0 else None

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
        "config": {"workload": "seed batch: 256 seeds x (G=4096 per sign, T=65, P=32, standard model, bfull, "
                               "--normalize, stop at chi^2 <= 1), one persistent launch, synthetic three-peak "
                               "ISF with rel. error 1e-4",
                   "precision": "fp64 in the reference's order on FP64 MFMA: bit-identical per seed to the "
                                "reference CPU build",
                   "generations_per_s": prof["generations"] / nb / loop_s,
                   "instances_stopped": stopped,
                   "wall_s_including_host_setup": wall,
                   "plain_fp64": {"value": plain_value, "unit": "genome-evals/s",
                                  "workload": "single instance G=4096, T=65, P=8, seed 1",
                                  "ms_per_generation": pprof["loop_ms"] / pprof["generations"]},
                   "batched_over_plain": value / plain_value,
                   "plain_fp64_batch_shape": {"value": plain32_value, "unit": "genome-evals/s",
                                              "workload": "the seed-batch instance shape (G=4096, T=65, P=32) "
                                                          "run seed by seed through the plain path, seed 1",
                                              "ms_per_generation": qprof["loop_ms"] / qprof["generations"]},
                   "batched_over_plain_batch_shape": value / plain32_value,
                   "uneven_batch": uneven,
                   "precision_sweep": sweep,
                   "precision_trajectories": trajectories},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "traffic_note": "HBM bytes per launch (one launch = the timed generations, up to the "
                                     "4096-generation slice) from profiles/batch_kernel_seed256_hbm_counters.md; "
                                     "algorithmic bytes per launch: " + str(alg_launch),
                     "kernel": "batch_kernel<1,2,1>", "flop_per_genome_eval": flop_per_eval,
                     "kernel_ms_per_generation": prof["evolve_ms"] / args.steps,
                     "hbm_algorithmic_GBps": evals * 2 * G * 8 / kern_s / 1e9,
                     "mfma_busy": mfma_busy_profiled(),
                     "mfma_busy_note": "SQ_VALU_MFMA_BUSY_CYCLES over the dispatch's cycles per SIMD, from "
                                       "profiles/batch_kernel_seed256_sq_counters.md"},
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_gens)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
