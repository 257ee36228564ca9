#!/usr/bin/env python3
"""Benchmark of the MI355X DEAC engine: genome evaluations per second.

Headline workload (one GPU): the seed batch -- 256 independent
reconstructions (seeds 1..256) of the synthetic three-peak ISF with 65
timeslices (relative error 1e-4), genome 4096 per sign, population 32,
omega_max 512, standard model, bfull, --normalize, per-instance stop at
chi^2 <= 1.0, all in one persistent launch (deac_batch_run), with the
model ISF's K.S on the FP32 matrix cores in the delta form (each member's
model ISF kept in FP64, a trial's = that + K.(trial - parent) in FP32;
chi^2, selection and everything else FP64; an exact FP64 re-evaluation of
the population every 1024 generations and at the end, so every reported
chi^2 is exact for its genome).

A step is one generation of every instance (256 x 32 genome evaluations: trial
vector, model ISF over 65 timeslices and 8192 frequencies, chi^2, selection).
The warm-up runs W generations; the K timed steps continue those same runs
(deac_batch_continue), so they are generations W+1..W+K of each run, the
segment's closing FP64 re-evaluation included.  `value` is genome-evals/s
over the timed generations, from HIP events on the stream the batch kernel
runs on, with everything resident in HBM.  Beside it: the same batch at
BF16x3 and at exact FP64 (bit-identical to the reference CPU build per seed),
a 400-generation steady-state figure for each, the plain single-instance
FP64 path (P=8, and P=32 run seed by seed) and the ratios that split the
speed-up into batching and precision.

`roofline` prices the FP32 batch kernel against HBM (its 16.25 flop/B lies
below the FP32-MFMA ridge): algorithmic bytes per launch (65536 per genome
evaluation) over the kernel's average launch duration from HIP events, with
the measured HBM traffic from the committed rocprofv3 summary;
`roofline_fp64_batch` prices the exact FP64 kernel against FP64 MFMA.
`cpu_baseline` times the reference CPU executable built from source
(oracle/_ref) on a bounded sample of the seed batch, one process per seed
on the host cores at once, as the reference is run for many seeds.

`config.batch_precision_accuracy` runs the seed batch at FP64, FP32 and
BF16x3 with statistics and reports each mixed precision's first divergent
generation against FP64, its final chi^2 and the exactness of the reported
chi^2.  `config.uneven_batch` times the uneven batch (tests/golden/
make_uneven.py: 64 wavevectors x 16 seeds with mixed T, G, omega_max, model
and zero-temperature kernels, population 16).  `config.precision_trajectories`
and `config.precision_sweep` cover the single-instance T=257, G=16384
hyperbolic workload (split loop, deac_model_isf).
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
FP32_PEAK_TFLOPS = 157.3         # MI355X FP32 matrix, dense
HBM_PEAK_GBS = 8000.0
HEADLINE = "fp32"                # the seed batch's K.S precision (FP32-MFMA)
HEADLINE_KERNEL = "batch_kernel<1,3,2,0> (FP32 K.S, 3 tile pairs per chain wave, 2 member tiles)"
HEADLINE_PROFILE = "profiles/batch_kernel_fp32_seed256_hbm_counters.md"

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
                "host": host_cpu(),
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


def host_cpu():
    """CPU model and core counts of the host (lscpu's model name, nproc)."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}


def traffic_per_generation(path):
    """HBM bytes per generation of the seed batch, from a committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (calibrated against a plain copy)."""
    import re
    try:
        m = re.search(r"traffic_bytes_per_generation:\s*([0-9.]+)", open(os.path.join(ROOT, path)).read())
    except OSError:
        return None
    return float(m.group(1)) if m else None


def batch_precision_accuracy(da, prob, gens):
    """The seed batch at FP64, FP32 and BF16x3 for `gens` generations with
    statistics: per mixed precision, the first generation at which each
    seed's best-fitness history departs from the FP64 batch's by more than
    1e-4 relative (100x the contraction's error: a different selection, not
    the evaluation), the final chi^2 against FP64's, and the reported chi^2
    against an exact FP64 evaluation of the reported genome (deac_model_isf)."""
    import numpy as np
    seeds = list(range(1, BATCH["seeds"] + 1))
    out = {"workload": f"the seed batch, {gens} generations, stop at chi^2 <= {BATCH['stop']}"}
    with da.Batch([prob] * len(seeds)) as b:
        runs = {prec: b.run(seeds, gens + 1, BATCH["stop"], True, precision=prec)
                for prec in ("fp64", "fp32", "bf16x3")}
    ref = runs["fp64"]
    for prec in ("fp32", "bf16x3"):
        o = runs[prec]
        first = []
        for a, r in zip(o, ref):
            n = min(len(a["minimum"]), len(r["minimum"]))
            far = np.nonzero(np.abs(a["minimum"][:n] - r["minimum"][:n]) / r["minimum"][:n] > 1e-4)[0]
            first.append(int(far[0]) if len(far) else None)
        div = [x for x in first if x is not None]
        fin = np.array([a["minimum_fitness"] for a in o])
        finr = np.array([r["minimum_fitness"] for r in ref])
        exact = []
        for k in (0, 63, 127, 255):
            dsf = o[k]["dsf"]
            g = np.concatenate([dsf[G - 1:], dsf[G - 1::-1]])
            _, chi, _ = da.model_isf(prob, g, "fp64")
            exact.append(abs(chi[0] - o[k]["minimum_fitness"]) / chi[0])
        # DSF L2 error per seed, and of the seed-averaged DSF against the
        # FP64 batch's average (which is the plain path's: bit-identical per
        # seed), where the spectrum has weight (0 <= omega <= 100: the far
        # negative-frequency genes do not enter chi^2 and drift without bound)
        sel = (ref[0]["frequency"] >= 0.0) & (ref[0]["frequency"] <= 100.0)
        rel = lambda a, r: float(np.linalg.norm(a[sel] - r[sel]) / np.linalg.norm(r[sel]))
        l2 = np.array([rel(a["dsf"], r["dsf"]) for a, r in zip(o, ref)])
        avg, avgr = np.mean([a["dsf"] for a in o], axis=0), np.mean([r["dsf"] for r in ref], axis=0)
        half = len(ref) // 2
        halves = rel(np.mean([r["dsf"] for r in ref[:half]], axis=0), np.mean([r["dsf"] for r in ref[half:]], axis=0))
        out[prec] = {"first_divergent_generation_median": float(np.median(div)) if div else None,
                     "seeds_not_diverged": sum(1 for x in first if x is None),
                     "final_chi2_median": float(np.median(fin)), "final_chi2_median_fp64": float(np.median(finr)),
                     "final_chi2_rel_diff_median": float(np.median(np.abs(fin - finr) / finr)),
                     "final_chi2_rel_diff_max": float(np.max(np.abs(fin - finr) / finr)),
                     "dsf_l2_rel_w0_100_median": float(np.median(l2)), "dsf_l2_rel_w0_100_max": float(np.max(l2)),
                     "seed_average_dsf_l2_rel_w0_100": rel(avg, avgr),
                     "fp64_seed_halves_l2_rel_w0_100": halves,
                     "reported_vs_fp64_reevaluated_rel_max": float(max(exact)),
                     "generation1_best_rel_diff_max": float(max(abs(a["minimum"][1] - r["minimum"][1]) /
                                                                r["minimum"][1] for a, r in zip(o, ref)))}
    return out


def uneven_batch(da, gens):
    """The uneven batch for `gens` timed generations (after one warm-up run):
    with the closed-form launch schedule, then with the schedule the
    autotuner (deac_batch_tune, 40 generations per candidate) picks, at FP64
    and at BF16x3 K.S."""
    import make_uneven as mu
    insts = mu.instances()
    probs = [da.make_problem(x["tau"], x["isf"], x["err"], da.frequency_grid(x["G"], x["omega_max"]), mu.TEMPERATURE,
                             model=x["model"], zero_temperature=x["zero_temperature"], population_size=x["P"],
                             normalize=True, spectra_type="bfull") for x in insts]
    seeds = [x["seed"] for x in insts]
    flop_per_gen = sum(x["P"] * 2 * (1 if x["zero_temperature"] else 2) * x["G"] * x["T"] for x in insts)

    def timed(b, prec="fp64"):
        b.run(seeds, 3, 0.0, track_stats=False, precision=prec)
        t0 = time.perf_counter()
        b.run(seeds, gens + 1, 0.0, track_stats=False, precision=prec)
        wall = time.perf_counter() - t0
        prof = b.profile
        loop_s = prof["loop_ms"] / 1e3
        return {"instance_generations_per_s": prof["generations"] / loop_s,
                "genome_evals_per_s": prof["generations"] * mu.POPULATION / loop_s,
                "fp64_equivalent_tflops": flop_per_gen * gens / loop_s / 1e12,
                "ms": prof["loop_ms"], "wall_s_including_host_setup": wall}

    with da.Batch(probs) as b:
        b.schedule = {"group_streams": 3, "group_order": 2, "instance_order": 0}  # the closed-form rule
        rule = timed(b)
        t0 = time.perf_counter()
        chosen = b.tune(seeds, 40)
        tune_s = time.perf_counter() - t0
        tuned = timed(b)
        mixed = timed(b, "bf16x3")
    out = {"workload": f"{len(insts)} instances (64 wavevectors x 16 seeds), T in {{33,65,129,257}}, "
                       "G in {1024..8192}, omega_max in [128,512], standard/normalization/zero-T, P=16, "
                       f"{gens} generations each, one deac_batch_run",
           "closed_form_schedule": rule, "tuned_schedule": chosen, "tuning_s": tune_s, "tuned": tuned,
           "tuned_bf16x3": mixed, "tuned_over_closed_form": tuned["genome_evals_per_s"] / rule["genome_evals_per_s"]}
    out.update(tuned)  # the headline uneven figures: FP64 with the tuned schedule
    return out


def small_population_batch(da, gens):
    """Seeds with small populations: 512 seeds of the G=4096, T=65 workload at
    P=8, exact FP64, for `gens` timed generations (after an 8-generation
    warm-up run), one instance per workgroup and with the automatic stacking
    (two instances in the 16 member rows of each workgroup)."""
    t, f, s = load_isf()
    n, P = 512, 8
    prob = da.make_problem(t, f, s, da.frequency_grid(G, OMEGA_MAX), TEMPERATURE, population_size=P, normalize=True,
                           spectra_type="bfull")
    seeds = list(range(1, n + 1))
    out = {"workload": f"{n} seeds x (G=4096 per sign, T=65, P=8, bfull, --normalize), FP64, {gens} generations"}
    with da.Batch([prob] * n) as b:
        for name, pack in (("one_per_workgroup", 1), ("stacked", 0)):
            sc = b.schedule
            sc["pack"] = pack
            b.schedule = sc
            b.run(seeds, 8, 0.0, track_stats=False)
            b.run(seeds, gens + 1, 0.0, track_stats=False)
            prof = b.profile
            out[name] = {"value": P * prof["generations"] / (prof["loop_ms"] / 1e3), "unit": "genome-evals/s",
                         "workgroups_per_launch": prof["evolve_workgroups"] / max(prof["evolve_launches"], 1),
                         "ms_per_generation": prof["loop_ms"] / gens}
    out["stacked_over_one_per_workgroup"] = out["stacked"]["value"] / out["one_per_workgroup"]["value"]
    return out


def wide_population_batch(da, gens):
    """Seeds with 64-member populations (four member tiles per workgroup):
    256 seeds of the G=4096, T=65 workload at P=64 for `gens` timed
    generations after a 100-generation run, at FP64, FP32 and BF16x3 K.S;
    kernel time of the continued launch."""
    t, f, s = load_isf()
    n, P = 256, 64
    prob = da.make_problem(t, f, s, da.frequency_grid(G, OMEGA_MAX), TEMPERATURE, population_size=P, normalize=True,
                           spectra_type="bfull")
    seeds = list(range(1, n + 1))
    out = {"workload": f"{n} seeds x (G=4096 per sign, T=65, P=64, bfull, --normalize), {gens} generations"}
    with da.Batch([prob] * n) as b:
        for prec in ("fp64", "fp32", "bf16x3"):
            b.run(seeds, 101, 0.0, track_stats=False, precision=prec)
            b.continue_run(gens + 1, 0.0, track_stats=False, precision=prec)
            prof = b.profile
            ms = prof["evolve_ms"] / gens
            out[prec] = {"kernel_ms_per_generation": ms, "genome_evals_per_s": P * n / ms * 1e3}
    out["fp32_over_fp64"] = out["fp32"]["genome_evals_per_s"] / out["fp64"]["genome_evals_per_s"]
    return out


def plain_default_population(da, gens):
    """The plain path at the reference CLI's default population (-P 512,
    deac.cpp:1935-1938): one instance of the G=4096, T=65 workload, exact
    FP64, `gens` timed generations after 20."""
    t, f, s = load_isf()
    P = 512
    prob = da.make_problem(t, f, s, da.frequency_grid(G, OMEGA_MAX), TEMPERATURE, population_size=P, normalize=True,
                           spectra_type="bfull")
    with da.Engine(prob) as e:
        e.run(1, 21, 0.0, track_stats=False)
        o = e.run(1, gens + 1, 0.0, track_stats=False)
    ms = o["profile"]["loop_ms"] / o["profile"]["generations"]
    return {"workload": f"single instance G=4096, T=65, P={P}, seed 1, {gens} generations",
            "ms_per_generation": ms, "genome_evals_per_s": P / ms * 1e3}


def detailed_balance_batch(da, gens):
    """The detailed-balance variant (bdsf, with the -1st-moment term) on the
    seed-batch shape: 256 seeds of G=4096, T=65, P=32 for `gens` generations
    on the batch's fold kernel, against one seed on the plain path."""
    t, f, s = load_isf()
    prob = da.make_problem(t, f, s, da.frequency_grid(G, OMEGA_MAX), TEMPERATURE, population_size=32,
                           normalize=True, spectra_type="bdsf", use_negative_first_moment=True)
    n = 256
    seeds = list(range(1, n + 1))
    with da.Batch([prob] * n) as b:
        b.run(seeds, 11, 0.0, track_stats=False)
        b.continue_run(gens + 1, 0.0, track_stats=False)
        prof = b.profile
    batch_ms = prof["loop_ms"] / gens
    with da.Engine(prob) as e:
        e.run(1, 11, 0.0, track_stats=False)
        out = e.run(1, gens + 1, 0.0, track_stats=False)
    plain_ms = out["profile"]["loop_ms"] / out["profile"]["generations"]
    return {"workload": f"{n} seeds x (G=4096, T=65, P=32, bdsf, --normalize, -1st moment), FP64, {gens} generations",
            "batch_ms_per_generation": batch_ms, "batch_genome_evals_per_s": n * 32 / (batch_ms / 1e3),
            "plain_ms_per_generation_one_seed": plain_ms, "batch_over_seed_by_seed": n * plain_ms / batch_ms}


def role_cycles(da, args):
    """Per-wavefront cycles of the batch kernels on the seed batch, one JSON
    line per precision: the first `--warmup` generations of a run, then
    `--steps` continued generations (profiling build only)."""
    t, f, s = load_isf()
    prob = da.make_problem(t, f, s, da.frequency_grid(G, OMEGA_MAX), TEMPERATURE, population_size=BATCH["P"],
                           normalize=True, spectra_type="bfull")
    seeds = list(range(1, BATCH["seeds"] + 1))
    with da.Batch([prob] * len(seeds)) as b:
        for prec in args.batch_precisions.split(","):
            da.role_cycles(prec, reset=True)
            b.run(seeds, args.warmup + 1, BATCH["stop"], track_stats=False, precision=prec)
            early = da.role_cycles(prec, reset=True)
            b.continue_run(args.steps + 1, BATCH["stop"], track_stats=False, precision=prec)
            late = da.role_cycles(prec, reset=True)
            print(json.dumps({"precision": prec, "generations_1_to": args.warmup, "early": early,
                              "continued_generations": args.steps, "steady": late}), flush=True)


def schedule_grid(da, gens):
    """One JSON line per launch schedule of the uneven batch (FP64, `gens`
    generations after a warm-up), then the autotuner's choice and its rate."""
    import make_uneven as mu
    insts = mu.instances()
    probs = [da.make_problem(x["tau"], x["isf"], x["err"], da.frequency_grid(x["G"], x["omega_max"]), mu.TEMPERATURE,
                             model=x["model"], zero_temperature=x["zero_temperature"], population_size=x["P"],
                             normalize=True, spectra_type="bfull") for x in insts]
    seeds = [x["seed"] for x in insts]
    with da.Batch(probs) as b:
        b.run(seeds, 3, 0.0, track_stats=False)
        for st in (1, 2, 3):
            for go in (0, 1, 2):
                for io in (0, 1):
                    b.schedule = {"group_streams": st, "group_order": go, "instance_order": io}
                    b.run(seeds, gens + 1, 0.0, track_stats=False)
                    p = b.profile
                    print(json.dumps({"group_streams": st, "group_order": go, "instance_order": io,
                                      "ms_per_generation": p["loop_ms"] / gens,
                                      "genome_evals_per_s": p["generations"] * mu.POPULATION / (p["loop_ms"] / 1e3)}),
                          flush=True)
        chosen = b.tune(seeds, 40)
        b.run(seeds, gens + 1, 0.0, track_stats=False)
        p = b.profile
        print(json.dumps({"autotuned": chosen, "ms_per_generation": p["loop_ms"] / gens,
                          "genome_evals_per_s": p["generations"] * mu.POPULATION / (p["loop_ms"] / 1e3)}), flush=True)


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
    ap.add_argument("--plain-steps", type=int, default=5000, help="timed generations of the plain path (0: skip)")
    ap.add_argument("--batch-precisions", default="fp32,bf16x3,fp64",
                    help="K.S precisions of the seed batch to time, the headline (fp32) first")
    ap.add_argument("--cpu-gens", type=int, default=150, help="generations per seed of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--uneven-gens", type=int, default=200, help="timed generations of the uneven batch (0: skip)")
    ap.add_argument("--no-precision-sweep", action="store_true")
    ap.add_argument("--bdsf-gens", type=int, default=100,
                    help="timed generations of the detailed-balance seed batch (0: skip)")
    ap.add_argument("--small-pop-gens", type=int, default=300,
                    help="timed generations of the 512-seed P=8 batch, stacked and not (0: skip)")
    ap.add_argument("--wide-pop-gens", type=int, default=300,
                    help="timed generations of the 256-seed P=64 batch at each precision (0: skip)")
    ap.add_argument("--plain-p512-gens", type=int, default=200,
                    help="timed generations of the plain path at P=512 (0: skip)")
    ap.add_argument("--accuracy-gens", type=int, default=2000,
                    help="generations of the seed batch's FP64/FP32/BF16x3 accuracy comparison (0: skip)")
    ap.add_argument("--trajectory-gens", type=int, default=2000,
                    help="generations of the T=257, G=16384 precision trajectories (0: skip)")
    ap.add_argument("--role-cycles", action="store_true",
                    help="only report the batch kernels' per-wavefront cycles on the seed batch "
                         "(needs the profiling build: make -C deac-cpp_amd -B ROLE_TIMERS=1)")
    ap.add_argument("--schedule-grid", action="store_true",
                    help="only time the uneven batch under every launch schedule and the autotuner's choice")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device copy of this many bytes (HBM counter calibration) and exit")
    args = ap.parse_args()
    if HEADLINE not in args.batch_precisions.split(","):
        ap.error(f"--batch-precisions must include {HEADLINE}, the headline's K.S precision")
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
    if args.schedule_grid:
        schedule_grid(da, args.uneven_gens)
        return
    if args.role_cycles:
        role_cycles(da, args)
        return
    t, f, s = load_isf()
    freq = da.frequency_grid(G, OMEGA_MAX)
    flop_per_eval = 4 * G * T
    bytes_per_eval = 2 * G * 8

    # ---- seed batch (headline): W warm-up generations, then K timed
    # generations of the same runs (deac_batch_continue)
    nb, Pb = BATCH["seeds"], BATCH["P"]
    prob = da.make_problem(t, f, s, freq, TEMPERATURE, population_size=Pb, normalize=True, spectra_type="bfull")
    seeds = list(range(1, nb + 1))
    batch = {}
    with da.Batch([prob] * nb) as b:
        for prec in args.batch_precisions.split(","):
            b.run(seeds, max(args.warmup, 1) + 1, BATCH["stop"], track_stats=False, precision=prec)
            t0 = time.perf_counter()
            outs = b.continue_run(args.steps + 1, BATCH["stop"], track_stats=False, precision=prec)
            wall = time.perf_counter() - t0
            prof = dict(b.profile)
            # steady state: 400 more generations of the same runs
            b.continue_run(401, BATCH["stop"], track_stats=False, precision=prec)
            steady = dict(b.profile)
            evals = Pb * prof["generations"]
            batch[prec] = {
                "value": evals / (prof["loop_ms"] / 1e3), "unit": "genome-evals/s",
                "ms_per_step": prof["loop_ms"] / args.steps,
                "kernel_ms_per_launch": prof["evolve_ms"] / max(prof["evolve_launches"], 1),
                "evolve_launches": prof["evolve_launches"], "eval_ms": prof["eval_ms"],
                "eval_launches": prof["eval_launches"], "evals": evals,
                "steady_ms_per_generation": steady["loop_ms"] / (steady["generations"] / nb),
                "instances_stopped": sum(1 for o in outs if o["generation"] < args.steps),
                "wall_s_including_host": wall}
    head = batch[HEADLINE]
    value = head["value"]
    evals_per_launch = head["evals"] / max(head["evolve_launches"], 1)
    kern_s = head["kernel_ms_per_launch"] / 1e3
    per_gen = traffic_per_generation(HEADLINE_PROFILE)
    gens_per_launch = evals_per_launch / (nb * Pb)  # the timed segment is sliced into launches
    traffic = per_gen * gens_per_launch if per_gen else None
    fp64b = batch.get("fp64")

    # ---- plain FP64 single instance
    plain_value = plain32_value = float("nan")
    pprof = qprof = {"loop_ms": float("nan"), "generations": 1}
    if args.plain_steps > 0:
        pp = da.make_problem(t, f, s, freq, TEMPERATURE, population_size=PLAIN["P"], normalize=True,
                             spectra_type="bfull")
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
    accuracy = batch_precision_accuracy(da, prob, args.accuracy_gens) if args.accuracy_gens > 0 else None
    uneven = uneven_batch(da, args.uneven_gens) if args.uneven_gens > 0 else None
    small = small_population_batch(da, args.small_pop_gens) if args.small_pop_gens > 0 else None
    wide = wide_population_batch(da, args.wide_pop_gens) if args.wide_pop_gens > 0 else None
    plain512 = plain_default_population(da, args.plain_p512_gens) if args.plain_p512_gens > 0 else None
    bdsf = detailed_balance_batch(da, args.bdsf_gens) if args.bdsf_gens > 0 else None
    sweep = None if args.no_precision_sweep else precision_sweep(da)
    trajectories = precision_trajectories(da, args.trajectory_gens) if args.trajectory_gens > 0 else None

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "genome-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 K.S (delta form), f64 chi2/selection",
        "data": "synthetic",
        "config": {"workload": "seed batch: 256 seeds x (G=4096 per sign, T=65, P=32, standard model, bfull, "
                               "--normalize, stop at chi^2 <= 1), one persistent launch, synthetic three-peak "
                               "ISF with rel. error 1e-4; FP32-MFMA K.S in the delta form (E_parent FP64 + "
                               "K.(trial - parent)), chi^2 and selection FP64, exact FP64 re-evaluation every "
                               "1024 generations and at the end of a run",
                   "timed_region": f"generations {max(args.warmup, 1) + 1}..{max(args.warmup, 1) + args.steps} of "
                                   "the runs the warm-up started (deac_batch_continue), including the exact FP64 "
                                   "re-evaluation that ends the segment",
                   "generations_per_s": 1e3 / head["ms_per_step"],
                   "steady_state_ms_per_generation": head["steady_ms_per_generation"],
                   "steady_state_genome_evals_per_s": nb * Pb / (head["steady_ms_per_generation"] / 1e3),
                   "seed_batch": batch,
                   "plain_fp64": {"value": plain_value, "unit": "genome-evals/s",
                                  "workload": "single instance G=4096, T=65, P=8, seed 1",
                                  "ms_per_generation": pprof["loop_ms"] / pprof["generations"]},
                   "batched_over_plain": value / plain_value,
                   "plain_fp64_batch_shape": {"value": plain32_value, "unit": "genome-evals/s",
                                              "workload": "the seed-batch instance shape (G=4096, T=65, P=32) "
                                                          "run seed by seed through the plain path, seed 1",
                                              "ms_per_generation": qprof["loop_ms"] / qprof["generations"]},
                   "batched_over_plain_batch_shape": value / plain32_value,
                   "batching_vs_precision": {"fp64_batch_over_plain_batch_shape": fp64b["value"] / plain32_value,
                                             **{f"{p}_over_fp64_batch": batch[p]["value"] / fp64b["value"]
                                                for p in batch if p != "fp64"}} if fp64b else None,
                   "batch_precision_accuracy": accuracy,
                   "uneven_batch": uneven,
                   "small_population_batch": small,
                   "wide_population_batch": wide,
                   "plain_fp64_default_population": plain512,
                   "detailed_balance_batch": bdsf,
                   "precision_sweep": sweep,
                   "precision_trajectories": trajectories},
        "roofline": {"bound": "hbm", "achieved": evals_per_launch * bytes_per_eval / kern_s / 1e9,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": evals_per_launch * bytes_per_eval / kern_s / 1e9 / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "kernel": HEADLINE_KERNEL,
                     "bound_note": "FP32 K.S: 16.25 flop/B is below the FP32-MFMA ridge (157.3 TF / 8 TB/s = "
                                   "19.7 flop/B), so HBM bounds it; algorithmic bytes per genome evaluation "
                                   "2 G 8 = 65536 (one FP64 read of the genome pair)",
                     "algorithmic_bytes_per_launch": evals_per_launch * bytes_per_eval,
                     "kernel_ms_per_launch": head["kernel_ms_per_launch"],
                     "fp32_mfma_tflops": evals_per_launch * flop_per_eval / kern_s / 1e12,
                     "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE (calibrated "
                                     f"against a plain copy) in {HEADLINE_PROFILE}, per generation x the launch's "
                                     "generations"},
    }
    if fp64b:  # the exact FP64 batch kernel against FP64 MFMA, when it was timed
        fp64_tf = fp64b["evals"] / max(fp64b["evolve_launches"], 1) * flop_per_eval / (fp64b["kernel_ms_per_launch"] / 1e3) / 1e12
        line["roofline_fp64_batch"] = {"bound": "mfma", "kernel": "batch_kernel<0,1,2,1>", "achieved": fp64_tf,
                                       "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": fp64_tf / FP64_PEAK_TFLOPS}
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_gens)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
