#!/usr/bin/env python3
"""Golden vectors from the reference CPU executables.

Runs oracle/_ref/deac_<variant>.e (compiled from the reference sources by
oracle/Makefile; `make -C oracle ref`) on small inputs and stores, per case,
the exact float64 outputs: the three --track_stats histories (mean, minimum,
squared mean of the population fitness per generation), the best-member DSF
and the mirrored frequency grid, together with the run's arguments.

Only the reference's own outputs are stored -- never its sources.  The
committed .npz files are what the tests compare the oracle and the GPU
engine against; this script is only needed to regenerate them.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")

SMOKE = dict(tau=[0.0, 0.2, 0.4, 0.6], isf=[-1.0, -0.85, -0.72, -0.61], err=[0.05] * 4)


def smooth_isf(T, sign=1.0, beta=1.0 / 1.3):
    tau = np.linspace(0.0, 0.5 * beta, T)
    return dict(tau=tau.tolist(), isf=(sign * np.exp(-2.0 * tau)).tolist(), err=[0.03] * T)


# name: (variant, input, args).  args are deac.e flags; --track_stats and
# --seed are added by run_case.
CASES = {
    # the reference smoke-test fixture (test/deac_smoke_test.py:41-51) run longer
    "smoke_default": ("standard", SMOKE, dict(T=1.0, N=300, P=8, M=8, omega_max=4.0, seed=1, stop=0.0)),
    "smoke_normalize": ("standard", SMOKE, dict(T=1.0, N=300, P=8, M=8, omega_max=4.0, seed=2, stop=0.0, normalize=True)),
    "smoke_first_moment": ("standard", SMOKE, dict(T=1.0, N=300, P=8, M=8, omega_max=4.0, seed=3, stop=0.0, first_moment=0.5)),
    "smoke_third_moment": ("standard", SMOKE, dict(T=1.0, N=300, P=8, M=8, omega_max=4.0, seed=5, stop=0.0,
                                                   third_moment=1.5, third_moment_error=0.3)),
    "smoke_stop": ("standard", SMOKE, dict(T=1.0, N=5000, P=8, M=8, omega_max=4.0, seed=4, stop=300.0)),
    # generation-0 comparison of test/deac_parity_test.py:45-73
    "parity_gen0": ("standard", SMOKE, dict(T=1.0, N=2, P=8, frequency=[0.0, 0.7, 1.8, 3.0], seed=17, stop=1e300)),
    # odd sizes: every tail of the in-order reductions
    "odd_T5_P5_G7": ("standard", smooth_isf(5, -1.0), dict(T=1.3, N=200, P=5, M=7, omega_max=5.0, seed=11, stop=0.0,
                                                           normalize=True, first_moment=0.7, third_moment=2.0,
                                                           third_moment_error=0.5)),
    "odd_T6_P6_G5": ("standard", smooth_isf(6, -1.0), dict(T=1.3, N=200, P=6, M=5, omega_max=5.0, seed=12, stop=0.0,
                                                           normalize=True)),
    "odd_T7_P7_G6": ("standard", smooth_isf(7, -1.0), dict(T=1.3, N=200, P=7, M=6, omega_max=5.0, seed=13, stop=0.0,
                                                           first_moment=0.4)),
    "odd_T3_P4_G3": ("standard", smooth_isf(3, -1.0), dict(T=1.3, N=200, P=4, M=3, omega_max=5.0, seed=14, stop=0.0)),
    "odd_T2_P9_G2": ("standard", smooth_isf(2, -1.0), dict(T=1.3, N=200, P=9, M=2, omega_max=5.0, seed=15, stop=0.0)),
    "odd_T9_P11_G13_freqfile": ("standard", smooth_isf(9, -1.0), dict(
        T=1.3, N=200, P=11, frequency=[0.0, 0.1, 0.35, 0.5, 0.9, 1.3, 1.31, 2.0, 2.6, 3.3, 4.0, 4.9, 6.0], seed=16,
        stop=0.0, normalize=True)),
}
for model in ("standard", "normalization", "hyperbolic"):
    for spectra in ("spbsf", "spfsf", "bfull", "ffull"):
        sign = 1.0 if spectra in ("bfull", "ffull") else -1.0
        CASES[f"model_{model}_{spectra}"] = (model, smooth_isf(9, sign), dict(
            T=1.3, N=300, P=8, M=16, omega_max=6.0, seed=21, stop=0.0, spectra=spectra, normalize=True,
            first_moment=0.6, third_moment=3.0, third_moment_error=0.7))
CASES["variant_single_particle"] = ("standard_sp", smooth_isf(9, -1.0), dict(
    T=1.3, N=300, P=8, M=15, omega_max=6.0, seed=31, stop=0.0, normalize=True))
CASES["variant_negative_weight"] = ("standard_neg", smooth_isf(9, -1.0), dict(
    T=1.3, N=300, P=8, M=16, omega_max=6.0, seed=32, stop=0.0, normalize=True))

# the bosonic detailed-balance builds (one sign, spectra type bdsf): every
# model with normalization, both moments and the -1st-moment term; a plain
# run; odd sizes; and the zero-temperature build (which forces the -1st
# moment off, deac.cpp:230-236)
for model in ("standard", "normalization", "hyperbolic"):
    CASES[f"bdsf_{model}"] = (f"{model}_bdsf", smooth_isf(9, 1.0), dict(
        T=1.3, N=300, P=8, M=16, omega_max=6.0, seed=41, stop=0.0, spectra="bdsf", normalize=True,
        first_moment=0.6, third_moment=3.0, third_moment_error=0.7, use_negative_first_moment=True))
CASES["bdsf_standard_plain"] = ("standard_bdsf", smooth_isf(9, 1.0), dict(
    T=1.3, N=300, P=8, M=16, omega_max=6.0, seed=42, stop=0.0, spectra="bdsf"))
CASES["bdsf_normalization_odd_T7_P5_G9"] = ("normalization_bdsf", smooth_isf(7, 1.0), dict(
    T=1.3, N=200, P=5, M=9, omega_max=5.0, seed=43, stop=0.0, spectra="bdsf", use_negative_first_moment=True))
CASES["bdsf_hyperbolic_odd_T6_P6_G5"] = ("hyperbolic_bdsf", smooth_isf(6, 1.0), dict(
    T=1.3, N=200, P=6, M=5, omega_max=5.0, seed=44, stop=0.0, spectra="bdsf", normalize=True,
    use_negative_first_moment=True))
CASES["zerot_bdsf_moments"] = ("zerot_bdsf", smooth_isf(9, 1.0), dict(
    T=1.3, N=300, P=8, M=16, omega_max=6.0, seed=45, stop=0.0, spectra="bdsf", normalize=True,
    first_moment=0.6, third_moment=3.0, third_moment_error=0.7, use_negative_first_moment=True))
CASES["zerot_bdsf_odd_T5_P7_G6"] = ("zerot_bdsf", smooth_isf(5, 1.0), dict(
    T=1.3, N=200, P=7, M=6, omega_max=5.0, seed=46, stop=0.0, spectra="bdsf"))


def synthetic(name):
    a = np.fromfile(os.path.join(HERE, name)).reshape(3, -1)
    return dict(tau=a[0].tolist(), isf=a[1].tolist(), err=a[2].tolist())


# benchmark workloads (full config-1 run; prefixes of the larger ones)
CASES["c1_full"] = ("standard", "isf_c1_T33.bin", dict(T=1.35, N=100000, P=8, M=512, omega_max=64.0, seed=1, stop=0.0,
                                                     normalize=True, spectra="bfull"))
CASES["c2_prefix2000"] = ("standard", "isf_c2_T65.bin", dict(T=1.35, N=2000, P=8, M=4096, omega_max=512.0, seed=1,
                                                           stop=0.0, normalize=True, spectra="bfull"))
CASES["c4_seed7_prefix300"] = ("standard", "isf_c2_T65.bin", dict(T=1.35, N=300, P=32, M=4096, omega_max=512.0,
                                                                seed=7, stop=1.0, normalize=True, spectra="bfull"))
# long horizons at the headline shapes: config 2 in full, and two config-4
# seeds (P = 32, stop at chi^2 <= 1) for 2000 generations
CASES["c2_full"] = ("standard", "isf_c2_T65.bin", dict(T=1.35, N=200000, P=8, M=4096, omega_max=512.0, seed=1,
                                                     stop=0.0, normalize=True, spectra="bfull"))
for _seed in (1, 200):
    CASES[f"c4_seed{_seed}_prefix2000"] = ("standard", "isf_c2_T65.bin", dict(
        T=1.35, N=2000, P=32, M=4096, omega_max=512.0, seed=_seed, stop=1.0, normalize=True, spectra="bfull"))
CASES["bdsf_c1_shape_prefix2000"] = ("standard_bdsf", "isf_c1_T33.bin", dict(
    T=1.35, N=2000, P=8, M=512, omega_max=64.0, seed=1, stop=0.0, normalize=True, spectra="bdsf",
    use_negative_first_moment=True))
CASES["zerot_c1_shape_prefix2000"] = ("zerot_bdsf", "isf_c1_T33.bin", dict(
    T=1.35, N=2000, P=8, M=512, omega_max=64.0, seed=1, stop=0.0, normalize=True, spectra="bdsf"))
CASES["c3_hyperbolic_prefix4"] = ("hyperbolic", "isf_c3_T257.bin", dict(T=1.35, N=4, P=64, M=16384, omega_max=512.0,
                                                                      seed=1, stop=0.0, spectra="bfull"))


def run_case(name, variant, data, args, workdir):
    if isinstance(data, str):
        data = synthetic(data)
    isf_path = os.path.join(workdir, "isf.bin")
    np.concatenate([data["tau"], data["isf"], data["err"]]).astype("<f8").tofile(isf_path)
    spectra = args.get("spectra", "spfsf" if data["isf"][0] < 0 else "bfull")
    cmd = [os.path.join(REF, f"deac_{variant}.e"), "-T", repr(args["T"]), "-N", str(args["N"]), "-P", str(args["P"]),
           "--seed", str(args["seed"]), "--stop_minimum_fitness", repr(args["stop"]), "--spectra_type", spectra,
           "--save_directory", workdir, "--track_stats"]
    if "frequency" in args:
        fpath = os.path.join(workdir, "freq.bin")
        np.asarray(args["frequency"], dtype="<f8").tofile(fpath)
        cmd += ["--frequency_file", fpath]
    else:
        cmd += ["-M", str(args["M"]), "--omega_max", repr(args["omega_max"])]
    if args.get("normalize"):
        cmd.append("--normalize")
    if args.get("use_negative_first_moment"):
        cmd.append("--use_negative_first_moment")
    for k in ("first_moment", "third_moment", "third_moment_error"):
        if k in args:
            cmd += ["--" + k, repr(args[k])]
    cmd.append(isf_path)
    out = subprocess.run(cmd, capture_output=True, text=True, check=True)
    uuid = str(args["seed"])
    zt = "zT" if variant.startswith("zerot") else ""  # deac.cpp:1688-1692
    pre = os.path.join(workdir, f"deac-{zt}{spectra}_")
    res = {k: np.fromfile(f"{pre}stats_fitness-{k}_{uuid}.bin") for k in ("mean", "minimum", "squared-mean")}
    res["dsf"] = np.fromfile(f"{pre}dsf_{uuid}.bin")
    res["frequency_out"] = np.fromfile(f"{pre}frequency_{uuid}.bin")
    gen = int([l for l in out.stdout.splitlines() if l.startswith("generation:")][0].split()[1])
    meta = dict(variant=variant, spectra=spectra, args=args, generation=gen,
                stdout_minimum_fitness=[l for l in out.stdout.splitlines() if l.startswith("minimum_fitness:")][0])
    arrays = dict(tau=np.asarray(data["tau"]), isf=np.asarray(data["isf"]), err=np.asarray(data["err"]),
                  mean=res["mean"], minimum=res["minimum"], squared_mean=res["squared-mean"], dsf=res["dsf"],
                  frequency_out=res["frequency_out"], meta=np.array(json.dumps(meta)))
    if name in ("c1_full", "c2_full"):
        # keep the fixture small: the full best-fitness history, 5000
        # generations of the other two histories
        arrays["mean"] = arrays["mean"][:5000]
        arrays["squared_mean"] = arrays["squared_mean"][:5000]
    return arrays


def main(names):
    for name in names or CASES:
        variant, data, args = CASES[name]
        with tempfile.TemporaryDirectory() as d:
            arrays = run_case(name, variant, data, args, d)
        np.savez_compressed(os.path.join(HERE, f"golden_{name}.npz"), **arrays)
        print(name, json.loads(str(arrays["meta"]))["generation"], flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
