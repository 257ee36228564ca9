"""The bosonic detailed-balance variant (spectra type bdsf) on the GPU.

The reference builds it with USE_BOSONIC_DETAILED_BALANCE_CONDITION_DSF
(src/deac/CMakeLists.txt:21,33-34): one sign of genes, the detailed-balance
kernel of each model, the -1st-moment term, G-length output, and -- in the
reference CPU build -- the trial population's model ISF summed as a fold of
rounded products.  Its zero-temperature build is the only zero-temperature
build that compiles.  The bdsf_* and zerot_* goldens come from those builds;
tests/test_gpu_parity.py and tests/test_split_gpu.py run them through
deac_run bit for bit.  Here: the command line against the same goldens, seed
batches on the batch kernel's fold contraction against the oracle, and a
mixed-precision run.
"""
import os
import subprocess

import numpy as np
import pytest

import deac_amd as da
import golden_cases as gc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "deac-cpp_amd", "deac.e")


def _cli(case, workdir, seed, extra=()):
    m, a = case["meta"], case["meta"]["args"]
    isf = workdir / "isf.bin"
    np.concatenate([case["tau"], case["isf"], case["err"]]).astype("<f8").tofile(isf)
    cmd = [EXE, "-T", repr(a["T"]), "-N", str(a["N"]), "-P", str(a["P"]), "-M", str(a["M"]),
           "--omega_max", repr(a["omega_max"]), "--seed", str(seed), "--stop_minimum_fitness", repr(a["stop"]),
           "--spectra_type", "bdsf", "--save_directory", str(workdir / "out"), "--track_stats",
           "--model", m["variant"].split("_")[0] if not m["variant"].startswith("zerot") else "standard"]
    if m["variant"].startswith("zerot"):
        cmd.append("--zero_temperature")
    if a.get("normalize"):
        cmd.append("--normalize")
    if a.get("use_negative_first_moment"):
        cmd.append("--use_negative_first_moment")
    for k in ("first_moment", "third_moment", "third_moment_error"):
        if k in a:
            cmd += ["--" + k, repr(a[k])]
    cmd += list(extra) + [str(isf)]
    return subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)


def _files(workdir, prefix, seed):
    out = workdir / "out"
    got = {k: np.fromfile(out / f"{prefix}_stats_fitness-{f}_{seed}.bin")
           for k, f in (("mean", "mean"), ("minimum", "minimum"), ("squared_mean", "squared-mean"))}
    got["dsf"] = np.fromfile(out / f"{prefix}_dsf_{seed}.bin")
    got["frequency"] = np.fromfile(out / f"{prefix}_frequency_{seed}.bin")
    log = dict(l.split(": ", 1) for l in (out / f"{prefix}_log_{seed}.dat").read_text().splitlines() if ": " in l)
    got["generation"] = int(log["generation"])
    return got


@pytest.mark.parametrize("name", ["bdsf_standard", "bdsf_hyperbolic", "zerot_bdsf_moments"])
def test_cli_bdsf_matches_reference_golden(tmp_path, name):
    print(f"[gpu] deac.e --spectra_type bdsf, golden {name}", flush=True)
    case = gc.load(name)
    a = case["meta"]["args"]
    r = _cli(case, tmp_path, a["seed"])
    assert r.returncode == 0, r.stdout
    zt = case["meta"]["variant"].startswith("zerot")
    if zt:  # deac.cpp:230-236
        assert "use_negative_first_moment disabled for zero temperature build" in r.stdout
    got = _files(tmp_path, "deac-zTbdsf" if zt else "deac-bdsf", a["seed"])
    errors = gc.compare_outputs(got, case)
    assert not errors, f"{name}: " + "; ".join(errors)


def test_bdsf_seed_batch_on_the_fold_kernel(tmp_path):
    # deac.e batches bdsf seeds (the batch kernel's fold contraction), each
    # equal to its own reference run
    print("[gpu] bdsf seeds batched by deac.e", flush=True)
    case = gc.load("bdsf_normalization_odd_T7_P5_G9")
    a = case["meta"]["args"]
    p = da.make_problem(case["tau"], case["isf"], case["err"], gc.frequency_of(case, da.frequency_grid), a["T"],
                        **gc.problem_kwargs(case))
    r = _cli(case, tmp_path, a["seed"], ["--number_of_seeds", "2"])
    assert r.returncode == 0, r.stdout
    assert "seeds run one after another" not in r.stdout, r.stdout
    errors = gc.compare_outputs(_files(tmp_path, "deac-bdsf", a["seed"]), case)
    assert not errors, "; ".join(errors)
    with da.Engine(p) as e:
        second = e.run(a["seed"] + 1, a["N"], a["stop"])
    got = _files(tmp_path, "deac-bdsf", a["seed"] + 1)
    for k in ("minimum", "mean", "squared_mean", "dsf", "frequency"):
        assert gc.bit_diff(got[k], second[k]) is None, k


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_bdsf_mixed_precision_split_loop(prec):
    # K.S at reduced precision in the split loop: the reported fitness is the
    # mixed-precision chi^2 of the best genome (standard bdsf: the DSF is the
    # genome itself), the history never increases
    print(f"[gpu] bdsf split loop at {prec}", flush=True)
    case = gc.load("bdsf_c1_shape_prefix2000")
    a = case["meta"]["args"]
    kw = dict(gc.problem_kwargs(case), use_negative_first_moment=False)
    p = da.make_problem(case["tau"], case["isf"], case["err"], gc.frequency_of(case, da.frequency_grid), a["T"], **kw)
    with da.Engine(p) as e:
        ref = e.run(a["seed"], 600, 0.0)
        got = e.run(a["seed"], 600, 0.0, precision=prec)
    assert len(got["dsf"]) == len(got["frequency"]) == a["M"]
    assert np.all(np.diff(got["minimum"]) <= 0)
    _, chi, _ = da.model_isf(p, got["dsf"], prec)
    assert chi[0] == got["minimum_fitness"]
    tol = {"fp32": 1e-4, "bf16x3": 1e-3}[prec]
    assert abs(got["minimum"][0] - ref["minimum"][0]) <= tol * ref["minimum"][0]


# bdsf batches against the oracle at shapes beyond the goldens: the fold
# kernel's lane layout (256 / T member groups), odd genome lengths (the fused
# last position), the -1st-moment term, moments, zero temperature, launch
# slices, uneven generation counts and a stop inside the batch
FOLD = {
    "standard_T65_G513_P32": dict(T=65, G=513, temp=1.0, kw=dict(population_size=32, use_negative_first_moment=True)),
    "hyperbolic_T33_G300_P12": dict(T=33, G=300, temp=1.2, kw=dict(model="hyperbolic", population_size=12,
                                                                    use_negative_first_moment=True,
                                                                    first_moment=0.6, third_moment=3.0,
                                                                    third_moment_error=0.5)),
    "zeroT_normalization_T17_G97_P8": dict(T=17, G=97, temp=0.0, kw=dict(model="normalization", population_size=8,
                                                                          zero_temperature=True, normalize=True)),
    "normalization_T100_G64_P20": dict(T=100, G=64, temp=0.8, kw=dict(model="normalization", population_size=20,
                                                                      use_negative_first_moment=True)),
}


@pytest.mark.parametrize("name", sorted(FOLD))
@pytest.mark.parametrize("slice_len", [None, "9"])
def test_bdsf_batch_matches_oracle(name, slice_len, monkeypatch):
    import oracle_binding as ob
    spec = FOLD[name]
    print(f"[gpu] bdsf batch {name} slice={slice_len}", flush=True)
    if slice_len:
        monkeypatch.setenv("DEAC_BATCH_SLICE", slice_len)
    T = spec["T"]
    beta = 1.0 / spec["temp"] if spec["temp"] > 0 else 1.0
    tau = np.linspace(0.0, beta / 2, T)
    isf = 0.7 * np.exp(-1.3 * tau) + 0.3 * np.exp(-0.2 * tau)
    err = 2e-3 * isf
    freq = da.frequency_grid(spec["G"], 16.0)
    prob = da.make_problem(tau, isf, err, freq, spec["temp"], spectra_type="bdsf", **spec["kw"])
    seeds, Ns = [3, 4, 5], [90, 61, 90]

    def oracle(sd, N, stop):
        return ob.run(ob.make_problem(tau, isf, err, ob.frequency_grid(spec["G"], 16.0), spec["temp"],
                                      spectra_type="bdsf", number_of_generations=N, stop_minimum_fitness=stop,
                                      **spec["kw"]), sd)
    stops = [0.0, 0.0, float(oracle(5, 90, 0.0)["minimum"][40])]
    with da.Batch([prob] * 3) as b:
        outs = b.run(seeds, Ns, stops, True)
    for sd, N, st, o in zip(seeds, Ns, stops, outs):
        r = oracle(sd, N, st)
        for k in ("minimum", "mean", "squared_mean", "dsf", "frequency"):
            assert gc.bit_diff(o[k], r[k]) is None, (sd, k, gc.bit_diff(o[k], r[k]))
        assert o["rng"] == r["rng"] and o["generation"] == r["generation"], sd


# One case per register-tile shape of the fold kernel (2 RL + XR: RL
# timeslices per lane, R = 64 RL / NM of them in the row blocks, XR extra
# timeslices past R), for each member-tile count NM (P <= 16, 32, 64), with
# odd and even genome lengths (the fused last position, partial chunks) and
# whole chunks of 128, 64 or 32 positions before the last
FOLD_SHAPES = {
    "nm1_rl1_x_T66_P16": dict(T=66, G=301, P=16),
    "nm1_rl2_T100_P10": dict(T=100, G=257, P=10),
    "nm1_rl2_T100_P10_one_chunk": dict(T=100, G=64, P=10),
    "nm2_rl1_T30_P24": dict(T=30, G=130, P=24),
    "nm2_rl1_x_T34_P24": dict(T=34, G=75, P=24),
    "nm2_rl2_T64_P20": dict(T=64, G=200, P=20),
    "nm4_rl1_T16_P40": dict(T=16, G=33, P=40),
    "nm4_rl1_x_T19_P36": dict(T=19, G=64, P=36),
    "nm4_rl2_T32_P48": dict(T=32, G=65, P=48),
    "nm4_rl2_x_T35_P64": dict(T=35, G=96, P=64),
    "nm4_rl4_T50_P40": dict(T=50, G=47, P=40),
    "nm4_rl4_x_T68_P48": dict(T=68, G=120, P=48),
    "nm4_rl8_T128_P64": dict(T=128, G=71, P=64),
}


@pytest.mark.parametrize("name", sorted(FOLD_SHAPES))
def test_bdsf_fold_shapes_match_oracle(name):
    import oracle_binding as ob
    spec = FOLD_SHAPES[name]
    print(f"[gpu] bdsf fold shape {name}", flush=True)
    T, G, P = spec["T"], spec["G"], spec["P"]
    tau = np.linspace(0.0, 0.4, T)
    isf = 0.6 * np.exp(-1.1 * tau) + 0.4 * np.exp(-0.3 * tau)
    err = 2e-3 * isf
    kw = dict(population_size=P, spectra_type="bdsf", use_negative_first_moment=True)
    prob = da.make_problem(tau, isf, err, da.frequency_grid(G, 16.0), 1.25, **kw)
    seeds, N = [11, 12], 70
    with da.Batch([prob] * 2) as b:
        outs = b.run(seeds, N, 0.0, True)
    for sd, o in zip(seeds, outs):
        r = ob.run(ob.make_problem(tau, isf, err, ob.frequency_grid(G, 16.0), 1.25, number_of_generations=N,
                                   stop_minimum_fitness=0.0, **kw), sd)
        for k in ("minimum", "mean", "squared_mean", "dsf", "frequency"):
            assert gc.bit_diff(o[k], r[k]) is None, (sd, k, gc.bit_diff(o[k], r[k]))
        assert o["rng"] == r["rng"] and o["generation"] == r["generation"], sd


def test_bdsf_batch_continues():
    case = gc.load("bdsf_standard")
    a = case["meta"]["args"]
    p = da.make_problem(case["tau"], case["isf"], case["err"], gc.frequency_of(case, da.frequency_grid), a["T"],
                        **gc.problem_kwargs(case))
    with da.Batch([p, p]) as b:
        whole = b.run([7, 8], 80, 0.0, True)
        first = b.run([7, 8], 31, 0.0, True)
        rest = b.continue_run(50, 0.0, True)
    for w, f, r in zip(whole, first, rest):
        assert gc.bit_diff(np.concatenate([f["minimum"], r["minimum"][1:]]), w["minimum"]) is None
        assert gc.bit_diff(r["dsf"], w["dsf"]) is None and r["rng"] == w["rng"]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("nfm,P", [(True, 24), (False, 24), (True, 48)])
def test_bdsf_mixed_batch(prec, nfm, P):
    # detailed-balance instances at FP32 / BF16x3 K.S run on the mixed kernel
    # (delta form; the -1st-moment term from the mixed model ISF): the
    # initial fitness is the FP64 evaluation's, generation 0's best matches
    # the FP64 (fold) batch to the precision of the contraction, where the
    # -1st-moment term moves it by far more, and runs repeat bit for bit
    print(f"[gpu] bdsf mixed batch {prec} nfm={nfm} P={P}", flush=True)
    T, G = 65, 513
    tau = np.linspace(0.0, 0.4, T)
    isf = 0.6 * np.exp(-1.1 * tau) + 0.4 * np.exp(-0.3 * tau)
    err = 2e-3 * isf

    def problem(use_nfm):
        return da.make_problem(tau, isf, err, da.frequency_grid(G, 16.0), 1.25, population_size=P,
                               spectra_type="bdsf", use_negative_first_moment=use_nfm, normalize=True)
    seeds, N = [3, 4, 5, 6], 60
    with da.Batch([problem(nfm)] * len(seeds)) as b:
        ref = b.run(seeds, N, 0.0, True, precision="fp64")
        out = b.run(seeds, N, 0.0, True, precision=prec)
        again = b.run(seeds, N, 0.0, True, precision=prec)
    if nfm:
        with da.Batch([problem(False)] * len(seeds)) as b:
            bare = b.run(seeds, N, 0.0, True, precision="fp64")
        assert all(abs(z["minimum"][1] - r["minimum"][1]) > 1e-3 * r["minimum"][1] for z, r in zip(bare, ref))
    for k, (o, r, a) in enumerate(zip(out, ref, again)):
        assert o["minimum"][0] == r["minimum"][0], k
        assert o["rng"] == r["rng"], k
        assert abs(o["minimum"][1] - r["minimum"][1]) <= 1e-5 * r["minimum"][1], (k, o["minimum"][1], r["minimum"][1])
        assert gc.bit_diff(o["minimum"], a["minimum"]) is None and gc.bit_diff(o["dsf"], a["dsf"]) is None, k
        assert o["dsf"].shape == r["dsf"].shape
    fin = np.median([o["minimum_fitness"] for o in out])
    finr = np.median([r["minimum_fitness"] for r in ref])
    assert 0.3 < fin / finr < 3.0, (fin, finr)
