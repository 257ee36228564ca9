"""deac.e command-line surface: the reference's smoke cases
(test/deac_smoke_test.py:9-34).  Validation and help cases exit before any
device call and run everywhere; the run cases need the GPU and compare the
written files with the reference executable's outputs."""
import os
import struct
import subprocess

import numpy as np
import pytest

import golden_cases as gc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "deac-cpp_amd", "deac.e")


def write_doubles(path, values):
    path.write_bytes(struct.pack("<" + "d" * len(values), *values))


def write_fixture(path, tau=None, isf=None, error=None):
    tau = [0.0, 0.2, 0.4, 0.6] if tau is None else tau
    isf = [-1.0, -0.85, -0.72, -0.61] if isf is None else isf
    error = [0.05] * 4 if error is None else error
    write_doubles(path, tau + isf + error)


def deac_command(workdir, fixture, N="2", P="8", M="8", omega_max="4.0", seed="7", extra=None):
    cmd = [EXE, "-T", "1.0", "-N", N, "-P", P, "-M", M, "--omega_max", omega_max,
           "--save_directory", str(workdir / "results"), "--seed", seed]
    return cmd + list(extra or []) + [str(fixture)]


def run(cmd, cwd):
    return subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)


@pytest.fixture(scope="module", autouse=True)
def need_exe():
    if not os.path.exists(EXE):
        pytest.fail(f"deac.e not built at {EXE} (make -C deac-cpp_amd)")


def test_help(tmp_path):
    r = run([EXE, "--help"], tmp_path)
    assert r.returncode == 0
    assert "Usage: deac-cpp" in r.stdout
    for flag in ("--temperature", "--number_of_generations", "--spectra_type", "--track_stats", "isf_file"):
        assert flag in r.stdout


def test_bad_spectra(tmp_path):
    fixture = tmp_path / "tiny-isf.bin"
    write_fixture(fixture)
    r = run([EXE, "-T", "1.0", "--spectra_type", "invalid", str(fixture)], tmp_path)
    assert r.returncode == 1
    assert "Please choose spectra_type" in r.stdout


def test_bad_ks_precision(tmp_path):
    fixture = tmp_path / "tiny-isf.bin"
    write_fixture(fixture)
    r = run(deac_command(tmp_path, fixture, extra=["--ks_precision", "fp16"]), tmp_path)
    assert r.returncode == 1
    assert "Please choose ks_precision from the following options: fp64, fp32, bf16x3" in r.stdout


def test_missing_isf_file_argument(tmp_path):
    r = run([EXE, "-T", "1.0"], tmp_path)
    assert r.returncode == 1
    assert "isf_file: 1 argument(s) expected. 0 provided." in r.stdout
    assert "Usage: deac-cpp" in r.stdout


def test_unknown_argument(tmp_path):
    r = run([EXE, "--bogus", "1", "x.bin"], tmp_path)
    assert r.returncode == 1
    assert "Unknown argument: --bogus" in r.stdout


def test_missing_file(tmp_path):
    r = run([EXE, "-T", "1.0", str(tmp_path / "nope.bin")], tmp_path)
    assert r.returncode == 1
    assert "File error:" in r.stdout


VALIDATION_CASES = {
    "bad_isf_byte_length": "does not contain a whole number of doubles",
    "empty_isf": "is empty",
    "uneven_isf_arrays": "ISF input file must contain tau, isf, and error arrays of equal length",
    "too_few_timeslices": "ISF input file must contain at least two timeslices",
    "nonfinite_isf": "ISF input file contains non-finite values",
    "positive_isf_single_particle": "positive ISF values are not supported for single-particle spectra",
    "bad_third_moment_error": "third_moment_error must be positive when third_moment is used",
    "too_few_generations": "number_of_generations must be at least 2",
    "too_small_population": "population_size must be at least 4",
    "too_small_genome": "genome_size must be at least 2",
    "bad_omega_max": "omega_max must be positive",
    "short_frequency_file": "frequency_file must contain at least two frequencies",
    "nonfinite_frequency": "frequencies must be finite and non-negative",
    "negative_frequency": "frequencies must be finite and non-negative",
    "unsorted_frequency": "frequencies must be sorted in non-decreasing order",
    "bad_temperature": "temperature must be positive for non-ZeroT builds",
    "nonpositive_error": "ISF input errors must be positive",
}


@pytest.mark.parametrize("case", sorted(VALIDATION_CASES))
def test_validation_case(tmp_path, case):
    # deac_smoke_test.py:180-259
    fixture = tmp_path / "invalid-isf.bin"
    freq = tmp_path / "frequency.bin"
    opts, extra = {}, []
    if case == "bad_isf_byte_length":
        fixture.write_bytes(b"not-a-double")
    elif case == "empty_isf":
        fixture.write_bytes(b"")
    elif case == "uneven_isf_arrays":
        write_doubles(fixture, [0.0, 0.2, -1.0, 0.05])
    elif case == "too_few_timeslices":
        write_fixture(fixture, tau=[0.0], isf=[-1.0], error=[0.05])
    elif case == "nonfinite_isf":
        write_fixture(fixture, isf=[-1.0, float("nan"), -0.72, -0.61])
    elif case == "positive_isf_single_particle":
        write_fixture(fixture, isf=[1.0, 0.85, 0.72, 0.61])
    elif case == "nonpositive_error":
        write_fixture(fixture, error=[0.05, 0.0, 0.05, 0.05])
    else:
        write_fixture(fixture)
    if case == "bad_third_moment_error":
        extra = ["--third_moment", "1.0", "--third_moment_error", "0.0"]
    elif case == "too_few_generations":
        opts["N"] = "1"
    elif case == "too_small_population":
        opts["P"] = "3"
    elif case == "too_small_genome":
        opts["M"] = "1"
    elif case == "bad_omega_max":
        opts["omega_max"] = "0.0"
    elif case in ("short_frequency_file", "nonfinite_frequency", "negative_frequency", "unsorted_frequency"):
        values = {"short_frequency_file": [0.0], "nonfinite_frequency": [0.0, float("inf")],
                  "negative_frequency": [0.0, -1.0], "unsorted_frequency": [0.0, 2.0, 1.0]}[case]
        write_doubles(freq, values)
        extra = ["--frequency_file", str(freq)]
    cmd = deac_command(tmp_path, fixture, extra=extra, **opts)
    if case == "bad_temperature":
        cmd[cmd.index("-T") + 1] = "0.0"
    r = run(cmd, tmp_path)
    assert r.returncode == 1, r.stdout
    assert VALIDATION_CASES[case] in r.stdout


# ---------------------------------------------------------------- GPU runs

def _golden_args(name):
    case = gc.load(name)
    a = case["meta"]["args"]
    return case, a


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["smoke_default", "smoke_normalize", "smoke_first_moment", "smoke_third_moment",
                                  "smoke_stop", "parity_gen0", "odd_T9_P11_G13_freqfile"])
def test_cli_run_matches_reference_files(tmp_path, name):
    case, a = _golden_args(name)
    fixture = tmp_path / "isf.bin"
    write_doubles(fixture, list(case["tau"]) + list(case["isf"]) + list(case["err"]))
    cmd = [EXE, "-T", repr(a["T"]), "-N", str(a["N"]), "-P", str(a["P"]), "--seed", str(a["seed"]),
           "--stop_minimum_fitness", repr(a["stop"]), "--spectra_type", case["meta"]["spectra"],
           "--save_directory", str(tmp_path / "out"), "--track_stats"]
    if "frequency" in a:
        write_doubles(tmp_path / "freq.bin", a["frequency"])
        cmd += ["--frequency_file", str(tmp_path / "freq.bin")]
    else:
        cmd += ["-M", str(a["M"]), "--omega_max", repr(a["omega_max"])]
    if a.get("normalize"):
        cmd.append("--normalize")
    for k in ("first_moment", "third_moment", "third_moment_error"):
        if k in a:
            cmd += ["--" + k, repr(a[k])]
    cmd.append(str(fixture))
    r = run(cmd, tmp_path)
    assert r.returncode == 0, r.stdout
    assert "minimum_fitness:" in r.stdout
    assert f"generation: {case['meta']['generation']}" in r.stdout
    assert case["meta"]["stdout_minimum_fitness"] in r.stdout
    pre = tmp_path / "out" / f"deac-{case['meta']['spectra']}_"
    uuid = str(a["seed"])
    got = {"dsf": np.fromfile(f"{pre}dsf_{uuid}.bin"), "frequency_out": np.fromfile(f"{pre}frequency_{uuid}.bin"),
           "mean": np.fromfile(f"{pre}stats_fitness-mean_{uuid}.bin"),
           "minimum": np.fromfile(f"{pre}stats_fitness-minimum_{uuid}.bin"),
           "squared_mean": np.fromfile(f"{pre}stats_fitness-squared-mean_{uuid}.bin")}
    for key in got:
        d = gc.bit_diff(got[key], case[key])
        assert d is None, f"{key}: {d}"
    log = (tmp_path / "out" / f"deac-{case['meta']['spectra']}_log_{uuid}.dat").read_text()
    for key in ("uuid:", "isf_file:", "build: USE_STANDARD_MODEL", "minimum_fitness:", "fitness_mean_filename:",
                "generation:"):
        assert key in log
    assert 'fitness_mean_filename: ""' not in log


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["default", "normalize", "first_moment", "track_stats"])
def test_cli_smoke_runs(tmp_path, case):
    # deac_smoke_test.py:128-177: tiny runs, file sizes and log keys
    fixture = tmp_path / "tiny-isf.bin"
    write_fixture(fixture)
    seed = {"default": "1", "normalize": "2", "first_moment": "3", "track_stats": "4"}[case]
    extra, N = [], "2"
    if case == "normalize":
        extra = ["--normalize"]
    elif case == "first_moment":
        extra = ["--first_moment", "0.5"]
    elif case == "track_stats":
        N, extra = "3", ["--track_stats"]
    r = run(deac_command(tmp_path, fixture, N=N, seed=seed, extra=extra), tmp_path)
    assert r.returncode == 0, r.stdout
    assert "minimum_fitness:" in r.stdout
    res = tmp_path / "results"
    assert (res / f"deac-spfsf_dsf_{seed}.bin").stat().st_size == (2 * 8 - 1) * 8
    assert (res / f"deac-spfsf_frequency_{seed}.bin").stat().st_size == (2 * 8 - 1) * 8
    assert "minimum_fitness:" in (res / f"deac-spfsf_log_{seed}.dat").read_text()
    if case == "track_stats":
        for k in ("mean", "minimum", "squared-mean"):
            assert (res / f"deac-spfsf_stats_fitness-{k}_{seed}.bin").stat().st_size == 3 * 8


# ---------------------------------------------------------------- seed batches

def test_number_of_seeds_must_be_positive(tmp_path):
    fixture = tmp_path / "tiny-isf.bin"
    write_fixture(fixture)
    r = run(deac_command(tmp_path, fixture, extra=["--number_of_seeds", "0"]), tmp_path)
    assert r.returncode == 1
    assert "number_of_seeds must be at least 1" in r.stdout


def test_uuid_with_several_seeds_is_rejected(tmp_path):
    fixture = tmp_path / "tiny-isf.bin"
    write_fixture(fixture)
    r = run(deac_command(tmp_path, fixture, extra=["--number_of_seeds", "2", "--uuid", "x"]), tmp_path)
    assert r.returncode == 1
    assert "--uuid cannot be combined with --number_of_seeds > 1" in r.stdout


def _log_without_paths(path):
    return [l for l in path.read_text().splitlines() if "filename:" not in l]


def _seed_batch_vs_single_runs(tmp_path, extra, seeds, expect_batched, P="8"):
    fixture = tmp_path / "tiny-isf.bin"
    write_fixture(fixture)
    batch_dir, single_dir = tmp_path / "batch", tmp_path / "single"
    cmd = deac_command(tmp_path, fixture, N="60", P=P, seed=str(seeds[0]),
                       extra=list(extra) + ["--track_stats", "--number_of_seeds", str(len(seeds)), "--average_dsf"])
    cmd[cmd.index("--save_directory") + 1] = str(batch_dir)
    r = run(cmd, tmp_path)
    assert r.returncode == 0, r.stdout
    assert ("seeds run one after another" not in r.stdout) == expect_batched, r.stdout
    for s in seeds:
        cmd = deac_command(tmp_path, fixture, N="60", P=P, seed=str(s), extra=list(extra) + ["--track_stats"])
        cmd[cmd.index("--save_directory") + 1] = str(single_dir)
        rs = run(cmd, tmp_path)
        assert rs.returncode == 0, rs.stdout
    for s in seeds:
        for kind in ("dsf", "frequency", "stats_fitness-mean", "stats_fitness-minimum", "stats_fitness-squared-mean"):
            a = np.fromfile(batch_dir / f"deac-spfsf_{kind}_{s}.bin")
            b = np.fromfile(single_dir / f"deac-spfsf_{kind}_{s}.bin")
            assert a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64)), (s, kind)
        assert _log_without_paths(batch_dir / f"deac-spfsf_log_{s}.dat") == \
            _log_without_paths(single_dir / f"deac-spfsf_log_{s}.dat")
    # tools/plot_simple_average_deac_results.py:8-28 over the single runs: keys
    # sorted as strings, the first DSF plus the others in that order, / count
    keys = np.sort(np.array([str(s) for s in seeds]))
    avg = np.fromfile(single_dir / f"deac-spfsf_dsf_{keys[0]}.bin")
    for k in keys[1:]:
        avg += np.fromfile(single_dir / f"deac-spfsf_dsf_{k}.bin")
    avg /= keys.shape[0]
    got = np.fromfile(batch_dir / f"deac-spfsf_seed-average_{seeds[0]}-{seeds[-1]}.bin")
    assert np.array_equal(got.view(np.uint64), avg.view(np.uint64))


@pytest.mark.gpu
def test_cli_seed_batch_matches_single_runs(tmp_path):
    # seeds 8..12 sort as strings 10, 11, 12, 8, 9
    _seed_batch_vs_single_runs(tmp_path, ["--normalize"], list(range(8, 13)), expect_batched=True)


@pytest.mark.gpu
def test_cli_seed_batch_with_moments_matches_single_runs(tmp_path):
    # moment constraints in the batch kernel (FP64): each seed exactly its own run
    _seed_batch_vs_single_runs(tmp_path, ["--normalize", "--first_moment", "0.5", "--third_moment", "2.0",
                                          "--third_moment_error", "0.5"], [3, 4], expect_batched=True)


@pytest.mark.gpu
def test_cli_seed_batch_64_members_matches_single_runs(tmp_path):
    # 33 to 64 members run in the FP64 batch kernel (four member tiles)
    _seed_batch_vs_single_runs(tmp_path, ["--normalize"], [3, 4, 5], expect_batched=True, P="64")


@pytest.mark.gpu
def test_cli_seed_batch_with_renormalization_matches_single_runs(tmp_path):
    # per-generation renormalization (single-particle spectra) in the batch kernel
    _seed_batch_vs_single_runs(tmp_path, ["--normalize", "--single_particle_fermionic_spectral_function"], [3, 4],
                               expect_batched=True)


@pytest.mark.gpu
def test_cli_seed_batch_unbatchable_runs_seed_by_seed(tmp_path):
    # more than 64 members run through the single-instance engine
    _seed_batch_vs_single_runs(tmp_path, ["--normalize"], [3, 4], expect_batched=False, P="70")


@pytest.mark.gpu
@pytest.mark.parametrize("prec,moments,P", [("fp32", False, 8), ("bf16x3", False, 8), ("fp32", True, 8),
                                             ("bf16x3", False, 48)])
def test_cli_mixed_precision_seed_batch(tmp_path, prec, moments, P):
    # fp32 / bf16x3 seeds run in the batch kernel (delta form), with moment
    # constraints too: each seed's files are that seed's instance of
    # deac_batch_run at the same precision, bit for bit, and its reported
    # fitness is the FP64 fitness of its genome
    import deac_amd as da
    fixture = tmp_path / "tiny-isf.bin"
    write_fixture(fixture)
    seeds = [5, 6, 7]
    mflags = ["--first_moment", "0.5", "--third_moment", "2.0", "--third_moment_error", "0.5"] if moments else []
    cmd = deac_command(tmp_path, fixture, N="300", P=str(P), seed=str(seeds[0]),
                       extra=["--normalize", "--ks_precision", prec, "--track_stats",
                              "--number_of_seeds", str(len(seeds))] + mflags)
    r = run(cmd, tmp_path)
    assert r.returncode == 0, r.stdout
    assert "seeds run one after another" not in r.stdout, r.stdout
    tau, isf, err = np.fromfile(fixture).reshape(3, -1)
    mkw = dict(first_moment=0.5, third_moment=2.0, third_moment_error=0.5) if moments else {}
    p = da.make_problem(tau, isf, err, da.frequency_grid(8, 4.0), 1.0, population_size=P, normalize=True,
                        spectra_type="spfsf", **mkw)
    with da.Batch([p] * len(seeds)) as b:
        outs = b.run(seeds, 300, 1.0, track_stats=True, precision=prec)
    res = tmp_path / "results"
    for s, o in zip(seeds, outs):
        n = o["generation"] + 1
        assert np.array_equal(np.fromfile(res / f"deac-spfsf_dsf_{s}.bin").view(np.uint64), o["dsf"].view(np.uint64))
        for kind, key in (("mean", "mean"), ("minimum", "minimum"), ("squared-mean", "squared_mean")):
            got = np.fromfile(res / f"deac-spfsf_stats_fitness-{kind}_{s}.bin")
            assert np.array_equal(got.view(np.uint64), o[key][:n].view(np.uint64)), (s, kind)
        log = dict(l.split(": ", 1) for l in (res / f"deac-spfsf_log_{s}.dat").read_text().splitlines() if ": " in l)
        assert int(log["generation"]) == o["generation"]
        assert abs(float(log["minimum_fitness"]) - o["minimum_fitness"]) <= 1e-5 * o["minimum_fitness"]
