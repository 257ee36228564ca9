"""The bosonic detailed-balance variant (spectra type bdsf) on the GPU.

The reference builds it with USE_BOSONIC_DETAILED_BALANCE_CONDITION_DSF
(src/deac/CMakeLists.txt:21,33-34): one sign of genes, the detailed-balance
kernel of each model, the -1st-moment term, G-length output, and -- in the
reference CPU build -- the trial population's model ISF summed as a fold of
rounded products.  Its zero-temperature build is the only zero-temperature
build that compiles.  The bdsf_* and zerot_* goldens come from those builds;
tests/test_gpu_parity.py and tests/test_split_gpu.py run them through
deac_run bit for bit.  Here: the command line against the same goldens, the
batch's refusal and the per-seed fallback, and a mixed-precision run.
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


def test_bdsf_seed_batch_runs_seed_by_seed(tmp_path):
    # the batch kernel refuses bdsf; deac.e runs the seeds one after another,
    # each equal to its own reference run
    print("[gpu] bdsf refused by the batch; deac.e seeds one after another", flush=True)
    case = gc.load("bdsf_normalization_odd_T7_P5_G9")  # no moment constraints: refused for bdsf itself
    a = case["meta"]["args"]
    p = da.make_problem(case["tau"], case["isf"], case["err"], gc.frequency_of(case, da.frequency_grid), a["T"],
                        **gc.problem_kwargs(case))
    with pytest.raises(da.DeacError, match="bdsf"):
        da.Batch([p, p])
    r = _cli(case, tmp_path, a["seed"], ["--number_of_seeds", "2"])
    assert r.returncode == 0, r.stdout
    assert "seeds run one after another" in r.stdout
    assert "batched runs do not take the detailed-balance spectra (bdsf)" in r.stdout
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
