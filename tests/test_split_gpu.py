"""The split generation loop on the GPU.

At FP64 the split loop (one kernel per step, the exact contraction) must give
the fused kernel's results bit for bit, and so the reference's: it runs every
golden case against the reference CPU build's outputs, and the options no
golden covers (renormalization, zero temperature) against the fused kernel.

At FP32 and BF16x3 the trajectory leaves FP64's at the first selection that
the contraction's rounding decides; what is pinned instead:
  * every fitness the loop reports is the FP64 chi^2 of the mixed-precision
    model ISF of its genome -- the best member's final fitness equals
    deac_model_isf at the same precision on the genome read back from the
    DSF, bit for bit (that entry point is itself bounded against FP64);
  * the initial best fitness is within the format's bound of FP64's;
  * the best-fitness history never increases.
"""
import ctypes as C

import numpy as np
import pytest

import deac_amd as da
import golden_cases as gc

pytestmark = pytest.mark.gpu


def _engine(case):
    a = case["meta"]["args"]
    freq = gc.frequency_of(case, da.frequency_grid)
    return da.Engine(da.make_problem(case["tau"], case["isf"], case["err"], freq, a["T"], **gc.problem_kwargs(case)))


@pytest.mark.parametrize("name", gc.names())
def test_split_loop_fp64_matches_reference_golden(name):
    print(f"[gpu] split loop fp64, golden {name}", flush=True)
    case = gc.load(name)
    a = case["meta"]["args"]
    # long runs on a prefix (the prefix of a run is itself a complete run)
    n = 3000 if a["N"] > 20000 else None
    with _engine(case) as e:
        got = e.run(a["seed"], n or a["N"], a["stop"], track_stats=True, split_loop=True)
    errors = gc.compare_outputs(got, case, n_hist=n)
    assert not errors, f"{name}: " + "; ".join(errors)


def _synthetic(fname):
    return np.fromfile(f"{gc.GOLDEN_DIR}/{fname}").reshape(3, -1)


@pytest.mark.parametrize("kw", [dict(normalize=True, renormalize=True), dict(zero_temperature=True),
                                dict(zero_temperature=True, model="normalization", spectra_type="ffull"),
                                dict(normalize=True, renormalize=True, first_moment=3.0)])
def test_split_loop_fp64_matches_fused(kw):
    print(f"[gpu] split loop fp64 vs fused, {kw}", flush=True)
    t, f, s = _synthetic("isf_c1_T33.bin")
    p = da.make_problem(t, f, s, da.frequency_grid(512, 64.0), 1.35, population_size=8, **kw)
    with da.Engine(p) as e:
        fused = e.run(3, 1500, 0.0)
        split = e.run(3, 1500, 0.0, split_loop=True)
    for k in ("minimum", "mean", "squared_mean", "dsf", "frequency"):
        assert gc.bit_diff(fused[k], split[k]) is None, k
    assert fused["rng"] == split["rng"] and fused["generation"] == split["generation"]


def _best_genome(dsf, G):
    # standard model, bfull (deac.cpp:1594-1675): dsf[G-1+i] = S+[i],
    # dsf[G-1-i] = S-[i]
    return np.concatenate([dsf[G - 1:], dsf[G - 1::-1]])


@pytest.mark.parametrize("prec,tol", [("fp32", 1e-4), ("bf16x3", 1e-3)])
def test_mixed_precision_trajectory(prec, tol):
    print(f"[gpu] mixed-precision trajectory {prec}", flush=True)
    t, f, s = _synthetic("isf_c2_T65.bin")
    G = 4096
    p = da.make_problem(t, f, s, da.frequency_grid(G, 512.0), 1.35, population_size=8, normalize=True)
    with da.Engine(p) as e:
        ref = e.run(1, 1000, 0.0)
        got = e.run(1, 1000, 0.0, precision=prec)
        again = e.run(1, 1000, 0.0, precision=prec)
    for k in ("minimum", "dsf"):
        assert gc.bit_diff(got[k], again[k]) is None, k  # deterministic
    assert np.all(np.diff(got["minimum"]) <= 0)
    assert abs(got["minimum"][0] - ref["minimum"][0]) <= tol * ref["minimum"][0]
    genome = _best_genome(got["dsf"], G)
    _, chi, _ = da.model_isf(p, genome, prec)
    assert chi[0] == got["minimum_fitness"] == got["minimum"][-1]
    _, chi64, _ = da.model_isf(p, genome, "fp64")
    rel = abs(got["minimum_fitness"] - ref["minimum_fitness"]) / ref["minimum_fitness"]
    l2 = np.linalg.norm(got["dsf"] - ref["dsf"]) / np.linalg.norm(ref["dsf"])
    print(f"[gpu] {prec}: final chi2 {got['minimum_fitness']:.6g} (fp64 re-evaluation {chi64[0]:.6g}) vs fp64 run "
          f"{ref['minimum_fitness']:.6g}: |dchi2|/chi2 = {rel:.3e}, DSF L2 error {l2:.3e}", flush=True)
    assert abs(chi64[0] - got["minimum_fitness"]) <= tol * chi64[0]


def test_split_loop_stop_matches_fused():
    # a stop inside a batch of the split loop: the later kernels of the batch
    # return at once, and the results and RNG state are those of the stop
    print("[gpu] split loop stop", flush=True)
    case = gc.load("smoke_stop")
    a = case["meta"]["args"]
    with _engine(case) as e:
        fused = e.run(a["seed"], a["N"], a["stop"])
        split = e.run(a["seed"], a["N"], a["stop"], split_loop=True)
        mixed = e.run(a["seed"], a["N"], a["stop"], precision="fp32")
    assert split["generation"] == fused["generation"] == case["meta"]["generation"]
    assert split["rng"] == fused["rng"]
    for k in ("minimum", "dsf"):
        assert gc.bit_diff(fused[k], split[k]) is None, k
    assert mixed["minimum"][-1] <= a["stop"] and np.all(mixed["minimum"][:-1] > a["stop"])


def test_run_rejects_unknown_precision():
    # an unknown ks_precision is an invalid argument, before any launch
    print("[gpu] deac_run with ks_precision 7", flush=True)
    case = gc.load("smoke_default")
    with _engine(case) as e:
        nout = da.lib().deac_output_length(C.byref(e.problem))
        dsf, freq = np.zeros(nout), np.zeros(nout)
        r = da.Result(0, 0.0, 0, da._ptr(dsf), da._ptr(freq), None, None, None, None)
        o = da.RunOptions(10, 0.0, 0, 7, 0)
        assert da.lib().deac_run(e.handle, da.seed_state(1), C.byref(o), C.byref(r)) == 1  # INVALID_ARGUMENT
        assert e.run(1, 10, 0.0)["generation"] == 9  # the handle still works


def test_batch_refuses_unknown_precision():
    # the batch takes FP64, FP32 and BF16X3 (tests/test_batch_mixed_gpu.py);
    # any other ks_precision is an invalid argument, refused before a launch
    print("[gpu] deac_batch_run with ks_precision 7", flush=True)
    case = gc.load("smoke_default")
    a = case["meta"]["args"]
    freq = gc.frequency_of(case, da.frequency_grid)
    p = da.make_problem(case["tau"], case["isf"], case["err"], freq, a["T"], **gc.problem_kwargs(case))
    with da.Batch([p]) as b:
        nout = da.lib().deac_output_length(C.byref(p))
        dsf, fr = np.zeros(nout), np.zeros(nout)
        r = (da.Result * 1)(da.Result(0, 0.0, 0, da._ptr(dsf), da._ptr(fr), None, None, None, None))
        o = (da.RunOptions * 1)(da.RunOptions(10, 0.0, 0, 7, 0))
        st = da.lib().deac_batch_run(b.handle, da.seed_state(1), o, r)
        assert st == 1 and "ks_precision" in da._last_message()


@pytest.mark.parametrize("T,G,P", [(12, 48, 520), (4100, 24, 6)])
def test_split_loop_fallback_shapes_match_fused(T, G, P):
    # the split kernels' paths for populations past the LDS-staged top of
    # generation (P > 512) and for ISFs past the LDS-staged residuals
    # (T > 4096) against the fused kernel, bit for bit
    print(f"[gpu] split loop fp64 vs fused, T={T} G={G} P={P}", flush=True)
    beta = 1.0 / 1.35
    t = np.linspace(0.0, beta / 2, T)
    f = 0.6 * np.exp(-2.0 * t) + 0.4 * np.exp(-9.0 * t)
    p = da.make_problem(t, f, 1e-3 * f, da.frequency_grid(G, 64.0), 1.35, population_size=P, normalize=True)
    with da.Engine(p) as e:
        fused = e.run(7, 60, 0.0)
        split = e.run(7, 60, 0.0, split_loop=True)
    for k in ("minimum", "mean", "squared_mean", "dsf", "frequency"):
        assert gc.bit_diff(fused[k], split[k]) is None, k
    assert fused["rng"] == split["rng"] and fused["generation"] == split["generation"]


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_split_loop_graphs_match_plain_launches(prec, monkeypatch):
    # the split loop's generations run as captured HIP graphs (8 generations
    # per graph, replayed across batches with plain launches and host
    # synchronizations between them); every replay's effect is checked
    # through the whole trajectory against plain launches, bit for bit
    print(f"[gpu] split loop graphs vs plain launches, {prec}", flush=True)
    t, f, s = _synthetic("isf_c1_T33.bin")
    p = da.make_problem(t, f, s, da.frequency_grid(512, 64.0), 1.35, population_size=8, normalize=True)
    kw = dict(split_loop=True) if prec == "fp64" else dict(precision=prec)
    with da.Engine(p) as e:
        graphs = e.run(5, 700, 0.0, **kw)
        again = e.run(5, 700, 0.0, **kw)
        monkeypatch.setenv("DEAC_SPLIT_GRAPH", "0")
        plain = e.run(5, 700, 0.0, **kw)
    assert graphs["profile"]["graph_launches"] >= 80 and plain["profile"]["graph_launches"] == 0
    for k in ("minimum", "mean", "squared_mean", "dsf", "frequency"):
        assert gc.bit_diff(graphs[k], plain[k]) is None, k
        assert gc.bit_diff(graphs[k], again[k]) is None, k
    assert graphs["rng"] == plain["rng"] and graphs["generation"] == plain["generation"]
    if prec == "fp64":
        fused = None
        with da.Engine(p) as e:
            fused = e.run(5, 700, 0.0)
        assert gc.bit_diff(graphs["minimum"], fused["minimum"]) is None
