"""Mixed-precision K.S in the batch kernel (deac_batch_run with ks_precision
FP32 or BF16X3) on the GPU.

The delta form keeps each member's model ISF in FP64 and adds K.(trial -
parent) at the mixed precision; the population is re-evaluated exactly in
FP64 before the run, every refresh slice and at the end.  Bars, each written
in its test:
  - the reported fitness is the exact FP64 chi^2 of the reported genome
    (relative difference <= 1e-12, well inside a 1e-6 requirement);
  - generation 0 is bit-identical to the FP64 batch (it is an exact FP64
    evaluation of the same initial population), and so is every returned
    RNG state (the draws do not depend on fitness values);
  - after one generation the best fitness agrees with the FP64 batch's to
    the contraction's error (relative 1e-5).
"""
import os

import numpy as np
import pytest

import deac_amd as da
import golden_cases as gc

pytestmark = pytest.mark.gpu

G4 = 4096


def _config4_problem():
    t, f, s = np.fromfile(os.path.join(gc.GOLDEN_DIR, "isf_c2_T65.bin")).reshape(3, -1)
    return da.make_problem(t, f, s, da.frequency_grid(G4, 512.0), 1.35, population_size=32, normalize=True,
                           spectra_type="bfull")


def _genome_of(dsf, G):
    # standard model, bfull (deac.cpp:1594-1675): dsf[G-1+i] = S+[i], dsf[G-1-i] = S-[i]
    return np.concatenate([dsf[G - 1:], dsf[G - 1::-1]])


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_seed_batch_against_fp64(prec, monkeypatch):
    # 64 seeds of the G=4096, T=65, P=32 workload for 300 generations, with
    # the exact re-evaluation every 128 generations
    print(f"[gpu] mixed seed batch {prec}: 64 seeds x 300 generations vs FP64", flush=True)
    monkeypatch.setenv("DEAC_BATCH_REFRESH", "128")
    p = _config4_problem()
    seeds = list(range(1, 65))
    N = 300
    with da.Batch([p] * len(seeds)) as b:
        ref = b.run(seeds, N, 1.0, True, precision="fp64")
        out = b.run(seeds, N, 1.0, True, precision=prec)
    for sd, o, r in zip(seeds, out, ref):
        assert o["minimum"][0] == r["minimum"][0] and o["mean"][0] == r["mean"][0], sd
        assert o["rng"] == r["rng"], sd
        assert abs(o["minimum"][1] - r["minimum"][1]) <= 1e-5 * r["minimum"][1], sd
        assert o["generation"] == r["generation"]
        assert np.all(o["dsf"] >= 0) and np.all(np.isfinite(o["dsf"]))
        # best fitness never rises by more than the contraction's error (the
        # exact re-evaluations may move it either way by that much)
        m = o["minimum"]
        assert np.all(m[1:] <= m[:-1] * (1 + 1e-5)), sd
    for k in (0, 17, 63):
        g = _genome_of(out[k]["dsf"], G4)
        _, chi, _ = da.model_isf(p, g, "fp64")
        assert abs(chi[0] - out[k]["minimum_fitness"]) <= 1e-12 * chi[0], (k, chi[0], out[k]["minimum_fitness"])
    fin = np.array([o["minimum_fitness"] for o in out])
    finr = np.array([r["minimum_fitness"] for r in ref])
    print(f"[gpu] {prec}: final chi2 median {np.median(fin):.6g} (FP64 {np.median(finr):.6g})", flush=True)
    # a different trajectory of the same search: the same quality of result
    assert 0.5 < np.median(fin) / np.median(finr) < 2.0


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_shapes_against_fp64(prec):
    # one member tile and two; odd T (rows past the last tile are zero rows);
    # zero temperature and the normalization model; genomes not a multiple of
    # the chunk; the largest T.  Tile pairs per chain wave: 1 (T = 18, 33),
    # 2 (50, 100), 3 (65, 180), 5 (129, 257), 9 (260)
    print(f"[gpu] mixed batch shapes {prec}", flush=True)
    cases = []
    for T, G, P, kw in ((18, 300, 8, {}), (33, 512, 16, dict(model="normalization")),
                        (65, 700, 24, dict(zero_temperature=True)), (129, 1000, 32, {}),
                        (257, 2048, 16, {}), (260, 96, 32, dict(model="hyperbolic")),
                        (50, 333, 30, {}), (100, 256, 12, dict(model="hyperbolic")), (180, 129, 10, {})):
        beta = 1.0 / 1.35
        t = np.linspace(0.0, beta / 2, T)
        f = 0.6 * np.exp(-2.0 * t) + 0.4 * np.exp(-9.0 * t)
        cases.append(da.make_problem(t, f, 1e-3 * f, da.frequency_grid(G, 40.0), 1.35, population_size=P,
                                     normalize=True, **kw))
    seeds = [3 + k for k in range(len(cases))]
    N = 40
    with da.Batch(cases) as b:
        ref = b.run(seeds, N, 0.0, True, precision="fp64")
        out = b.run(seeds, N, 0.0, True, precision=prec)
    for k, (o, r) in enumerate(zip(out, ref)):
        assert o["minimum"][0] == r["minimum"][0], k
        assert o["rng"] == r["rng"], k
        assert abs(o["minimum"][1] - r["minimum"][1]) <= 1e-5 * r["minimum"][1], (k, o["minimum"][1], r["minimum"][1])
        assert np.all(np.isfinite(o["dsf"]))


def test_mixed_and_fp64_instances_in_one_batch():
    # per-instance precision: FP64 instances stay bit-exact beside mixed ones
    print("[gpu] FP64 and mixed instances in one batch", flush=True)
    p = _config4_problem()
    seeds = [1, 2, 3, 4, 5, 6]
    precs = ["fp64", "fp32", "bf16x3", "fp64", "bf16x3", "fp32"]
    with da.Batch([p] * len(seeds)) as b:
        ref = b.run(seeds, 30, 1.0, True, precision="fp64")
        out = b.run(seeds, 30, 1.0, True, precision=precs)
    for o, r, pr in zip(out, ref, precs):
        if pr == "fp64":
            for k in ("minimum", "mean", "squared_mean", "dsf"):
                assert gc.bit_diff(o[k], r[k]) is None, k
        else:
            assert o["minimum"][0] == r["minimum"][0]
        assert o["rng"] == r["rng"]


def test_batch_refuses_the_split_loop():
    # the split loop is deac_run's; the batch reports it with a message
    import ctypes as C
    p = _config4_problem()
    with da.Batch([p]) as b:
        o = [np.zeros(2 * G4 - 1), np.zeros(2 * G4 - 1)]
        res = (da.Result * 1)(da.Result(0, 0.0, 0, da._ptr(o[0]), da._ptr(o[1]), None, None, None, None))
        opts = (da.RunOptions * 1)(da.RunOptions(5, 1.0, 0, 1, 1))
        states = (C.c_uint64 * 4)(*da.seed_state(1))
        st = da.lib().deac_batch_run(b.handle, states, opts, res)
        assert st == 5 and "split loop" in da._last_message()


# (T, G, P, problem options, ISF error / ISF): errors large enough that the
# first-moment term alone moves the fitness
MOMENTS = [
    (65, 700, 24, dict(first_moment=3.0, third_moment=40.0, third_moment_error=2.0, normalize=True), 1e-3),
    (33, 256, 8, dict(first_moment=3.0), 0.5),
    (129, 333, 32, dict(third_moment=40.0, third_moment_error=0.5, model="hyperbolic"), 1e-3),
    (40, 300, 16, dict(first_moment=0.2, zero_temperature=True, normalize=True), 1e-3),
]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("T,G,P,kw,e", MOMENTS)
def test_mixed_batch_with_moment_constraints(prec, T, G, P, kw, e):
    # the mixed kernel sums each trial's first and third moments in FP64 (its
    # own order): generation 0's selections match the FP64 batch's to the
    # precision of the contraction, the initial and final fitness are the
    # FP64 evaluation's, and the runs repeat bit for bit
    print(f"[gpu] mixed batch {prec} with moments T={T} G={G} P={P} {kw}", flush=True)
    beta = 1.0 / 1.35
    t = np.linspace(0.0, beta / 2, T)
    f = 0.6 * np.exp(-2.0 * t) + 0.4 * np.exp(-9.0 * t)
    temp = 0.0 if kw.get("zero_temperature") else 1.35
    p = da.make_problem(t, f, e * f, da.frequency_grid(G, 40.0), temp, population_size=P, **kw)
    seeds, N = [3, 4, 5, 6], 60
    with da.Batch([p] * len(seeds)) as b:
        ref = b.run(seeds, N, 0.0, True, precision="fp64")
        out = b.run(seeds, N, 0.0, True, precision=prec)
        again = b.run(seeds, N, 0.0, True, precision=prec)
    # the moment terms matter at this tolerance: without them the fitness of
    # the best trial after generation 0 moves by far more than 1e-5
    plain_kw = {k: v for k, v in kw.items() if "moment" not in k}
    p0 = da.make_problem(t, f, e * f, da.frequency_grid(G, 40.0), temp, population_size=P, **plain_kw)
    with da.Batch([p0] * len(seeds)) as b:
        bare = b.run(seeds, N, 0.0, True, precision="fp64")
    assert all(abs(z["minimum"][1] - r["minimum"][1]) > 1e-3 * r["minimum"][1] for z, r in zip(bare, ref))
    for k, (o, r, a) in enumerate(zip(out, ref, again)):
        assert o["minimum"][0] == r["minimum"][0], k
        assert o["rng"] == r["rng"], k
        assert abs(o["minimum"][1] - r["minimum"][1]) <= 1e-5 * r["minimum"][1], (k, o["minimum"][1], r["minimum"][1])
        assert gc.bit_diff(o["minimum"], a["minimum"]) is None and gc.bit_diff(o["dsf"], a["dsf"]) is None, k
        m = o["minimum"]
        assert np.all(m[1:] <= m[:-1] * (1 + 1e-5)), k
    fin = np.median([o["minimum_fitness"] for o in out])
    finr = np.median([r["minimum_fitness"] for r in ref])
    assert 0.3 < fin / finr < 3.0, (fin, finr)
