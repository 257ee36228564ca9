"""Batched engine (deac_batch_*) on the GPU against the reference and the oracle.

Bar: bit-exact per instance.  Every instance of a batch must reproduce what
the reference CPU build (the golden files) or the oracle computes for the
same problem and seed: per-generation statistics, DSF, frequency grid,
generation count and the final RNG state.
"""
import os

import numpy as np
import pytest

import deac_amd as da
import golden_cases as gc
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def batch_compatible(case):
    m = case["meta"]
    a = m["args"]
    # the complete 200000-generation config-2 run is the plain engine's case
    # (one instance is a single workgroup here); config 4's long seeds run
    # inside a 256-seed batch below
    return a["P"] <= 32 and a["N"] <= 20000


def _engine_problem(case):
    a = case["meta"]["args"]
    freq = gc.frequency_of(case, da.frequency_grid)
    return da.make_problem(case["tau"], case["isf"], case["err"], freq, a["T"], **gc.problem_kwargs(case))


def _compare(got, ref, keys=("minimum", "mean", "squared_mean", "dsf", "frequency")):
    errors = []
    for k in keys:
        d = gc.bit_diff(got[k], ref[k])
        if d:
            errors.append(f"{k}: {d}")
    return errors


def test_batch_of_reference_goldens():
    # every golden case the batch kernel runs, all in one heterogeneous batch
    # (different T, G, P, spectra, models, N and stop thresholds)
    names = [n for n in gc.names() if batch_compatible(gc.load(n))]
    print(f"[gpu] batch of {len(names)} goldens: {names}", flush=True)
    cases = [gc.load(n) for n in names]
    with da.Batch([_engine_problem(c) for c in cases]) as b:
        outs = b.run([c["meta"]["args"]["seed"] for c in cases], [c["meta"]["args"]["N"] for c in cases],
                     [c["meta"]["args"]["stop"] for c in cases], True)
    bad = {}
    for n, c, o in zip(names, cases, outs):
        errors = gc.compare_outputs(o, c)
        if errors:
            bad[n] = errors
    assert not bad, bad


SYN = {
    "zeroT_T17_G37_P5": dict(T=17, G=37, wmax=8.0, temp=0.0, seeds=[9, 10],
                             kw=dict(model="normalization", spectra_type="ffull", zero_temperature=True,
                                     population_size=5, normalize=True)),
    "hyper_negw_T30_G100_P12": dict(T=30, G=100, wmax=10.0, temp=1.0, seeds=[4],
                                    kw=dict(model="hyperbolic", allow_negative=True, population_size=12)),
    "T20_G300_P32": dict(T=20, G=300, wmax=20.0, temp=1.0, seeds=[11],
                         kw=dict(population_size=32, normalize=True)),
    "T47_G129_P17": dict(T=47, G=129, wmax=30.0, temp=0.8, seeds=[3],
                         kw=dict(population_size=17, spectra_type="bfull")),
}


def _synthetic_case(spec):
    T = spec["T"]
    beta = 1.0 / spec["temp"] if spec["temp"] > 0 else 1.0
    tau = np.linspace(0, beta / 2, T)
    isf = np.exp(-1.5 * tau) + np.exp(-0.3 * tau)
    return tau, isf, 1e-3 * isf


@pytest.mark.parametrize("name", sorted(SYN))
@pytest.mark.parametrize("slice_len", [None, "7"])
def test_batch_matches_oracle(name, slice_len, monkeypatch):
    # shapes the goldens do not cover: zero temperature, odd T (16k+r rows
    # split between the matrix cores and the remainder chains), P not a
    # multiple of 16; and launches cut every 7 generations
    spec = SYN[name]
    print(f"[gpu] batch vs oracle {name} slice={slice_len}", flush=True)
    if slice_len:
        monkeypatch.setenv("DEAC_BATCH_SLICE", slice_len)
    tau, isf, err = _synthetic_case(spec)
    freq = ob.frequency_grid(spec["G"], spec["wmax"])
    N = 250
    with da.Batch([da.make_problem(tau, isf, err, freq, spec["temp"], **spec["kw"])] * len(spec["seeds"])) as b:
        outs = b.run(spec["seeds"], N, 0.0, True)
    for sd, o in zip(spec["seeds"], outs):
        r = ob.run(ob.make_problem(tau, isf, err, freq, spec["temp"], number_of_generations=N,
                                   stop_minimum_fitness=0.0, **spec["kw"]), sd)
        assert not _compare(o, r), (sd, _compare(o, r))
        assert o["rng"] == r["rng"] and o["generation"] == r["generation"]


def test_batch_matches_plain_engine_and_continues_rng():
    # the batch and deac_run agree, and each instance's returned RNG state
    # continues the reference stream (the state at the first generation not
    # evolved)
    print("[gpu] batch vs plain engine", flush=True)
    case = gc.load("smoke_stop")
    p = _engine_problem(case)
    a = case["meta"]["args"]
    with da.Engine(p) as e:
        plain = [e.run(sd, a["N"], a["stop"]) for sd in (a["seed"], 5, 6)]
    with da.Batch([p] * 3) as b:
        outs = b.run([a["seed"], 5, 6], a["N"], a["stop"], True)
    for pl, o in zip(plain, outs):
        assert not _compare(o, pl)
        assert o["rng"] == pl["rng"] and o["generation"] == pl["generation"]


def test_seed_batch_full_size():
    # 256 seeds of the G=4096, T=65, P=32 workload in one launch; seeds 1 and
    # 256 checked bit for bit against the oracle, the rest by properties
    print("[gpu] 256-seed batch", flush=True)
    t, f, s = np.fromfile(os.path.join(gc.GOLDEN_DIR, "isf_c2_T65.bin")).reshape(3, -1)
    freq = da.frequency_grid(4096, 512.0)
    N = 25
    prob = da.make_problem(t, f, s, freq, 1.35, population_size=32, normalize=True, spectra_type="bfull")
    seeds = list(range(1, 257))
    with da.Batch([prob] * 256) as b:
        outs = b.run(seeds, N, 1.0, True)
        assert b.profile["generations"] == 256 * (N - 1)
    for sd in (1, 256):
        r = ob.run(ob.make_problem(t, f, s, freq, 1.35, number_of_generations=N, population_size=32,
                                   normalize=True, stop_minimum_fitness=1.0), sd)
        assert not _compare(outs[sd - 1], r), _compare(outs[sd - 1], r)
        assert outs[sd - 1]["rng"] == r["rng"]
    finals = np.array([o["minimum_fitness"] for o in outs])
    assert len(set(finals.tolist())) > 250
    for o in outs:
        assert np.all(np.diff(o["minimum"]) <= 0)
        assert np.all(o["dsf"] >= 0)
    avg = np.mean([o["dsf"] for o in outs], axis=0)
    assert np.all(np.isfinite(avg)) and avg.shape == (2 * 4096 - 1,)


def test_seed_batch_config4_long_goldens():
    # the 256-seed config-4 batch for 2000 generations; seeds 1 and 200
    # against the reference executable's own 2000-generation runs (stop at
    # chi^2 <= 1), bit for bit: histories, DSF, frequency, generation
    print("[gpu] 256-seed batch, 2000 generations vs reference goldens (seeds 1, 200)", flush=True)
    cases = {sd: gc.load(f"c4_seed{sd}_prefix2000") for sd in (1, 200)}
    c = cases[1]
    a = c["meta"]["args"]
    prob = _engine_problem(c)
    with da.Batch([prob] * 256) as b:
        outs = b.run(list(range(1, 257)), a["N"], a["stop"], True)
    for sd, case in cases.items():
        errors = gc.compare_outputs(outs[sd - 1], case)
        assert not errors, (sd, errors)
    for o in outs:
        assert np.all(np.diff(o["minimum"]) <= 0)


def _oracle_of(x, N, mu):
    return ob.run(ob.make_problem(x["tau"], x["isf"], x["err"], ob.frequency_grid(x["G"], x["omega_max"]),
                                  mu.TEMPERATURE, model=x["model"], zero_temperature=x["zero_temperature"],
                                  population_size=x["P"], normalize=True, spectra_type="bfull",
                                  number_of_generations=N, stop_minimum_fitness=0.0), x["seed"])


def test_uneven_batch():
    # the uneven batch (tests/golden/make_uneven.py): 64 wavevectors x 16
    # seeds of mixed T, G, omega_max, model and zero-temperature kernels in
    # one deac_batch_run; a spread of 24 instances checked bit for bit
    # against the single-instance engine, the rest by properties
    import sys
    sys.path.insert(0, gc.GOLDEN_DIR)
    import make_uneven as mu
    insts = mu.instances()
    print(f"[gpu] uneven batch of {len(insts)} instances", flush=True)
    N = 8
    probs = [da.make_problem(x["tau"], x["isf"], x["err"], da.frequency_grid(x["G"], x["omega_max"]), mu.TEMPERATURE,
                             model=x["model"], zero_temperature=x["zero_temperature"], population_size=x["P"],
                             normalize=True, spectra_type="bfull") for x in insts]
    with da.Batch(probs) as b:
        outs = b.run([x["seed"] for x in insts], N, 0.0, True)
        assert b.profile["generations"] == len(insts) * (N - 1)
    for i in range(0, len(insts), 43):
        x = insts[i]
        with da.Engine(probs[i]) as e:
            ref = e.run(x["seed"], N, 0.0)
        assert not _compare(outs[i], ref), (i, x["k"], x["seed"], _compare(outs[i], ref))
        assert outs[i]["rng"] == ref["rng"]
    # and against the oracle: one instance of every (T, zero-temperature)
    # class, the largest genome of each, bit for bit
    pick = {}
    for i, x in enumerate(insts):
        key = (x["T"], x["zero_temperature"], x["model"])
        if key not in pick or x["G"] > insts[pick[key]]["G"]:
            pick[key] = i
    for key, i in sorted(pick.items()):
        x = insts[i]
        print(f"[gpu] uneven instance {i} T={x['T']} G={x['G']} {x['model']} zeroT={x['zero_temperature']} "
              "vs oracle", flush=True)
        r = _oracle_of(x, N, mu)
        assert not _compare(outs[i], r), (i, key, _compare(outs[i], r))
        assert outs[i]["rng"] == r["rng"]
    for x, o in zip(insts, outs):
        nout = x["G"] if x["zero_temperature"] else 2 * x["G"] - 1
        assert o["dsf"].shape == (nout,) and np.all(np.isfinite(o["dsf"]))
        assert np.all(np.diff(o["minimum"]) <= 0)


def _synthetic_arrays(T):
    beta = 1.0 / 1.35
    t = np.linspace(0.0, beta / 2, T)
    f = 0.6 * np.exp(-2.0 * t) + 0.4 * np.exp(-9.0 * t)
    return t, f, 1e-3 * f


def _synthetic_problem(T, G, P):
    t, f, s = _synthetic_arrays(T)
    return da.make_problem(t, f, s, da.frequency_grid(G, 64.0), 1.35, population_size=P, normalize=True)


def _synthetic_oracle(T, G, P, seed, N):
    t, f, s = _synthetic_arrays(T)
    return ob.run(ob.make_problem(t, f, s, ob.frequency_grid(G, 64.0), 1.35, population_size=P, normalize=True,
                                  number_of_generations=N, stop_minimum_fitness=0.0), seed)


def test_batch_shape_limits():
    # the largest T the batch kernel takes (16 full 16-row tiles on four
    # chain waves plus 4 rows on 4x4x4 MFMAs) runs bit-exact against
    # deac_run; one more timeslice, 65 members, or 33 members past T = 132
    # is refused
    print("[gpu] batch shape limits: T=260 vs deac_run, T=261, P=65, P=33 at T=133 refused", flush=True)
    p = _synthetic_problem(260, 96, 32)
    with da.Engine(p) as e:
        plain = [e.run(sd, 40, 0.0) for sd in (2, 3)]
    with da.Batch([p] * 2) as b:
        outs = b.run([2, 3], 40, 0.0, True)
    for sd, pl, o in zip((2, 3), plain, outs):
        assert not _compare(o, pl), _compare(o, pl)
        assert o["rng"] == pl["rng"]
        r = _synthetic_oracle(260, 96, 32, sd, 40)
        assert not _compare(o, r), _compare(o, r)
        assert o["rng"] == r["rng"]
    for T, P in ((261, 32), (65, 65), (133, 33), (143, 64)):
        with pytest.raises(da.DeacError) as err:
            da.Batch([_synthetic_problem(T, 96, P)])
        assert err.value.status == 5  # DEAC_ERR_UNSUPPORTED
    with pytest.raises(da.DeacError, match="260 timeslices"):
        da.Batch([_synthetic_problem(261, 96, 32)])
    for T, P in ((132, 64), (127, 33), (128, 48)):
        da.Batch([_synthetic_problem(T, 96, P)]).close()  # taken


@pytest.mark.parametrize("T,P", [(18, 8), (35, 32), (51, 20)])
def test_batch_remainder_rows_match_plain_engine(T, P):
    # every count of rows past the last full 16-row tile that runs on the
    # 4x4x4 remainder chains (T = 16k + 2, 3; 51 = 48 + 3 with 20 members,
    # two member tiles) against deac_run, bit for bit
    print(f"[gpu] batch remainder rows: T={T} ({T % 16} rows) P={P} vs deac_run", flush=True)
    p = _synthetic_problem(T, 80, P)
    with da.Engine(p) as e:
        plain = [e.run(sd, 50, 0.0) for sd in (4, 9)]
    with da.Batch([p] * 2) as b:
        outs = b.run([4, 9], 50, 0.0, True)
    for sd, pl, o in zip((4, 9), plain, outs):
        assert not _compare(o, pl), _compare(o, pl)
        assert o["rng"] == pl["rng"]
        r = _synthetic_oracle(T, 80, P, sd, 50)
        assert not _compare(o, r), _compare(o, r)
        assert o["rng"] == r["rng"]


def test_all_zero_rng_state_is_refused():
    # the all-zero xoshiro256+ state is a fixed point (every draw 0, donor
    # draws without end): refused before anything is launched
    print("[gpu] all-zero rng state refused", flush=True)
    p = _synthetic_problem(18, 40, 8)
    with da.Engine(p) as e:
        with pytest.raises(da.DeacError) as err:
            e.run([0, 0, 0, 0], 10, 0.0)
        assert err.value.status == 1
        ok = e.run(4, 10, 0.0)                   # the handle still works
    with da.Batch([p] * 2) as b:
        with pytest.raises(da.DeacError) as err:
            b.run([4, [0, 0, 0, 0]], 10, 0.0)
        assert err.value.status == 1
        with pytest.raises(ValueError):
            b.run([4], 10, 0.0)                  # one seed per problem
        outs = b.run([4, 9], 10, 0.0)
    assert not _compare(outs[0], ok)


def test_continue_equals_one_longer_run():
    # deac_batch_continue: a run of N1 generations continued by N2 is bit for
    # bit one run of N1 + N2 - 1 (population, fitness, CR/F and random state
    # stay on the device); histories join at the shared entry
    print("[gpu] batch continue = one longer run", flush=True)
    p = _synthetic_problem(65, 300, 32)
    q = _synthetic_problem(18, 200, 12)
    seeds = [5, 6, 7]
    with da.Batch([p, q, p]) as b:
        whole = b.run(seeds, 60, 0.0, True)
        first = b.run(seeds, 21, 0.0, True)
        rest = b.continue_run(40, 0.0, True)
        more = b.continue_run(3, 0.0, True)  # and again
    with da.Batch([p, q, p]) as b:
        whole2 = b.run(seeds, 62, 0.0, True)
    for w, f, r in zip(whole, first, rest):
        for k in ("minimum", "mean", "squared_mean"):
            assert f[k][-1] == r[k][0]
            joined = np.concatenate([f[k], r[k][1:]])
            assert gc.bit_diff(joined, w[k]) is None, k
        assert gc.bit_diff(r["dsf"], w["dsf"]) is None
        assert r["rng"] == w["rng"] and r["generation"] == 39 and w["generation"] == 59
    for w2, m in zip(whole2, more):
        assert gc.bit_diff(m["dsf"], w2["dsf"]) is None and m["rng"] == w2["rng"]
        assert m["minimum"][-1] == w2["minimum"][-1]


def test_tuned_schedule_changes_no_result():
    # deac_batch_tune times the launch schedules (streams, group order,
    # instance order) and keeps the fastest; results under any schedule are
    # those of deac_run, bit for bit
    print("[gpu] batch schedule tuning", flush=True)
    probs = [_synthetic_problem(18, 80, 8), _synthetic_problem(65, 300, 16), _synthetic_problem(130, 200, 32),
             _synthetic_problem(260, 96, 12)]
    seeds = [2, 3, 4, 5]
    with da.Batch(probs) as b:
        ref = b.run(seeds, 30, 0.0, True)
        chosen = b.tune(seeds, 10)
        assert 1 <= chosen["group_streams"] <= 3 and chosen["ms_per_generation"] > 0
        assert b.schedule["group_streams"] == chosen["group_streams"]
        for sc in ({"group_streams": 1, "group_order": 1, "instance_order": 1},
                   {"group_streams": 2, "group_order": 0, "instance_order": 0}):
            b.schedule = sc
            out = b.run(seeds, 30, 0.0, True)
            for o, r in zip(out, ref):
                assert not _compare(o, r)
                assert o["rng"] == r["rng"]
        with pytest.raises(da.DeacError):
            b.schedule = {"group_streams": 4, "group_order": 0, "instance_order": 0}
    with da.Batch(probs) as b2:  # the process keeps the tuned schedule for these shapes
        assert b2.schedule["group_streams"] == chosen["group_streams"]
        assert b2.schedule["group_order"] == chosen["group_order"]


def test_batch_moments_match_oracle():
    # moment constraints in the batch kernel (the walk wave folds each
    # member's trial values chunk by chunk in the reference's order) at the
    # seed batch's T with an odd genome (the fused last position of each
    # sign, a partial last chunk), two signs with P = 32 and one sign
    # (zero temperature) with P = 12, in one batch, against the oracle
    print("[gpu] batch with moment constraints vs oracle", flush=True)
    t, f, s = np.fromfile(f"{gc.GOLDEN_DIR}/isf_c2_T65.bin").reshape(3, -1)
    G, N = 1023, 120
    kw = dict(normalize=True, first_moment=30.0, third_moment=1e9, third_moment_error=1e8)
    shapes = [dict(P=32, zero_temperature=False), dict(P=12, zero_temperature=True)]
    probs = [da.make_problem(t, f, s, da.frequency_grid(G, 512.0), 1.35, population_size=x["P"],
                             zero_temperature=x["zero_temperature"], **kw) for x in shapes]
    with da.Batch(probs) as b:
        outs = b.run([3, 4], N, 0.0, True)
    for sd, x, o in zip((3, 4), shapes, outs):
        r = ob.run(ob.make_problem(t, f, s, ob.frequency_grid(G, 512.0), 1.35, population_size=x["P"],
                                   zero_temperature=x["zero_temperature"], number_of_generations=N,
                                   stop_minimum_fitness=0.0, **kw), sd)
        assert not _compare(o, r), (x, _compare(o, r))
        assert o["rng"] == r["rng"]


# 33 to 64 members: four 16-member tiles per workgroup (FP64, at most two
# kernel tiles per chain wave).  T = 51 puts the three remainder rows on all
# four member tiles' 4x4x4 chains; T = 132 is the largest T with four tiles
# of members (two full kernel tiles per chain wave plus four remainder rows).
WIDE = [
    (65, 300, 64, dict(normalize=True)),
    (51, 200, 40, dict(model="normalization", spectra_type="ffull", zero_temperature=True, normalize=True)),
    (132, 120, 64, dict(model="hyperbolic", allow_negative=True)),
    (128, 97, 33, dict(first_moment=30.0, third_moment=1e9, third_moment_error=1e8, normalize=True)),
    # whole 16-row tiles only and one tile per chain wave (4 block pairs of
    # K per chunk at this member count)
    (16, 64, 40, dict(normalize=True)),
    (32, 33, 48, dict(model="hyperbolic")),
    (64, 130, 64, dict(normalize=True)),
]


@pytest.mark.parametrize("T,G,P,kw", WIDE)
def test_batch_up_to_64_members_matches_oracle(T, G, P, kw):
    print(f"[gpu] batch with {P} members, T={T} G={G} {kw}", flush=True)
    t, f, s = _synthetic_arrays(T)
    temp = 0.0 if kw.get("zero_temperature") else 1.35
    N, seeds = 150, [5, 6]
    prob = da.make_problem(t, f, s, da.frequency_grid(G, 64.0), temp, population_size=P, **kw)
    with da.Batch([prob] * len(seeds)) as b:
        outs = b.run(seeds, [N, N - 37], [0.0, 0.0], True)
    for sd, n_gen, o in zip(seeds, (N, N - 37), outs):
        r = ob.run(ob.make_problem(t, f, s, ob.frequency_grid(G, 64.0), temp, population_size=P,
                                   number_of_generations=n_gen, stop_minimum_fitness=0.0, **kw), sd)
        assert not _compare(o, r), (sd, _compare(o, r))
        assert o["rng"] == r["rng"] and o["generation"] == r["generation"]


# Every FP64 kernel shape (NT 16-row kernel tiles per chain wave, NM
# 16-member tiles, remainder rows REM) against the oracle: T picks NT and
# REM (64: NT 1; 35: NT 1 + 3 rows; 90: NT 2, 10 rows padded to a tile;
# 114: NT 2 + 2 rows; 160: NT 4; 196: NT 4 + 4 rows), P picks NM (12, 28,
# 50 -> 1, 2, 4; 4 member tiles take at most 2 kernel tiles per wave)
SHAPE_SWEEP = [(T, P) for T in (64, 35, 90, 114, 160, 196) for P in (12, 28, 50) if not (T > 132 and P > 32)]


@pytest.mark.parametrize("T,P", SHAPE_SWEEP)
def test_batch_kernel_shapes_match_oracle(T, P):
    print(f"[gpu] batch kernel shape T={T} P={P}", flush=True)
    t, f, s = _synthetic_arrays(T)
    G, N, seeds = 45 + (T % 7) * 11, 60, [9, 10]
    kw = dict(normalize=(T % 2 == 0), model="hyperbolic" if P == 28 else "standard")
    prob = da.make_problem(t, f, s, da.frequency_grid(G, 48.0), 1.35, population_size=P, **kw)
    with da.Batch([prob] * len(seeds)) as b:
        outs = b.run(seeds, N, 0.0, True)
    for sd, o in zip(seeds, outs):
        r = ob.run(ob.make_problem(t, f, s, ob.frequency_grid(G, 48.0), 1.35, population_size=P,
                                   number_of_generations=N, stop_minimum_fitness=0.0, **kw), sd)
        assert not _compare(o, r), (sd, _compare(o, r))
        assert o["rng"] == r["rng"] and o["generation"] == r["generation"]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("T,G,P,kw", [
    (65, 300, 64, dict(normalize=True)),
    (16, 64, 40, dict(normalize=True)),
    (132, 120, 48, dict(model="hyperbolic", first_moment=2.0)),
    (100, 97, 33, dict(spectra_type="ffull", model="normalization", zero_temperature=True, normalize=True)),
])
def test_batch_64_members_mixed_precision(prec, T, G, P, kw):
    # 33 to 64 members on the mixed kernels (each chain wave one member tile,
    # every kernel tile): generation 0's best matches the FP64 batch to the
    # precision of the contraction, the initial fitness is the FP64
    # evaluation's, runs repeat bit for bit, and the search reaches the FP64
    # batch's quality
    print(f"[gpu] mixed batch {prec} with {P} members, T={T} G={G} {kw}", flush=True)
    t, f, s = _synthetic_arrays(T)
    temp = 0.0 if kw.get("zero_temperature") else 1.35
    prob = da.make_problem(t, f, s, da.frequency_grid(G, 64.0), temp, population_size=P, **kw)
    seeds, N = [5, 6, 7], 80
    with da.Batch([prob] * len(seeds)) as b:
        ref = b.run(seeds, N, 0.0, True, precision="fp64")
        out = b.run(seeds, N, 0.0, True, precision=prec)
        again = b.run(seeds, N, 0.0, True, precision=prec)
    for k, (o, r, a) in enumerate(zip(out, ref, again)):
        assert o["minimum"][0] == r["minimum"][0], k
        assert o["rng"] == r["rng"], k
        assert abs(o["minimum"][1] - r["minimum"][1]) <= 1e-5 * r["minimum"][1], (k, o["minimum"][1], r["minimum"][1])
        assert gc.bit_diff(o["minimum"], a["minimum"]) is None and gc.bit_diff(o["dsf"], a["dsf"]) is None, k
    fin = np.median([o["minimum_fitness"] for o in out])
    finr = np.median([r["minimum_fitness"] for r in ref])
    assert 0.3 < fin / finr < 3.0, (fin, finr)


# Per-generation renormalization (the reference's single-particle build,
# deac.cpp:1342-1364): each generation's trials are scaled to the zeroth
# moment before their contraction, in a pass of their own.
RENORM = [
    (33, 200, 8, dict(spectra_type="spfsf"), 1.0),
    (65, 301, 32, dict(spectra_type="spbsf", model="hyperbolic"), 1.0),
    (17, 90, 12, dict(spectra_type="spfsf", zero_temperature=True, first_moment=0.4), 0.0),
    (51, 128, 48, dict(spectra_type="bfull", model="normalization", third_moment=2.0, third_moment_error=0.5), 1.0),
]


@pytest.mark.parametrize("T,G,P,kw,temp", RENORM)
@pytest.mark.parametrize("slice_len", [None, "11"])
def test_batch_renormalization_matches_oracle(T, G, P, kw, temp, slice_len, monkeypatch):
    print(f"[gpu] batch with renormalization T={T} G={G} P={P} {kw} slice={slice_len}", flush=True)
    if slice_len:
        monkeypatch.setenv("DEAC_BATCH_SLICE", slice_len)
    t, f, s = _synthetic_arrays(T)
    if kw["spectra_type"].startswith("sp"):
        f, s = -f, s  # single-particle ISFs are negative
    N, seeds = 120, [7, 8, 9]
    prob = da.make_problem(t, f, s, da.frequency_grid(G, 32.0), temp, normalize=True, renormalize=True,
                           population_size=P, **kw)
    with da.Batch([prob] * len(seeds)) as b:
        outs = b.run(seeds, [N, N - 29, N], [0.0, 0.0, 0.0], True)
    for sd, n_gen, o in zip(seeds, (N, N - 29, N), outs):
        r = ob.run(ob.make_problem(t, f, s, ob.frequency_grid(G, 32.0), temp, normalize=True, renormalize=True,
                                   population_size=P, number_of_generations=n_gen, stop_minimum_fitness=0.0,
                                   **kw), sd)
        assert not _compare(o, r), (sd, _compare(o, r))
        assert o["rng"] == r["rng"] and o["generation"] == r["generation"]


# mixed-kernel shapes where renormalization moves
# generation 0's best fitness by more than 1e-3 (it barely does for the
# hyperbolic model); the zero-temperature case has a first moment too
MIXED_RENORM = [
    (33, 200, 8, dict(spectra_type="spfsf"), 1.0),
    (65, 301, 32, dict(spectra_type="spbsf"), 1.0),
    (17, 90, 12, dict(spectra_type="spfsf", zero_temperature=True, first_moment=0.4), 0.0),
    (50, 150, 24, dict(spectra_type="spfsf"), 1.0),
    (51, 128, 48, dict(spectra_type="spfsf"), 1.0),  # four member tiles
]


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
@pytest.mark.parametrize("T,G,P,kw,temp", MIXED_RENORM)
def test_batch_renormalization_mixed_precision(prec, T, G, P, kw, temp):
    # the mixed kernels' renormalization pass: the workers sum each trial's
    # zeroth moment in FP64 (their own order), and the evolution pass scales
    # the FP64 trial before its change from the parent goes to the
    # contraction.  Generation 0's best matches the FP64 batch to the
    # precision of the contraction where renormalization moves it by far
    # more; runs repeat bit for bit.
    print(f"[gpu] mixed batch {prec} with renormalization T={T} G={G} P={P} {kw}", flush=True)
    t, f, s = _synthetic_arrays(T)
    if kw["spectra_type"].startswith("sp"):
        f, s = -f, s

    def problem(renorm):
        return da.make_problem(t, f, s, da.frequency_grid(G, 32.0), temp, normalize=True, renormalize=renorm,
                               population_size=P, **kw)
    N, seeds = 60, [7, 8, 9]
    with da.Batch([problem(True)] * len(seeds)) as b:
        ref = b.run(seeds, N, 0.0, True, precision="fp64")
        out = b.run(seeds, N, 0.0, True, precision=prec)
        again = b.run(seeds, N, 0.0, True, precision=prec)
    with da.Batch([problem(False)] * len(seeds)) as b:
        bare = b.run(seeds, N, 0.0, True, precision="fp64")
    assert all(abs(z["minimum"][1] - r["minimum"][1]) > 1e-3 * r["minimum"][1] for z, r in zip(bare, ref))
    for k, (o, r, a) in enumerate(zip(out, ref, again)):
        assert o["minimum"][0] == r["minimum"][0], k
        assert o["rng"] == r["rng"], k
        assert abs(o["minimum"][1] - r["minimum"][1]) <= 1e-5 * r["minimum"][1], (k, o["minimum"][1], r["minimum"][1])
        assert gc.bit_diff(o["minimum"], a["minimum"]) is None and gc.bit_diff(o["dsf"], a["dsf"]) is None, k
