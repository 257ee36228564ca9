"""Stacked populations in the batch kernel: FP64 instances of one problem whose
populations leave room in the 16 member rows of their kernel shape (at most
8 members) share a workgroup, up to 4 of them.  Each instance keeps its own
random stream, donors, statistics, stop test and state; the results are those
of deac_run and the oracle, bit for bit, whatever the packing.  The automatic
setting packs only batches larger than the device's compute units, so these
tests ask for packing explicitly (schedule "pack").
"""
import numpy as np
import pytest

import deac_amd as da
import golden_cases as gc
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _arrays(T):
    tau = np.linspace(0.0, 0.37, T)
    isf = 0.6 * np.exp(-2.0 * tau) + 0.4 * np.exp(-0.5 * tau)
    return tau, isf, 2e-3 * isf


KW = {
    "bfull_P8": dict(T=33, G=200, temp=1.35, kw=dict(population_size=8, normalize=True)),
    "zeroT_P5": dict(T=17, G=90, temp=0.0, kw=dict(population_size=5, spectra_type="ffull", zero_temperature=True,
                                                   model="normalization", normalize=True)),
    "negw_P4": dict(T=21, G=150, temp=1.0, kw=dict(population_size=4, allow_negative=True, model="hyperbolic")),
    "moments_P6": dict(T=25, G=120, temp=1.0, kw=dict(population_size=6, first_moment=1.0, third_moment=3.0,
                                                      third_moment_error=0.5)),
    "renorm_P8": dict(T=33, G=150, temp=1.0, kw=dict(population_size=8, spectra_type="bfull", renormalize=True,
                                                     normalize=True)),
    "bigT_P4": dict(T=150, G=80, temp=1.35, kw=dict(population_size=4, model="hyperbolic")),
}


def _problems(name):
    spec = KW[name]
    tau, isf, err = _arrays(spec["T"])
    freq = da.frequency_grid(spec["G"], 24.0)
    prob = da.make_problem(tau, isf, err, freq, spec["temp"], **spec["kw"])

    def oracle(seed, N, stop):
        return ob.run(ob.make_problem(tau, isf, err, ob.frequency_grid(spec["G"], 24.0), spec["temp"],
                                      number_of_generations=N, stop_minimum_fitness=stop, **spec["kw"]), seed)
    return prob, oracle


def _pack(b, pack):
    sc = b.schedule
    sc["pack"] = pack
    b.schedule = sc


def _compare(got, ref):
    errors = []
    for k in ("minimum", "mean", "squared_mean", "dsf", "frequency"):
        d = gc.bit_diff(got[k], ref[k])
        if d:
            errors.append(f"{k}: {d}")
    if got["rng"] != ref["rng"]:
        errors.append("rng")
    if got["generation"] != ref["generation"]:
        errors.append(f"generation {got['generation']} != {ref['generation']}")
    return errors


@pytest.mark.parametrize("name", sorted(KW))
@pytest.mark.parametrize("slice_len", [None, "13"])
def test_packed_instances_match_oracle(name, slice_len, monkeypatch):
    # 9 seeds of one problem (packs of up to 16 // P plus a partial one), with
    # different generation counts and one instance whose stop test fires
    # while its workgroup's other instances go on
    print(f"[gpu] packed batch {name} slice={slice_len}", flush=True)
    if slice_len:
        monkeypatch.setenv("DEAC_BATCH_SLICE", slice_len)
    prob, oracle = _problems(name)
    P = KW[name]["kw"]["population_size"]
    seeds = list(range(21, 30))
    Ns = [120, 120, 97, 120, 120, 64, 120, 120, 111]
    free = oracle(seeds[1], 120, 0.0)
    stops = [0.0] * len(seeds)
    stops[1] = float(free["minimum"][70])  # fires at generation 70 or before
    with da.Batch([prob] * len(seeds)) as b:
        assert b.schedule["pack"] == 0
        _pack(b, 4)
        outs = b.run(seeds, Ns, stops, True)
        prof = b.profile
    refs = [oracle(sd, N, st) for sd, N, st in zip(seeds, Ns, stops)]
    assert refs[1]["generation"] < 119
    bad = {sd: _compare(o, r) for sd, o, r in zip(seeds, outs, refs) if _compare(o, r)}
    assert not bad, bad
    q = min(4, 16 // P)
    if not slice_len:
        # the first launch holds ceil(9 / q) workgroups
        assert prof["evolve_workgroups"] >= -(-len(seeds) // q)
        assert prof["evolve_workgroups"] < len(seeds) * max(1, prof["evolve_launches"])


def test_pack_setting_changes_no_result():
    # pack 1 (one instance per workgroup), 2, 4 and automatic (five
    # instances: no packing) give the same bits
    print("[gpu] packed batch: pack settings", flush=True)
    prob, oracle = _problems("negw_P4")
    seeds = [3, 4, 5, 6, 7]
    results, groups = [], []
    with da.Batch([prob] * len(seeds)) as b:
        for pack in (1, 2, 4, 0):
            _pack(b, pack)
            results.append(b.run(seeds, 80, 0.0, True))
            groups.append(b.profile["evolve_workgroups"])
        with pytest.raises(da.DeacError):
            _pack(b, 5)
    assert groups == [5, 3, 2, 5]
    for outs in results[1:]:
        for a, c in zip(results[0], outs):
            assert not _compare(c, a)
    r = oracle(7, 80, 0.0)
    assert not _compare(results[2][4], r)
    assert not _compare(results[1][4], r)


def test_packed_continue_equals_one_longer_run():
    # a continuation repacks the instances by the generation each reached
    print("[gpu] packed batch: continue", flush=True)
    prob, _ = _problems("bfull_P8")
    seeds = [8, 9, 10]
    with da.Batch([prob] * 3) as b:
        _pack(b, 4)
        whole = b.run(seeds, 70, 0.0, True)
        first = b.run(seeds, [30, 41, 30], 0.0, True)
        rest = b.continue_run([41, 30, 41], 0.0, True)
    for w, f, r in zip(whole, first, rest):
        for k in ("minimum", "mean", "squared_mean"):
            joined = np.concatenate([f[k], r[k][1:]])
            assert gc.bit_diff(joined, w[k]) is None, k
        assert gc.bit_diff(r["dsf"], w["dsf"]) is None
        assert r["rng"] == w["rng"] and f["generation"] + r["generation"] == w["generation"]


def test_packed_next_to_mixed_and_other_problems():
    # packed FP64 instances beside mixed-precision instances of the same
    # problem (never packed) and another problem's instances
    print("[gpu] packed batch: mixed company", flush=True)
    prob, oracle = _problems("bfull_P8")
    prob2, oracle2 = _problems("zeroT_P5")
    probs = [prob, prob, prob2, prob, prob2, prob]
    seeds = [1, 2, 3, 4, 5, 6]
    precs = ["fp64", "fp32", "fp64", "fp64", "fp64", "bf16x3"]
    with da.Batch(probs) as b:
        _pack(b, 4)
        outs = b.run(seeds, 90, 0.0, True, precision=precs)
    for i in (0, 3):
        assert not _compare(outs[i], oracle(seeds[i], 90, 0.0)), i
    for i in (2, 4):
        assert not _compare(outs[i], oracle2(seeds[i], 90, 0.0)), i
    for i in (1, 5):
        r = oracle(seeds[i], 90, 0.0)
        assert outs[i]["generation"] == 89
        assert np.isfinite(outs[i]["minimum"]).all()
        assert outs[i]["minimum"][-1] < 10 * r["minimum"][-1] + 1.0
