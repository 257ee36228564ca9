"""deac_batch_run_noalloc: deac_batch_run without device allocation (the
statistics staged in the library's per-GPU device array), called through the
C ABI and checked against deac_batch_run on the same inputs -- bit for bit,
on two calls in a row, the second with a smaller input -- and its refusals
(histories beyond the array, mixed tiles not yet built) before any launch."""
import ctypes as C

import numpy as np
import pytest

import deac_amd as da
import golden_cases as gc

pytestmark = pytest.mark.gpu

SCRATCH_DOUBLES = 1 << 22


def _problem(T, G, P):
    beta = 1.0 / 1.35
    t = np.linspace(0.0, beta / 2, T)
    f = 0.6 * np.exp(-2.0 * t) + 0.4 * np.exp(-9.0 * t)
    return da.make_problem(t, f, 1e-3 * f, da.frequency_grid(G, 40.0), 1.35, population_size=P, normalize=True)


def _call(entry, batch, seeds, N, prec=0, track=1):
    """One call of a deac_batch_run-shaped entry point; returns (status, outputs)."""
    n = len(seeds)
    states = (C.c_uint64 * (4 * n))()
    for i, sd in enumerate(seeds):
        for k, w in enumerate(da.seed_state(sd)):
            states[4 * i + k] = w
    opts = (da.RunOptions * n)(*[da.RunOptions(N, 0.0, track, prec, 0) for _ in range(n)])
    res = (da.Result * n)()
    outs = []
    for i, p in enumerate(batch.problems):
        nout = da.lib().deac_output_length(C.byref(p))
        o = {k: np.zeros(m) for k, m in (("dsf", nout), ("frequency", nout), ("mean", N), ("minimum", N),
                                          ("squared_mean", N), ("fitness", p.population_size))}
        res[i] = da.Result(0, 0.0, 0, *(da._ptr(o[k]) for k in ("dsf", "frequency", "mean", "minimum",
                                                                 "squared_mean", "fitness")))
        outs.append(o)
    st = entry(batch.handle, states, opts, res)
    for i, o in enumerate(outs):
        o["generation"] = res[i].generation
        o["minimum_fitness"] = res[i].minimum_fitness
        o["rng"] = [states[4 * i + k] for k in range(4)]
    return st, outs


def _same(a, b):
    for x, y in zip(a, b):
        for k in ("dsf", "frequency", "mean", "minimum", "squared_mean", "fitness"):
            assert gc.bit_diff(x[k], y[k]) is None, k
        assert x["generation"] == y["generation"] and x["rng"] == y["rng"]
        assert x["minimum_fitness"] == y["minimum_fitness"]


def test_noalloc_matches_batch_run_twice():
    print("[gpu] deac_batch_run_noalloc vs deac_batch_run: N=300 then N=120", flush=True)
    L = da.lib()
    probs = [_problem(33, 300, 16), _problem(65, 200, 32), _problem(18, 120, 8)]
    seeds = [3, 4, 5]
    with da.Batch(probs) as b:
        for N in (300, 120):  # the second call smaller
            st, ref = _call(L.deac_batch_run, b, seeds, N)
            assert st == 0
            st, got = _call(L.deac_batch_run_noalloc, b, seeds, N)
            assert st == 0, da._last_message()
            _same(got, ref)


def test_noalloc_mixed_precision_and_refusals():
    print("[gpu] deac_batch_run_noalloc: mixed precision, refusals", flush=True)
    L = da.lib()
    probs = [_problem(65, 256, 32)] * 2
    seeds = [7, 8]
    with da.Batch(probs) as b:
        # the FP32 tiles are not built yet: refused, nothing launched
        st, _ = _call(L.deac_batch_run_noalloc, b, seeds, 50, prec=1)
        assert st == 5 and "first deac_batch_run" in da._last_message()
        st, ref = _call(L.deac_batch_run, b, seeds, 50, prec=1)
        assert st == 0
        st, got = _call(L.deac_batch_run_noalloc, b, seeds, 50, prec=1)
        assert st == 0
        _same(got, ref)
        # histories beyond the scratch array: refused before any launch
        N = SCRATCH_DOUBLES // 6 + 1
        st, _ = _call(L.deac_batch_run_noalloc, b, seeds, N, prec=0)
        assert st == 1 and "at most" in da._last_message()
        # without statistics the same N needs no scratch (kept short here)
        st, got = _call(L.deac_batch_run_noalloc, b, seeds, 40, prec=0, track=0)
        assert st == 0


def test_noalloc_fold_kernel_moments_and_stacked_populations():
    # the detailed-balance fold kernel, moment terms on the mixed kernels at
    # four member tiles, and FP64 populations stacked in one workgroup: the
    # same bits as deac_batch_run, two calls in a row, the second smaller
    print("[gpu] deac_batch_run_noalloc: bdsf, mixed moments, stacked", flush=True)
    L = da.lib()
    beta = 1.0 / 1.35

    def prob(T, G, P, **kw):
        t = np.linspace(0.0, beta / 2, T)
        f = 0.6 * np.exp(-2.0 * t) + 0.4 * np.exp(-9.0 * t)
        return da.make_problem(t, f, 1e-3 * f, da.frequency_grid(G, 40.0), 1.35, population_size=P, normalize=True,
                               **kw)
    probs = [prob(65, 200, 24, spectra_type="bdsf", use_negative_first_moment=True),
             prob(40, 150, 48, first_moment=2.0), prob(33, 100, 6), prob(33, 100, 6)]
    seeds = [3, 4, 5, 6]
    with da.Batch(probs) as b:
        sc = b.schedule
        sc["pack"] = 4
        b.schedule = sc
        for prec in (0, 1):  # FP64 (fold kernel, stacked pair); FP32 (mixed kernels)
            for N in (90, 40):
                st, ref = _call(L.deac_batch_run, b, seeds, N, prec=prec)
                assert st == 0, da._last_message()
                st, got = _call(L.deac_batch_run_noalloc, b, seeds, N, prec=prec)
                assert st == 0, da._last_message()
                _same(got, ref)
