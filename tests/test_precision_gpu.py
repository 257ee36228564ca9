"""K.S precision sweep (deac_model_isf) on the GPU.

FP64 must be the engine's exact evaluation: bit for bit the reference's FMA
chains (checked against the oracle's chain) and, for the reference's own final
best genome of the config-2 prefix, its logged fitness.  FP32 and BF16x3 are
bounded against FP64 (tolerances in the tests) and their |dchi^2|/chi^2 is
printed for the record.
"""
import numpy as np
import pytest

import deac_amd as da
import golden_cases as gc
import oracle_binding as ob

pytestmark = pytest.mark.gpu


def _oracle_model_isf(problem, genomes):
    """E[n][t] as the reference's chain: K+ then K- in one accumulator."""
    T, G = problem.number_of_timeslices, problem.genome_size
    kp, km = np.zeros(T * G), np.zeros(T * G)
    ob.lib().dor_kernel_terms(problem, ob.ptr(kp), ob.ptr(km))
    nsign = 1 if problem.zero_temperature else 2
    g = genomes.reshape(-1, nsign * G)
    E = np.zeros((g.shape[0], T))
    for n in range(g.shape[0]):
        sp = np.ascontiguousarray(g[n, :G])
        sm = np.ascontiguousarray(g[n, G:]) if nsign == 2 else None
        for t in range(T):
            acc = ob.lib().dor_fma_chain(0.0, ob.ptr(kp[t * G:(t + 1) * G]), ob.ptr(sp), G)
            if nsign == 2:
                acc = ob.lib().dor_fma_chain(acc, ob.ptr(km[t * G:(t + 1) * G]), ob.ptr(sm), G)
            E[n, t] = acc
    return E


@pytest.mark.parametrize("zero_t", [False, True])
def test_fp64_is_the_reference_chain(zero_t):
    print(f"[gpu] model_isf fp64 vs oracle chain, zero_t={zero_t}", flush=True)
    T, G, P, seed, N = 17, 37, 6, 4, 30
    tau = np.linspace(0.0, 0.37, T)
    isf = np.exp(-2.0 * tau) + np.exp(-2.0 * (0.74 - tau))
    err = 1e-3 * isf
    freq = ob.frequency_grid(G, 8.0)
    op = ob.make_problem(tau, isf, err, freq, 0.0 if zero_t else 1.35, spectra_type="bfull", zero_temperature=zero_t,
                         number_of_generations=N, population_size=P, normalize=True, stop_minimum_fitness=0.0)
    pop = ob.run(op, seed, want_population=True)["population"].reshape(P, -1)
    ref = _oracle_model_isf(op, pop)
    dp = da.make_problem(tau, isf, err, da.frequency_grid(G, 8.0), 0.0 if zero_t else 1.35, spectra_type="bfull",
                         zero_temperature=zero_t, population_size=P, normalize=True)
    E, chi, _ = da.model_isf(dp, pop, "fp64")
    assert np.array_equal(E.view(np.uint64), ref.view(np.uint64))
    for n in range(P):
        c = ob.lib().dor_chi_square(ob.ptr(isf), ob.ptr(np.ascontiguousarray(ref[n])), ob.ptr(err), T) / T
        assert chi[n] == c


def _genome_from_dsf(dsf, G):
    # standard model, bfull (deac.cpp:1594-1675): dsf[G-1+i] = S+[i],
    # dsf[G-1-i] = S-[i]
    return np.concatenate([dsf[G - 1:], dsf[G - 1::-1]])


def test_reference_best_genome_fitness_and_precisions():
    # the reference's final best genome of the config-2 prefix (2000
    # generations): FP64 gives its logged fitness exactly; FP32 and BF16x3
    # are measured against it
    case = gc.load("c2_prefix2000")
    a = case["meta"]["args"]
    G = a["M"]
    dp = da.make_problem(case["tau"], case["isf"], case["err"], gc.frequency_of(case, da.frequency_grid), a["T"],
                         **gc.problem_kwargs(case))
    genome = _genome_from_dsf(case["dsf"], G)
    _, chi64, _ = da.model_isf(dp, genome, "fp64")
    assert chi64[0] == case["minimum"][-1]
    for prec, tol in (("fp32", 1e-5), ("bf16x3", 1e-5)):
        E, chi, ms = da.model_isf(dp, genome, prec)
        rel = abs(chi[0] - chi64[0]) / chi64[0]
        print(f"[gpu] config-2 best genome ({chi64[0]:.6g}): {prec} |dchi2|/chi2 = {rel:.3e}", flush=True)
        assert rel < tol


@pytest.mark.parametrize("prec,tol", [("fp32", 5e-6), ("bf16x3", 3e-5)])
def test_mixed_precision_model_isf_error(prec, tol):
    # positive genomes and kernels: the relative error of every E element
    # stays within the bound of the format -- FP32 inputs (2^-24) with FP32
    # accumulation over 1024 positions; for BF16x3 the two-bfloat16 split
    # represents each input to about 2^-16
    rng = np.random.default_rng(11)
    T, G, P = 65, 512, 48
    tau = np.linspace(0.0, 0.37, T)
    isf = np.exp(-3.0 * tau) + np.exp(-3.0 * (0.74 - tau))
    dp = da.make_problem(tau, isf, 1e-4 * isf, da.frequency_grid(G, 64.0), 1.35, spectra_type="bfull",
                         population_size=P, normalize=True)
    genomes = rng.uniform(0.0, 1.0, (P, 2 * G)) / G
    E64, _, _ = da.model_isf(dp, genomes, "fp64")
    E, _, _ = da.model_isf(dp, genomes, prec)
    rel = np.max(np.abs(E - E64) / np.abs(E64))
    print(f"[gpu] {prec}: max relative E error {rel:.3e} (L = {2 * G})", flush=True)
    assert rel < tol


@pytest.mark.parametrize("prec,tol_isf,tol_chi", [("fp32", 5e-6, 5e-3), ("bf16x3", 1e-5, 2e-2)])
def test_mixed_precision_at_config3_shape(prec, tol_isf, tol_chi):
    # K.S at the precision sweep's own shape: T = 257, G = 16384 per sign,
    # hyperbolic model, bfull.  64 genomes near the solution (the true
    # spectrum on the genome grid as 2 S(w) e^{beta w / 2} on both signs,
    # 0.1 % multiplicative noise): every model-ISF element within tol_isf
    # (relative) of FP64, and |dchi^2|/chi^2 within tol_chi (the residuals
    # are (O - E) / sigma with sigma = 1e-4 F, so a relative E error d moves
    # each by about d 1e4)
    import os
    import sys
    sys.path.insert(0, gc.GOLDEN_DIR)
    import make_inputs as mi
    tau, f, s = np.fromfile(os.path.join(gc.GOLDEN_DIR, "isf_c3_T257.bin")).reshape(3, -1)
    G, temp, wmax = 16384, 1.35, 512.0
    beta = 1.0 / temp
    w = da.frequency_grid(G, wmax)
    grid = np.linspace(0.0, wmax, 65536)
    norm = np.trapezoid(sum(wt * mi.gaussian(grid, mu, sd) for mu, sd, wt in mi.CONFIG2_PEAKS), grid)
    S = sum(wt * mi.gaussian(w, mu, sd) for mu, sd, wt in mi.CONFIG2_PEAKS) / norm
    g0 = np.concatenate([2.0 * S * np.exp(0.5 * beta * w)] * 2)
    pop = g0 * (1.0 + 1e-3 * np.random.default_rng(3).standard_normal((64, 2 * G)))
    dp = da.make_problem(tau, f, s, w, temp, model="hyperbolic", spectra_type="bfull", population_size=64,
                         normalize=True)
    E64, c64, _ = da.model_isf(dp, pop, "fp64")
    E, c, _ = da.model_isf(dp, pop, prec)
    rel_isf = np.max(np.abs(E - E64) / np.abs(E64))
    rel_chi = np.abs(c - c64) / c64
    print(f"[gpu] T=257 G=16384 {prec}: max rel E error {rel_isf:.3e}, |dchi2|/chi2 median "
          f"{np.median(rel_chi):.3e} max {np.max(rel_chi):.3e} (chi2 ~ {np.median(c64):.3g})", flush=True)
    assert rel_isf < tol_isf
    assert np.max(rel_chi) < tol_chi
