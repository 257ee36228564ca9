"""Independent pins of the oracle beyond the golden vectors: published
known answers of the generators, exact-arithmetic restatements of the
reductions on tiny inputs, closed forms of the kernel matrix, and invariants
of differential evolution that a dropped term, a wrong sign or a transposed
operand would break.
"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_binding as ob

L = ob.lib()


def fma_exact(a, b, c):
    """Correctly rounded a*b+c via exact rational arithmetic."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def arr(x):
    return np.ascontiguousarray(x, dtype=np.float64)


# ---------------------------------------------------------------- generators

def test_splitmix64_known_answers():
    # splitmix64 from state 0 (Vigna's reference implementation)
    s = C.c_uint64(0)
    out = [L.dor_splitmix64(C.byref(s)) for _ in range(3)]
    assert out == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_xoshiro_step_against_python():
    s = (C.c_uint64 * 4)(1, 2, 3, 4)
    st = [1, 2, 3, 4]
    M = (1 << 64) - 1
    for _ in range(100):
        res = (st[0] + st[3]) & M
        t = (st[1] << 17) & M
        st[2] ^= st[0]; st[3] ^= st[1]; st[1] ^= st[2]; st[0] ^= st[3]; st[2] ^= t
        st[3] = ((st[3] << 45) | (st[3] >> 19)) & M
        assert L.dor_xoshiro256p_next(s) == res
    assert list(s) == st


def test_seeding_keeps_32_bits_per_word():
    # rng.hpp:46-59 keeps the low and high halves of each splitmix64 output
    s = ob.seed_state(1)
    sm = C.c_uint64(1407513600 + 1)
    a = L.dor_splitmix64(C.byref(sm))
    b = L.dor_splitmix64(C.byref(sm))
    assert list(s) == [a & 0xFFFFFFFF, a >> 32, b & 0xFFFFFFFF, b >> 32]


def test_jump_commutes_with_step():
    # the jump is a power of the transition, so it commutes with one step
    a = (C.c_uint64 * 4)(*ob.seed_state(3))
    b = (C.c_uint64 * 4)(*ob.seed_state(3))
    L.dor_xoshiro256p_jump(a)
    L.dor_xoshiro256p_next(a)
    L.dor_xoshiro256p_next(b)
    L.dor_xoshiro256p_jump(b)
    assert list(a) == list(b)


def _step_matrix():
    """The xoshiro256+ transition as a 256x256 matrix over GF(2): state bit
    64w+b is bit b of word w; column j is the step applied to unit vector j
    (the step is linear over GF(2): xors, a shift and a rotation)."""
    M = np.zeros((256, 256), dtype=np.uint8)
    for j in range(256):
        s = (C.c_uint64 * 4)(0, 0, 0, 0)
        s[j // 64] = 1 << (j % 64)
        L.dor_xoshiro256p_next(s)
        for w in range(4):
            M[64 * w:64 * w + 64, j] = [(s[w] >> b) & 1 for b in range(64)]
    return M


def _gf2_mul(A, B):
    # entries of the integer product are at most 256: exact in float64
    return (A.astype(np.float64) @ B.astype(np.float64)).astype(np.int64).astype(np.uint8) & 1


def _apply(M, state):
    v = np.array([(state[w] >> b) & 1 for w in range(4) for b in range(64)], dtype=np.float64)
    o = (M.astype(np.float64) @ v).astype(np.int64) & 1
    return [int(sum(int(o[64 * w + b]) << b for b in range(64))) for w in range(4)]


def test_step_matrix_matches_generator():
    # linearity: the matrix reproduces the generator on arbitrary states
    M = _step_matrix()
    for seed in (1, 7, 256):
        s = (C.c_uint64 * 4)(*ob.seed_state(seed))
        want = _apply(M, list(s))
        L.dor_xoshiro256p_next(s)
        assert list(s) == want


def test_jump_is_two_to_the_128_steps():
    # xoshiro256p_jump (rng.hpp:64-87) advances the state by 2^128 steps:
    # T^(2^128) from 128 squarings of the one-step matrix.  A wrong jump
    # polynomial, an xor taken after the step or a dropped final copy all
    # give a different state.
    J = _step_matrix()
    for _ in range(128):
        J = _gf2_mul(J, J)
    for seed in (1, 3, 99):
        s = (C.c_uint64 * 4)(*ob.seed_state(seed))
        want = _apply(J, list(s))
        L.dor_xoshiro256p_jump(s)
        assert list(s) == want
    # and not a smaller power: 2^127 steps give another state
    H = _step_matrix()
    for _ in range(127):
        H = _gf2_mul(H, H)
    s = (C.c_uint64 * 4)(*ob.seed_state(1))
    half = _apply(H, list(s))
    L.dor_xoshiro256p_jump(s)
    assert list(s) != half


def test_mutant_indices_distinct_and_rejection_order():
    s = ob.seed_state(9)
    ref = [int(v) for v in s]
    mi = (C.c_size_t * 3)()
    for P in (4, 5, 8, 32):
        for i0 in range(P):
            L.dor_set_mutant_indices(s, mi, i0, P)
            assert len({mi[0], mi[1], mi[2], i0}) == 4
            assert all(0 <= m < P for m in mi)
    # replay the first call by hand: rejection on x % P, in order
    st = (C.c_uint64 * 4)(*ref)
    got = []
    while len(got) < 3:
        x = L.dor_xoshiro256p_next(st) % 4
        if x != 0 and x not in got:
            got.append(x)
    s2 = (C.c_uint64 * 4)(*ref)
    L.dor_set_mutant_indices(s2, mi, 0, 4)
    assert list(mi) == got


# ---------------------------------------------------------------- reductions

@pytest.mark.parametrize("n", range(1, 14))
def test_fold4_dot_exact_rounding(n):
    rng = np.random.default_rng(n)
    a, b = rng.random(n) * 3, rng.random(n)
    acc0 = 0.25
    got = L.dor_dot_fold4(acc0, ob.ptr(arr(a)), ob.ptr(arr(b)), n)
    # 4-blocks and a 2-block of rounded products, a fused last odd element
    acc = acc0
    i = 0
    if n >= 4:
        while i + 4 <= n:
            for l in range(4):
                acc = float(np.float64(acc) + np.float64(a[i + l]) * np.float64(b[i + l]))
            i += 4
    if n - i >= 2:
        acc = float(np.float64(acc) + np.float64(a[i]) * np.float64(b[i]))
        acc = float(np.float64(acc) + np.float64(a[i + 1]) * np.float64(b[i + 1]))
        i += 2
    if n - i == 1:
        acc = fma_exact(a[i], b[i], acc)
    assert got == acc


@pytest.mark.parametrize("n", [1, 5, 64, 257])
def test_fma_chain_exact(n):
    rng = np.random.default_rng(100 + n)
    a, b = rng.normal(size=n), rng.normal(size=n)
    acc = 0.0
    for x, y in zip(a, b):
        acc = fma_exact(x, y, acc)
    assert L.dor_fma_chain(0.0, ob.ptr(arr(a)), ob.ptr(arr(b)), n) == acc


@pytest.mark.parametrize("n", range(1, 12))
def test_chi_square_order(n):
    rng = np.random.default_rng(200 + n)
    o, e, s = rng.normal(size=n), rng.normal(size=n), rng.random(n) + 0.1
    got = L.dor_chi_square(ob.ptr(arr(o)), ob.ptr(arr(e)), ob.ptr(arr(s)), n)
    d = (o - e) / s
    acc = 0.0
    for t in range(n):
        if n % 2 == 1 and t == n - 1:
            acc = fma_exact(d[t], d[t], acc)
        else:
            acc = float(np.float64(acc) + d[t] * d[t])
    assert got == acc
    # and it is the chi^2 statistic
    assert got == pytest.approx(float(np.sum(d * d)), rel=1e-14)


@pytest.mark.parametrize("n", range(1, 11))
def test_squared_mean_tail(n):
    a = np.random.default_rng(300 + n).random(n) * 10
    got = L.dor_squared_mean(ob.ptr(arr(a)), n)
    acc = 0.0
    for i in range(n):
        if i >= n - n % 4:
            acc = fma_exact(a[i], a[i], acc)
        else:
            acc = float(np.float64(acc) + a[i] * a[i])
    assert got == acc / n
    assert L.dor_mean(ob.ptr(arr(a)), n) == pytest.approx(float(np.mean(a)), rel=1e-15)


# ---------------------------------------------------------------- kernel matrix

def _kernel(model, spectra, T=5, G=9, temperature=1.3, zero_t=False, tau=None):
    tau = np.linspace(0, 0.4, T) if tau is None else np.asarray(tau, dtype=np.float64)
    T = len(tau)
    f = ob.frequency_grid(G, 6.0)
    p = ob.make_problem(tau, -np.ones(T), np.ones(T), f, temperature, model=model, spectra_type=spectra,
                        zero_temperature=zero_t)
    kp, km = np.zeros(T * G), np.zeros(T * G)
    L.dor_kernel_terms(C.byref(p), ob.ptr(kp), ob.ptr(km))
    return tau, f, kp.reshape(T, G), km.reshape(T, G)


def _weights(f):
    df = np.empty_like(f)
    df[0] = 0.5 * (f[1] - f[0])
    df[-1] = 0.5 * (f[-1] - f[-2])
    df[1:-1] = 0.5 * (f[2:] - f[:-2])
    return df


@pytest.mark.parametrize("spectra", ["spbsf", "spfsf", "bfull", "ffull"])
def test_kernel_closed_forms(spectra):
    beta = 1 / 1.3
    per = -1.0 if spectra in ("spfsf", "ffull") else 1.0
    tau, f, kp, km = _kernel("standard", spectra)
    df = _weights(f)
    np.testing.assert_allclose(kp, df * np.exp(-np.outer(tau, f)), rtol=1e-15)
    np.testing.assert_allclose(km, per * df * np.exp(np.outer(tau - beta, f)), rtol=1e-15)
    tau, f, kp, km = _kernel("normalization", spectra)
    np.testing.assert_allclose(kp, df * np.exp(-np.outer(tau, f)) / (1 + np.exp(-beta * f)), rtol=1e-14)
    e = np.exp(np.outer(tau - beta, f))
    np.testing.assert_allclose(km, per * df * (e - e / (1 + np.exp(-beta * f))), rtol=1e-14, atol=1e-300)
    tau, f, kp, km = _kernel("hyperbolic", spectra)
    np.testing.assert_allclose(kp, per * df * np.exp(np.outer(-0.5 * beta - tau, f)) / 2, rtol=1e-14)
    np.testing.assert_allclose(km, per * df * np.exp(-np.outer(1.5 * beta - tau, f)) / 2, rtol=1e-14)


def test_models_agree_on_physical_isf():
    # The models parametrise one spectrum differently.  Standard genome:
    # S(+w), S(-w) = e^{-beta w} S(w) (detailed balance).  The normalization
    # model's genome carries factors (1+e^{-+beta w}) and reproduces the same
    # F(tau).  The hyperbolic model's genome is 2 DSF e^{-+beta w/2}; its
    # positive-frequency kernel (deac.cpp:323, 363) carries tau + beta, so its
    # positive half equals the standard positive half at tau + beta and its
    # negative half equals the standard negative half at tau.
    beta = 1 / 1.3
    tau = np.linspace(0, 0.4, 6)
    _, f, kps, kms = _kernel("standard", "bfull", G=40, tau=tau)
    _, _, kpn, kmn = _kernel("normalization", "bfull", G=40, tau=tau)
    _, _, kph, kmh = _kernel("hyperbolic", "bfull", G=40, tau=tau)
    _, _, kps_b, _ = _kernel("standard", "bfull", G=40, tau=tau + beta)
    S = np.exp(-(f - 2.0) ** 2)
    Sp, Sm = S, np.exp(-beta * f) * S
    E_std = kps @ Sp + kms @ Sm
    E_norm = kpn @ (Sp * (1 + np.exp(-beta * f))) + kmn @ (Sm * (1 + np.exp(beta * f)))
    np.testing.assert_allclose(E_norm, E_std, rtol=1e-12)
    np.testing.assert_allclose(kph @ (2 * Sp * np.exp(-0.5 * beta * f)), kps_b @ Sp, rtol=1e-12)
    np.testing.assert_allclose(kmh @ (2 * Sm * np.exp(0.5 * beta * f)), kms @ Sm, rtol=1e-12)


def test_zero_temperature_kernel():
    tau, f, kp, _ = _kernel("standard", "bfull", zero_t=True)
    np.testing.assert_allclose(kp, _weights(f) * np.exp(-np.outer(tau, f)), rtol=1e-15)


# ---------------------------------------------------------------- invariants

def _run(N=60, P=8, G=16, T=6, seed=3, **kw):
    tau = np.linspace(0, 0.4, T)
    isf = np.exp(-2 * tau)
    f = ob.frequency_grid(G, 6.0)
    p = ob.make_problem(tau, isf, np.full(T, 0.02), f, 1.3, number_of_generations=N, population_size=P,
                        stop_minimum_fitness=0.0, **kw)
    return ob.run(p, seed, want_population=True), p, tau, isf, f


def _uniforms(seed, n):
    s = ob.seed_state(seed)
    return np.array([(L.dor_xoshiro256p_next(s) >> 11) * 2.0 ** -53 for _ in range(n)])


def test_best_fitness_is_monotone_and_min_of_population():
    out, *_ = _run(N=400)
    m = out["minimum"]
    assert np.all(np.diff(m) <= 0)
    assert m[-1] == out["fitness"].min() == out["minimum_fitness"]
    assert out["fitness"][out["best_index"]] == out["minimum_fitness"]


def test_no_crossover_freezes_population():
    frozen = dict(crossover_probability=0.0, self_adapting_crossover_probability=0.0)
    out0, *_ = _run(N=2, **frozen)
    out, *_ = _run(N=50, **frozen)
    assert np.all(out["minimum"] == out["minimum"][0])
    assert np.all(out["mean"] == out["mean"][0])
    np.testing.assert_array_equal(out["population"], out0["population"])


def test_normalize_sets_zeroth_moment():
    out, p, tau, isf, f = _run(N=2, normalize=True, crossover_probability=0.0,
                               self_adapting_crossover_probability=0.0)
    df = _weights(f)
    pop = out["population"]
    norms = pop[:, 0, :] @ df + pop[:, 1, :] @ df
    np.testing.assert_allclose(norms, isf[0], rtol=1e-13)


def test_fitness_is_reduced_chi_square_of_best_member():
    out, p, tau, isf, f = _run(N=80)
    beta = 1 / 1.3
    df = _weights(f)
    best = out["population"][out["best_index"]]
    E = (df * np.exp(-np.outer(tau, f))) @ best[0] + (df * np.exp(np.outer(tau - beta, f))) @ best[1]
    chi = np.sum(((isf - E) / 0.02) ** 2) / len(tau)
    assert out["minimum_fitness"] == pytest.approx(chi, rel=1e-12)
    # and the DSF output of bfull/standard is the best genome, mirrored
    G = len(f)
    np.testing.assert_array_equal(out["dsf"][G - 1:], best[0])
    np.testing.assert_array_equal(out["dsf"][:G - 1], best[1][1:][::-1])
    assert math.copysign(1.0, out["frequency"][G - 1]) == -1.0  # the centre is -0.0


def test_zero_frequency_match():
    out, *_ = _run(N=100)
    pop = out["population"]
    np.testing.assert_array_equal(pop[:, 0, 0], pop[:, 1, 0])


def test_trials_stay_non_negative():
    out, *_ = _run(N=200)
    assert np.all(out["population"] >= 0)
    out, *_ = _run(N=200, allow_negative=True)
    assert np.any(out["population"] < 0)


# ---------------------------------------------------------------- zero temperature
# The reference's ZeroT build does not compile (deac.cpp:345, 397); its
# intent is a single-sign kernel K = df e^{-tau w} over the positive
# frequencies with a G-length output (deac.cpp:392-399, 531-535, 1669-1673).
# These checks pin the oracle's zero-temperature branch of dor_run.

def _zero_t_kernel(tau, f):
    return _weights(f) * np.exp(-np.outer(tau, f))


def test_zero_temperature_initial_population_is_one_draw_per_gene():
    # CR = 0 and tau_CR = 0 freeze the population: it is the initial one, one
    # uniform per (member, gene) in member-major order -- not two per gene
    # as with two signs, and without the zero-frequency copy.
    P, G = 5, 7
    out, *_ = _run(N=4, P=P, G=G, zero_temperature=True, crossover_probability=0.0,
                   self_adapting_crossover_probability=0.0)
    pop = out["population"]
    assert pop.shape == (P, 1, G)
    np.testing.assert_array_equal(pop[:, 0, :].ravel(), _uniforms(3, P * G))


def test_zero_temperature_best_fitness_monotone():
    out, *_ = _run(N=300, zero_temperature=True)
    m = out["minimum"]
    assert np.all(np.diff(m) <= 0)
    assert m[-1] < m[0]
    assert m[-1] == out["fitness"].min() == out["minimum_fitness"]


def test_zero_temperature_fitness_is_chi_square_of_reported_dsf():
    out, p, tau, isf, f = _run(N=120, zero_temperature=True)
    G = len(f)
    dsf, freq = out["dsf"], out["frequency"]
    assert len(dsf) == G and len(freq) == G
    np.testing.assert_array_equal(freq, f)                    # no mirror, no -0.0 centre
    best = out["population"][out["best_index"], 0]
    np.testing.assert_array_equal(dsf, best)                  # the DSF is the genome
    E = _zero_t_kernel(tau, f) @ dsf
    chi = np.sum(((isf - E) / 0.02) ** 2) / len(tau)
    assert out["minimum_fitness"] == pytest.approx(chi, rel=1e-12)
    # a two-sign evaluation of the same genome (the standard negative kernel
    # added) gives a different fitness: the branch really is single-sign
    beta = 1 / 1.3
    E2 = E + (_weights(f) * np.exp(np.outer(tau - beta, f))) @ dsf
    assert abs(np.sum(((isf - E2) / 0.02) ** 2) / len(tau) - chi) > 1e-3 * chi


def test_zero_temperature_normalize_uses_single_sign_weights():
    out, p, tau, isf, f = _run(N=2, zero_temperature=True, normalize=True, crossover_probability=0.0,
                               self_adapting_crossover_probability=0.0)
    pop = out["population"][:, 0, :]
    np.testing.assert_allclose(pop @ _weights(f), isf[0], rtol=1e-13)
    # the normalization is a rescaling of the one-draw-per-gene population
    raw = _uniforms(3, pop.size).reshape(pop.shape)
    ratio = pop / raw
    np.testing.assert_allclose(ratio, ratio[:, :1] * np.ones_like(ratio), rtol=1e-14)


def test_zero_temperature_moment_terms():
    tau = np.linspace(0, 0.4, 3)
    f = ob.frequency_grid(9, 6.0)
    df = _weights(f)
    p = ob.make_problem(tau, np.ones(3), np.ones(3), f, 1.3, zero_temperature=True)
    want = {0: df, 1: df * f, 3: df * f ** 4}
    for which, w in want.items():
        tp = np.zeros(9)
        L.dor_moment_terms(C.byref(p), which, ob.ptr(tp), None)
        np.testing.assert_allclose(tp, w, rtol=1e-15)


def test_zero_temperature_moment_constraints_enter_fitness():
    # first and third moments of the best genome under the single-sign terms
    # reproduce the reported fitness (deac.cpp:1550-1555, CPU formula)
    m1, m3, e3 = 0.8, 2.5, 0.4
    out, p, tau, isf, f = _run(N=80, zero_temperature=True, first_moment=m1, third_moment=m3,
                               third_moment_error=e3)
    best = out["population"][out["best_index"], 0]
    df = _weights(f)
    E = _zero_t_kernel(tau, f) @ best
    fit = np.sum(((isf - E) / 0.02) ** 2) / len(tau)
    fit += (best @ (df * f) - m1) ** 2 / m1
    fit += ((m3 - best @ (df * f ** 4)) / e3) ** 2
    assert out["minimum_fitness"] == pytest.approx(fit, rel=1e-12)


@pytest.mark.parametrize("tau_c", [0.0, 1.0])
def test_zero_temperature_draws_per_generation(tau_c):
    # One generation consumes the jDE draws of ONE sign for every member (a
    # check for CR and one for F, each followed by a redraw when it passes:
    # deac.cpp:1180-1213), then per member the donor draws by rejection
    # (deac.cpp:200-216) and G crossover draws (deac.cpp:1231-1240); the
    # initial population took P*G.  The returned state must sit exactly that
    # far along the stream: two-sign jDE draws or 2G crossover draws land
    # elsewhere.
    P, G, seed = 6, 11, 5
    out, *_ = _run(N=2, P=P, G=G, seed=seed, zero_temperature=True, self_adapting_crossover_probability=tau_c,
                   self_adapting_differential_weight_probability=tau_c)
    s = ob.seed_state(seed)
    nxt = lambda: L.dor_xoshiro256p_next(s)
    for _ in range(P * G):
        nxt()
    for _ in range(P * (4 if tau_c == 1.0 else 2)):
        nxt()
    for i in range(P):
        got = []
        while len(got) < 3:
            x = nxt() % P
            if x != i and x not in got:
                got.append(x)
        for _ in range(G):
            nxt()
    assert out["rng"] == [int(v) for v in s]
