"""ctypes binding of the oracle library (oracle/libdeac_oracle.so).

The oracle is test infrastructure: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg use this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libdeac_oracle.so")

MODELS = {"standard": 0, "hyperbolic": 1, "normalization": 2}
SPECTRA = {"spbsf": 0, "spfsf": 1, "bfull": 2, "ffull": 3, "bdsf": 4}

_dp = C.POINTER(C.c_double)


class DorProblem(C.Structure):
    _fields_ = [
        ("number_of_timeslices", C.c_size_t),
        ("imaginary_time", _dp),
        ("isf", _dp),
        ("isf_error", _dp),
        ("genome_size", C.c_size_t),
        ("frequency", _dp),
        ("temperature", C.c_double),
        ("model", C.c_int),
        ("spectra_type", C.c_int),
        ("zero_temperature", C.c_int),
        ("allow_negative_spectral_weight", C.c_int),
        ("renormalize_each_generation", C.c_int),
        ("number_of_generations", C.c_size_t),
        ("population_size", C.c_size_t),
        ("normalize", C.c_int),
        ("first_moment", C.c_double),
        ("third_moment", C.c_double),
        ("third_moment_error", C.c_double),
        ("crossover_probability", C.c_double),
        ("self_adapting_crossover_probability", C.c_double),
        ("differential_weight", C.c_double),
        ("self_adapting_differential_weight_probability", C.c_double),
        ("stop_minimum_fitness", C.c_double),
        ("track_stats", C.c_int),
        ("use_negative_first_moment", C.c_int),
    ]


class DorResult(C.Structure):
    _fields_ = [
        ("generation", C.c_size_t),
        ("minimum_fitness", C.c_double),
        ("best_index", C.c_size_t),
        ("dsf", _dp),
        ("frequency_out", _dp),
        ("stats_mean", _dp),
        ("stats_minimum", _dp),
        ("stats_squared_mean", _dp),
        ("final_fitness", _dp),
        ("final_population", _dp),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        _lib = C.CDLL(ORACLE_SO)
        _lib.dor_run.argtypes = [C.POINTER(DorProblem), C.POINTER(C.c_uint64), C.POINTER(DorResult)]
        _lib.dor_xoshiro256p_init.argtypes = [C.c_uint64, C.POINTER(C.c_uint64)]
        _lib.dor_xoshiro256p_next.argtypes = [C.POINTER(C.c_uint64)]
        _lib.dor_xoshiro256p_next.restype = C.c_uint64
        _lib.dor_xoshiro256p_jump.argtypes = [C.POINTER(C.c_uint64)]
        _lib.dor_splitmix64.argtypes = [C.POINTER(C.c_uint64)]
        _lib.dor_splitmix64.restype = C.c_uint64
        _lib.dor_frequency_grid.argtypes = [C.c_size_t, C.c_double, _dp]
        _lib.dor_chi_square.argtypes = [_dp, _dp, _dp, C.c_size_t]
        _lib.dor_chi_square.restype = C.c_double
        _lib.dor_dot_fold4.argtypes = [C.c_double, _dp, _dp, C.c_size_t]
        _lib.dor_dot_fold4.restype = C.c_double
        _lib.dor_fma_chain.argtypes = [C.c_double, _dp, _dp, C.c_size_t]
        _lib.dor_fma_chain.restype = C.c_double
        _lib.dor_squared_mean.argtypes = [_dp, C.c_size_t]
        _lib.dor_squared_mean.restype = C.c_double
        _lib.dor_mean.argtypes = [_dp, C.c_size_t]
        _lib.dor_mean.restype = C.c_double
        _lib.dor_kernel_terms.argtypes = [C.POINTER(DorProblem), _dp, _dp]
        _lib.dor_moment_terms.argtypes = [C.POINTER(DorProblem), C.c_int, _dp, _dp]
        _lib.dor_set_mutant_indices.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_size_t), C.c_size_t, C.c_size_t]
    return _lib


def ptr(a):
    return a.ctypes.data_as(_dp)


def seed_state(seed):
    """RNG state of the reference for --seed (deac.cpp:2089-2091)."""
    s = (C.c_uint64 * 4)()
    lib().dor_xoshiro256p_init(C.c_uint64(1407513600 + int(seed)), s)
    return s


def frequency_grid(genome_size, omega_max):
    f = np.empty(genome_size)
    lib().dor_frequency_grid(genome_size, omega_max, ptr(f))
    return f


def make_problem(tau, isf, err, frequency, temperature, *, model="standard", spectra_type="bfull",
                 zero_temperature=False, allow_negative=False, renormalize=False,
                 number_of_generations=100, population_size=8, normalize=False,
                 first_moment=-1.0, third_moment=-1.0, third_moment_error=0.0,
                 crossover_probability=0.9, self_adapting_crossover_probability=0.1,
                 differential_weight=0.9, self_adapting_differential_weight_probability=0.1,
                 stop_minimum_fitness=1.0, track_stats=True, use_negative_first_moment=False):
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (tau, isf, err, frequency)]
    p = DorProblem(
        len(keep[0]), ptr(keep[0]), ptr(keep[1]), ptr(keep[2]), len(keep[3]), ptr(keep[3]),
        temperature, MODELS[model], SPECTRA[spectra_type], int(zero_temperature), int(allow_negative),
        int(renormalize), number_of_generations, population_size, int(normalize), first_moment,
        third_moment, third_moment_error, crossover_probability, self_adapting_crossover_probability,
        differential_weight, self_adapting_differential_weight_probability, stop_minimum_fitness,
        int(track_stats), int(use_negative_first_moment))
    p._keep = keep
    return p


def run(problem, seed, want_population=False):
    """Run the oracle; returns a dict of numpy outputs."""
    G, P, N = problem.genome_size, problem.population_size, problem.number_of_generations
    nsign = 1 if problem.zero_temperature or problem.spectra_type == SPECTRA["bdsf"] else 2
    nout = G if nsign == 1 else 2 * G - 1
    out = {k: np.zeros(n) for k, n in (("dsf", nout), ("frequency", nout), ("mean", N), ("minimum", N),
                                        ("squared_mean", N), ("fitness", P))}
    pop = np.zeros(P * nsign * G) if want_population else None
    r = DorResult(0, 0.0, 0, ptr(out["dsf"]), ptr(out["frequency"]), ptr(out["mean"]), ptr(out["minimum"]),
                  ptr(out["squared_mean"]), ptr(out["fitness"]), ptr(pop) if pop is not None else None)
    s = seed_state(seed) if not isinstance(seed, C.Array) else seed
    lib().dor_run(C.byref(problem), s, C.byref(r))
    n = r.generation + 1
    for k in ("mean", "minimum", "squared_mean"):
        out[k] = out[k][:n]
    out.update(generation=r.generation, minimum_fitness=r.minimum_fitness, best_index=r.best_index,
               rng=[int(v) for v in s])
    if pop is not None:
        out["population"] = pop.reshape(P, nsign, G)
    return out
