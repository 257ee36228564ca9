"""ctypes front end of libdeac_amd.so (the C ABI in include/deac_engine.h).

Mirrors the reference's deac() call (src/deac/src/deac.cpp:218-229): one call
runs a whole reconstruction and returns the best-member DSF, the mirrored
frequency grid, the generation count, the best fitness and the per-generation
statistics that --track_stats writes.  The library is required: importing
this module without a built libdeac_amd.so raises, there is no fallback.
"""
import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdeac_amd.so")

MODELS = {"standard": 0, "hyperbolic": 1, "normalization": 2}
SPECTRA = {"spbsf": 0, "spfsf": 1, "bfull": 2, "ffull": 3, "bdsf": 4}
STATUS = {0: "success", 1: "invalid argument", 2: "no usable gfx950 device", 3: "out of memory",
          4: "device error", 5: "unsupported"}

EXPORTED = ["deac_create", "deac_run", "deac_destroy", "deac_output_length", "deac_seed_state",
            "deac_rng_advance", "deac_frequency_grid", "deac_status_string", "deac_last_error_message", "deac_last_run_profile",
            "deac_batch_create", "deac_batch_run", "deac_batch_continue", "deac_batch_run_noalloc", "deac_batch_tune", "deac_batch_set_schedule",
            "deac_batch_get_schedule", "deac_batch_role_cycles", "deac_batch_destroy", "deac_batch_size", "deac_batch_last_profile", "deac_genome_length", "deac_model_isf"]

_dp = C.POINTER(C.c_double)


class Problem(C.Structure):
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
        ("population_size", C.c_size_t),
        ("normalize", C.c_int),
        ("first_moment", C.c_double),
        ("third_moment", C.c_double),
        ("third_moment_error", C.c_double),
        ("crossover_probability", C.c_double),
        ("self_adapting_crossover_probability", C.c_double),
        ("differential_weight", C.c_double),
        ("self_adapting_differential_weight_probability", C.c_double),
        ("use_negative_first_moment", C.c_int),
    ]


class RunOptions(C.Structure):
    _fields_ = [("number_of_generations", C.c_size_t), ("stop_minimum_fitness", C.c_double),
                ("track_stats", C.c_int), ("ks_precision", C.c_int), ("split_loop", C.c_int)]


class Result(C.Structure):
    _fields_ = [
        ("generation", C.c_size_t),
        ("minimum_fitness", C.c_double),
        ("best_index", C.c_size_t),
        ("dsf", _dp),
        ("frequency", _dp),
        ("stats_mean", _dp),
        ("stats_minimum", _dp),
        ("stats_squared_mean", _dp),
        ("final_fitness", _dp),
    ]


class RunProfile(C.Structure):
    _fields_ = [("loop_ms", C.c_double), ("walk_ms", C.c_double), ("mask_ms", C.c_double),
                ("evolve_ms", C.c_double), ("generations", C.c_size_t), ("evolve_launches", C.c_size_t),
                ("eval_ms", C.c_double), ("eval_launches", C.c_size_t), ("graph_launches", C.c_size_t),
                ("evolve_workgroups", C.c_size_t)]


class Schedule(C.Structure):
    _fields_ = [("group_streams", C.c_int), ("group_order", C.c_int), ("instance_order", C.c_int),
                ("ms_per_generation", C.c_double), ("pack", C.c_int)]


class DeacError(RuntimeError):
    def __init__(self, status, message=None):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}" + (f": {message}" if message else ""))


_lib = None


def lib():
    """Load libdeac_amd.so; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libdeac_amd.so not built at {LIB_PATH}; run make -C deac-cpp_amd")
        L = C.CDLL(LIB_PATH)
        L.deac_create.argtypes = [C.POINTER(Problem), C.POINTER(C.c_void_p), C.POINTER(C.c_char_p)]
        L.deac_create.restype = C.c_int
        L.deac_run.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(RunOptions), C.POINTER(Result)]
        L.deac_run.restype = C.c_int
        L.deac_destroy.argtypes = [C.c_void_p]
        L.deac_output_length.argtypes = [C.POINTER(Problem)]
        L.deac_output_length.restype = C.c_size_t
        L.deac_seed_state.argtypes = [C.c_uint64, C.POINTER(C.c_uint64)]
        L.deac_rng_advance.argtypes = [C.POINTER(C.c_uint64), C.c_uint64]
        L.deac_frequency_grid.argtypes = [C.c_size_t, C.c_double, _dp]
        L.deac_frequency_grid.restype = C.c_int
        L.deac_status_string.argtypes = [C.c_int]
        L.deac_status_string.restype = C.c_char_p
        L.deac_last_error_message.argtypes = []
        L.deac_last_error_message.restype = C.c_char_p
        L.deac_last_run_profile.argtypes = [C.c_void_p, C.POINTER(RunProfile)]
        L.deac_last_run_profile.restype = C.c_int
        L.deac_batch_create.argtypes = [C.POINTER(Problem), C.c_size_t, C.POINTER(C.c_void_p), C.POINTER(C.c_char_p)]
        L.deac_batch_create.restype = C.c_int
        L.deac_batch_run.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(RunOptions), C.POINTER(Result)]
        L.deac_batch_run.restype = C.c_int
        L.deac_batch_continue.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(RunOptions), C.POINTER(Result)]
        L.deac_batch_continue.restype = C.c_int
        L.deac_batch_run_noalloc.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(RunOptions), C.POINTER(Result)]
        L.deac_batch_run_noalloc.restype = C.c_int
        L.deac_batch_tune.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(RunOptions), C.c_size_t,
                                      C.POINTER(Schedule)]
        L.deac_batch_tune.restype = C.c_int
        L.deac_batch_role_cycles.argtypes = [C.c_int, C.POINTER(C.c_ulonglong), C.c_int]
        L.deac_batch_role_cycles.restype = C.c_int
        L.deac_batch_set_schedule.argtypes = [C.c_void_p, C.POINTER(Schedule)]
        L.deac_batch_set_schedule.restype = C.c_int
        L.deac_batch_get_schedule.argtypes = [C.c_void_p, C.POINTER(Schedule)]
        L.deac_batch_get_schedule.restype = C.c_int
        L.deac_batch_destroy.argtypes = [C.c_void_p]
        L.deac_batch_size.argtypes = [C.c_void_p]
        L.deac_batch_size.restype = C.c_size_t
        L.deac_batch_last_profile.argtypes = [C.c_void_p, C.POINTER(RunProfile)]
        L.deac_batch_last_profile.restype = C.c_int
        L.deac_genome_length.argtypes = [C.POINTER(Problem)]
        L.deac_genome_length.restype = C.c_size_t
        L.deac_model_isf.argtypes = [C.POINTER(Problem), _dp, C.c_size_t, C.c_int, _dp, _dp,
                                     C.POINTER(C.c_double), C.POINTER(C.c_char_p)]
        L.deac_model_isf.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _last_message():
    m = lib().deac_last_error_message()
    return m.decode() if m else None


# Engines and batches still open when the interpreter exits are closed by an
# atexit hook, while the HIP runtime is certainly alive: a handle destroyed
# later, from a finalizer during interpreter teardown, would reach
# hipStreamSynchronize/hipFree after the runtime (or a profiler attached to
# it) has shut down.  Finalizers that run while the interpreter is
# finalizing leave the handle to the process's exit instead.
_open = weakref.WeakSet()


@atexit.register
def _close_open_handles():
    for obj in list(_open):
        try:
            obj.close()
        except Exception:
            pass


def _finalize(obj):
    if sys.is_finalizing():
        return
    try:
        obj.close()
    except Exception:
        pass


def seed_state(seed):
    s = (C.c_uint64 * 4)()
    lib().deac_seed_state(C.c_uint64(int(seed)), s)
    return s


def rng_advance(state, n):
    s = (C.c_uint64 * 4)(*[int(v) for v in state])
    lib().deac_rng_advance(s, C.c_uint64(int(n)))
    return [int(v) for v in s]


def frequency_grid(genome_size, omega_max):
    f = np.empty(genome_size)
    st = lib().deac_frequency_grid(genome_size, omega_max, _ptr(f))
    if st:
        raise DeacError(st)
    return f


def make_problem(tau, isf, err, frequency, temperature, *, model="standard", spectra_type="bfull",
                 zero_temperature=False, allow_negative=False, renormalize=False, population_size=8,
                 normalize=False, first_moment=-1.0, third_moment=-1.0, third_moment_error=0.0,
                 crossover_probability=0.9, self_adapting_crossover_probability=0.1,
                 differential_weight=0.9, self_adapting_differential_weight_probability=0.1,
                 use_negative_first_moment=False):
    """spectra_type "bdsf" is the reference's bosonic detailed-balance build (one sign of genes, G-length
    output); use_negative_first_moment adds its -1st-moment term above zero temperature."""
    keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (tau, isf, err, frequency)]
    p = Problem(len(keep[0]), _ptr(keep[0]), _ptr(keep[1]), _ptr(keep[2]), len(keep[3]), _ptr(keep[3]),
                temperature, MODELS[model], SPECTRA[spectra_type], int(zero_temperature), int(allow_negative),
                int(renormalize), population_size, int(normalize), first_moment, third_moment,
                third_moment_error, crossover_probability, self_adapting_crossover_probability,
                differential_weight, self_adapting_differential_weight_probability, int(use_negative_first_moment))
    p._keep = keep
    return p


class Engine:
    """One deac_handle: problem uploaded once, any number of runs."""

    def __init__(self, problem):
        self.problem = problem
        h = C.c_void_p()
        msg = C.c_char_p()
        self.handle = None
        st = lib().deac_create(C.byref(problem), C.byref(h), C.byref(msg))
        if st:
            raise DeacError(st, msg.value.decode() if msg.value else None)
        self.handle = h
        _open.add(self)

    def close(self):
        if self.handle:
            lib().deac_destroy(self.handle)
            self.handle = None
        _open.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        _finalize(self)

    def run(self, seed, number_of_generations, stop_minimum_fitness=1.0, track_stats=True, precision="fp64",
            split_loop=False):
        """Run from --seed (int) or an explicit 4-word RNG state.

        precision: K.S of the model ISF, "fp64" (exact), "fp32" or "bf16x3"
        (mixed, through the split generation loop); split_loop=True runs
        fp64 through the split loop as well."""
        p = self.problem
        N = number_of_generations
        nout = lib().deac_output_length(C.byref(p))
        out = {k: np.zeros(n) for k, n in (("dsf", nout), ("frequency", nout), ("mean", N), ("minimum", N),
                                            ("squared_mean", N), ("fitness", p.population_size))}
        r = Result(0, 0.0, 0, _ptr(out["dsf"]), _ptr(out["frequency"]), _ptr(out["mean"]), _ptr(out["minimum"]),
                   _ptr(out["squared_mean"]), _ptr(out["fitness"]))
        if isinstance(seed, (int, np.integer)):
            s = seed_state(seed)
        else:
            s = (C.c_uint64 * 4)(*[int(v) for v in seed])
        o = RunOptions(N, stop_minimum_fitness, int(track_stats), PRECISIONS[precision], int(split_loop))
        st = lib().deac_run(self.handle, s, C.byref(o), C.byref(r))
        if st:
            raise DeacError(st, _last_message())
        n = r.generation + 1
        for k in ("mean", "minimum", "squared_mean"):
            out[k] = out[k][:n]
        prof = RunProfile()
        lib().deac_last_run_profile(self.handle, C.byref(prof))
        out.update(generation=r.generation, minimum_fitness=r.minimum_fitness, best_index=r.best_index,
                   rng=[int(v) for v in s], loop_ms=prof.loop_ms, loop_generations=prof.generations,
                   profile={k: getattr(prof, k) for k, _ in RunProfile._fields_})
        return out


def _per_instance(v, n):
    return list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v] * n


class Batch:
    """One deac_batch: independent reconstructions run by one persistent launch.

    Instance i computes what Engine(problems[i]).run(seeds[i], ...) computes:
    the multi-seed workflow of the reference (one deac.e per seed) in one call.
    """

    def __init__(self, problems):
        self.problems = list(problems)
        n = len(self.problems)
        arr = (Problem * n)(*self.problems)
        self._arr = arr
        h = C.c_void_p()
        msg = C.c_char_p()
        self.handle = None
        st = lib().deac_batch_create(arr, n, C.byref(h), C.byref(msg))
        if st:
            raise DeacError(st, msg.value.decode() if msg.value else None)
        self.handle = h
        _open.add(self)

    def close(self):
        if self.handle:
            lib().deac_batch_destroy(self.handle)
            self.handle = None
        _open.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        _finalize(self)

    def __len__(self):
        return len(self.problems)

    def run(self, seeds, number_of_generations, stop_minimum_fitness=1.0, track_stats=True, precision="fp64"):
        """seeds: list of --seed ints or 4-word states; N, stop, track_stats, precision ("fp64" exact, "fp32"
        or "bf16x3": mixed K.S in the delta form, FP64 everything else): scalar or per instance."""
        n = len(self.problems)
        if len(seeds) != n:
            raise ValueError(f"one seed or state per problem: {len(seeds)} given for {n} problems")
        return self._run(seeds, number_of_generations, stop_minimum_fitness, track_stats, precision, False)

    def tune(self, seeds, generations=20, precision="fp64"):
        """deac_batch_tune: time the candidate launch schedules on `generations` generations and keep the
        fastest; returns it as a dict."""
        n = len(self.problems)
        precs = _per_instance(precision, n)
        states = (C.c_uint64 * (4 * n))()
        for i, sd in enumerate(seeds):
            w = seed_state(sd) if isinstance(sd, (int, np.integer)) else sd
            for k in range(4):
                states[4 * i + k] = int(w[k])
        opts = (RunOptions * n)(*[RunOptions(generations + 1, 0.0, 0, PRECISIONS[precs[i]], 0) for i in range(n)])
        sc = Schedule()
        st = lib().deac_batch_tune(self.handle, states, opts, generations, C.byref(sc))
        if st:
            raise DeacError(st, _last_message())
        return {k: getattr(sc, k) for k, _ in Schedule._fields_}

    @property
    def schedule(self):
        sc = Schedule()
        lib().deac_batch_get_schedule(self.handle, C.byref(sc))
        return {k: getattr(sc, k) for k, _ in Schedule._fields_}

    @schedule.setter
    def schedule(self, d):
        sc = Schedule(d["group_streams"], d["group_order"], d["instance_order"], d.get("ms_per_generation", 0.0),
                      d.get("pack", 0))
        st = lib().deac_batch_set_schedule(self.handle, C.byref(sc))
        if st:
            raise DeacError(st)

    def continue_run(self, number_of_generations, stop_minimum_fitness=1.0, track_stats=True, precision="fp64"):
        """deac_batch_continue: N more generations (N - 1 evolution steps) of every instance from where the
        last run left it."""
        return self._run(None, number_of_generations, stop_minimum_fitness, track_stats, precision, True)

    def _run(self, seeds, number_of_generations, stop_minimum_fitness, track_stats, precision, resume):
        n = len(self.problems)
        Ns = _per_instance(number_of_generations, n)
        stops = _per_instance(stop_minimum_fitness, n)
        tracks = _per_instance(track_stats, n)
        precs = _per_instance(precision, n)
        states = (C.c_uint64 * (4 * n))()
        for i, sd in enumerate(seeds or []):
            w = seed_state(sd) if isinstance(sd, (int, np.integer)) else sd
            for k in range(4):
                states[4 * i + k] = int(w[k])
        opts = (RunOptions * n)(*[RunOptions(Ns[i], stops[i], int(tracks[i]), PRECISIONS[precs[i]], 0)
                                  for i in range(n)])
        outs, res = [], (Result * n)()
        for i, p in enumerate(self.problems):
            nout = lib().deac_output_length(C.byref(p))
            o = {k: np.zeros(m) for k, m in (("dsf", nout), ("frequency", nout), ("mean", Ns[i]), ("minimum", Ns[i]),
                                              ("squared_mean", Ns[i]), ("fitness", p.population_size))}
            res[i] = Result(0, 0.0, 0, _ptr(o["dsf"]), _ptr(o["frequency"]), _ptr(o["mean"]), _ptr(o["minimum"]),
                            _ptr(o["squared_mean"]), _ptr(o["fitness"]))
            outs.append(o)
        st = (lib().deac_batch_continue if resume else lib().deac_batch_run)(self.handle, states, opts, res)
        if st:
            raise DeacError(st, _last_message())
        prof = RunProfile()
        lib().deac_batch_last_profile(self.handle, C.byref(prof))
        self.profile = {k: getattr(prof, k) for k, _ in RunProfile._fields_}
        for i, o in enumerate(outs):
            r = res[i]
            k = r.generation + 1
            for key in ("mean", "minimum", "squared_mean"):
                o[key] = o[key][:k]
            o.update(generation=r.generation, minimum_fitness=r.minimum_fitness, best_index=r.best_index,
                     rng=[int(states[4 * i + w]) for w in range(4)])
        return outs


ROLES = ["chain0", "chain1", "chain2", "chain3"] + [f"worker{i}" for i in range(8)] + ["walk", "mask0", "mask1",
                                                                                    "mask2"]


def role_cycles(precision="fp64", reset=False):
    """deac_batch_role_cycles (profiling builds): per wavefront role, cycles
    per generation spent working (launch cycles minus barrier waits), and the
    launch cycles per generation."""
    buf = (C.c_ulonglong * 48)()
    st = lib().deac_batch_role_cycles(PRECISIONS[precision], buf, int(reset))
    if st:
        raise DeacError(st, _last_message())
    v = np.array(list(buf), dtype=np.float64).reshape(16, 3)
    gens = max(v[0, 2], 1.0)
    return {"generations": int(v[0, 2]), "cycles_per_generation": v[0, 1] / gens,
            "work_cycles_per_generation": {n: (v[i, 1] - v[i, 0]) / gens for i, n in enumerate(ROLES)}}


def run(problem, seed, number_of_generations, stop_minimum_fitness=1.0, track_stats=True):
    with Engine(problem) as e:
        return e.run(seed, number_of_generations, stop_minimum_fitness, track_stats)


PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2}


def model_isf(problem, genomes, precision="fp64"):
    """deac_model_isf: model ISF (count x T) and reduced chi^2 (count) of a
    stack of genomes at a K.S precision, plus the contraction's device time
    in ms."""
    L = lib().deac_genome_length(C.byref(problem))
    g = np.ascontiguousarray(genomes, dtype=np.float64).reshape(-1, L)
    n, T = g.shape[0], problem.number_of_timeslices
    e = np.zeros((n, T))
    chi = np.zeros(n)
    ms = C.c_double(0.0)
    msg = C.c_char_p()
    st = lib().deac_model_isf(C.byref(problem), _ptr(g), n, PRECISIONS[precision], _ptr(e), _ptr(chi),
                              C.byref(ms), C.byref(msg))
    if st:
        raise DeacError(st, msg.value.decode() if msg.value else None)
    return e, chi, ms.value
