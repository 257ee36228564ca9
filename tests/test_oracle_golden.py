"""The oracle against the reference executables' golden vectors.

Every case under tests/golden/ was produced by the reference CPU build
(compiled from its sources by oracle/Makefile); the oracle must reproduce its
per-generation statistics, DSF and frequency grid bit for bit.
"""
import pytest

import golden_cases as gc
import oracle_binding as ob

# the long histories (config 1 in full, config 2 in full, two 2000-generation
# config-4 seeds) are checked on a prefix here (the prefix of a longer run is
# itself a complete run) and in full by the GPU tests
CPU_PREFIX = {"c1_full": 20000, "c2_full": 400, "c4_seed1_prefix2000": 120, "c4_seed200_prefix2000": 120}


@pytest.mark.parametrize("name", gc.names())
def test_oracle_matches_reference(name):
    case = gc.load(name)
    a = case["meta"]["args"]
    n = CPU_PREFIX.get(name)
    N = n if n else a["N"]
    freq = gc.frequency_of(case, ob.frequency_grid)
    prob = ob.make_problem(case["tau"], case["isf"], case["err"], freq, a["T"], number_of_generations=N,
                           stop_minimum_fitness=a["stop"], track_stats=True, **gc.problem_kwargs(case))
    got = ob.run(prob, a["seed"])
    errors = gc.compare_outputs(got, case, n_hist=n)
    assert not errors, f"{name}: " + "; ".join(errors)
