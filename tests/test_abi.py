"""The C ABI without a GPU: libdeac_amd.so loads, exports every entry point
include/deac_engine.h declares, and its host-side pieces (seeding, jump-ahead,
frequency grid, problem validation) agree with the oracle and the reference's
messages.  No call here reaches a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import deac_amd as da
import oracle_binding as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "deac_engine.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(deac_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_abi():
    assert set(declared_functions()) == set(da.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = da.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_seed_state_matches_reference_seeding():
    for seed in (0, 1, 7, 17, 123456789):
        assert list(da.seed_state(seed)) == list(ob.seed_state(seed))


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 1000, 8192, 16384 * 2 + 7])
def test_rng_advance_equals_stepping(n):
    s = ob.seed_state(42)
    start = [int(v) for v in s]
    for _ in range(n):
        ob.lib().dor_xoshiro256p_next(s)
    assert da.rng_advance(start, n) == [int(v) for v in s]


def test_rng_advance_composes():
    s = [int(v) for v in ob.seed_state(5)]
    a = da.rng_advance(da.rng_advance(s, 12345), 67890)
    assert a == da.rng_advance(s, 12345 + 67890)
    big = (1 << 63) + 12345
    assert da.rng_advance(da.rng_advance(s, big), 1) == da.rng_advance(s, big + 1)


@pytest.mark.parametrize("G,omega", [(2, 1.0), (8, 4.0), (512, 64.0), (4096, 512.0), (16384, 512.0), (1000, 3.7)])
def test_frequency_grid_matches_reference(G, omega):
    np.testing.assert_array_equal(da.frequency_grid(G, omega), ob.frequency_grid(G, omega))


def _tiny(**kw):
    tau = np.array([0, .2, .4, .6])
    isf = np.array([-1, -.85, -.72, -.61])
    err = np.full(4, .05)
    args = dict(tau=tau, isf=isf, err=err, frequency=np.linspace(0, 4, 8), temperature=1.0)
    args.update({k: kw.pop(k) for k in list(kw) if k in args})
    return da.make_problem(args["tau"], args["isf"], args["err"], args["frequency"], args["temperature"], **kw)


# (problem overrides, reference message) -- deac.cpp:2075-2151
VALIDATION = [
    (dict(tau=np.array([0.0]), isf=np.array([-1.0]), err=np.array([0.05])),
     "ISF input file must contain at least two timeslices"),
    (dict(isf=np.array([-1, np.nan, -.72, -.61])), "ISF input file contains non-finite values"),
    (dict(isf=np.array([1, .85, .72, .61]), spectra_type="spfsf"),
     "positive ISF values are not supported for single-particle spectra"),
    (dict(isf=np.array([1, .85, .72, .61]), spectra_type="spbsf"),
     "positive ISF values are not supported for single-particle spectra"),
    (dict(err=np.array([.05, 0, .05, .05])), "ISF input errors must be positive"),
    (dict(temperature=0.0), "temperature must be positive for non-ZeroT builds"),
    (dict(temperature=float("inf")), "temperature must be positive for non-ZeroT builds"),
    (dict(population_size=3), "population_size must be at least 4"),
    (dict(frequency=np.array([0.0])), "genome_size must be at least 2"),
    (dict(frequency=np.array([0.0, np.inf])), "frequencies must be finite and non-negative"),
    (dict(frequency=np.array([0.0, -1.0])), "frequencies must be finite and non-negative"),
    (dict(frequency=np.array([0.0, 2.0, 1.0])), "frequencies must be sorted in non-decreasing order"),
    (dict(first_moment=np.nan), "first_moment must be finite"),
    (dict(third_moment=1.0, third_moment_error=0.0), "third_moment_error must be positive when third_moment is used"),
]


@pytest.mark.parametrize("overrides,message", VALIDATION)
def test_create_rejects_invalid_problems(overrides, message):
    kw = dict(spectra_type="spfsf")
    kw.update(overrides)
    p = _tiny(**kw)
    with pytest.raises(da.DeacError) as e:
        da.Engine(p)
    assert e.value.status == 1
    assert message in str(e.value)


def test_zero_temperature_skips_temperature_check_and_halves_output():
    p = _tiny(temperature=0.0, zero_temperature=True, spectra_type="bfull", isf=np.array([1, .85, .72, .61]))
    assert da.lib().deac_output_length(C.byref(p)) == 8
    p2 = _tiny(spectra_type="bfull", isf=np.array([1, .85, .72, .61]))
    assert da.lib().deac_output_length(C.byref(p2)) == 15


def test_bdsf_has_one_sign():
    # the detailed-balance build: G genes and G output values (deac.cpp:1596-1606)
    p = _tiny(spectra_type="bdsf", isf=np.array([1, .85, .72, .61]), use_negative_first_moment=True)
    assert da.lib().deac_output_length(C.byref(p)) == 8
    assert da.lib().deac_genome_length(C.byref(p)) == 8


def test_status_strings():
    for code, text in da.STATUS.items():
        assert da.lib().deac_status_string(code).decode() == text


def _batch_problem(T=17, **kw):
    tau = np.linspace(0.0, 0.37, T)
    isf = np.exp(-tau)
    return da.make_problem(tau, isf, 1e-3 * isf, da.frequency_grid(64, 8.0), 1.0, **kw)


@pytest.mark.parametrize("kw,status,text", [
    (dict(population_size=70), 5, "at most 64 members"),
    (dict(population_size=40, T=140), 5, "at most 132 timeslices"),
    (dict(spectra_type="bdsf", T=130, population_size=20), 5, "detailed-balance (bdsf) runs take at most 128"),
    (dict(population_size=2), 1, "population_size"),
])
def test_batch_create_rejects_before_touching_a_device(kw, status, text):
    # deac_batch_create validates and plans every problem on the host first
    with pytest.raises(da.DeacError) as e:
        da.Batch([_batch_problem(), _batch_problem(**kw)])
    assert e.value.status == status
    assert "problem 1" in str(e.value) and text in str(e.value)


def test_batch_create_without_device_reports_no_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(da.DeacError) as e:
        da.Batch([_batch_problem()] * 3)
    assert e.value.status == 2


def test_model_isf_rejects_bad_arguments_before_the_device():
    import ctypes as C
    tau = np.linspace(0.0, 0.3, 4)
    p = da.make_problem(tau, np.exp(-tau), 1e-3 * np.exp(-tau), da.frequency_grid(8, 4.0), 1.0,
                        spectra_type="bfull")
    assert da.lib().deac_genome_length(C.byref(p)) == 16
    g = np.zeros((2, 16))
    msg = C.c_char_p()
    st = da.lib().deac_model_isf(C.byref(p), g.ctypes.data_as(C.POINTER(C.c_double)), 2, 7, None, None, None,
                                 C.byref(msg))
    assert st == 1 and b"precision" in msg.value
    st = da.lib().deac_model_isf(C.byref(p), g.ctypes.data_as(C.POINTER(C.c_double)), 0, 0, None, None, None,
                                 C.byref(msg))
    assert st == 1
    pz = da.make_problem(tau, np.exp(-tau), 1e-3 * np.exp(-tau), da.frequency_grid(8, 4.0), 0.0,
                         spectra_type="bfull", zero_temperature=True)
    assert da.lib().deac_genome_length(C.byref(pz)) == 8


def test_binding_structs_match_the_header_layout(tmp_path):
    # the ctypes mirrors of the ABI structs have the C compiler's sizes and
    # field offsets (a C program built against the header prints them)
    structs = {"deac_problem": da.Problem, "deac_run_options": da.RunOptions, "deac_result": da.Result,
               "deac_run_profile": da.RunProfile, "deac_batch_schedule": da.Schedule}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "deac_engine.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HEADER), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in structs.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, f"{cname}.{field}"
