"""Golden cases produced by tests/golden/make_goldens.py from the reference
executables, and helpers that replay a case through the oracle or the engine.
"""
import glob
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN_DIR = os.path.join(HERE, "golden")

VARIANT_FLAGS = {
    "standard": dict(model="standard"),
    "normalization": dict(model="normalization"),
    "hyperbolic": dict(model="hyperbolic"),
    "standard_sp": dict(model="standard", renormalize=True),
    "standard_neg": dict(model="standard", allow_negative=True),
    "standard_bdsf": dict(model="standard"),
    "normalization_bdsf": dict(model="normalization"),
    "hyperbolic_bdsf": dict(model="hyperbolic"),
    "zerot_bdsf": dict(model="standard", zero_temperature=True),
}


def names():
    return sorted(os.path.basename(p)[len("golden_"):-len(".npz")]
                  for p in glob.glob(os.path.join(GOLDEN_DIR, "golden_*.npz")))


def load(name):
    z = np.load(os.path.join(GOLDEN_DIR, f"golden_{name}.npz"), allow_pickle=False)
    case = {k: z[k] for k in z.files if k != "meta"}
    case["meta"] = json.loads(str(z["meta"]))
    return case


def frequency_of(case, grid):
    a = case["meta"]["args"]
    if "frequency" in a:
        return np.asarray(a["frequency"], dtype=np.float64)
    return grid(a["M"], a["omega_max"])


def problem_kwargs(case):
    m = case["meta"]
    a = m["args"]
    kw = dict(VARIANT_FLAGS[m["variant"]])
    kw.update(spectra_type=m["spectra"], population_size=a["P"], normalize=bool(a.get("normalize", False)),
              first_moment=a.get("first_moment", -1.0), third_moment=a.get("third_moment", -1.0),
              third_moment_error=a.get("third_moment_error", 0.0),
              use_negative_first_moment=bool(a.get("use_negative_first_moment", False)))
    return kw


def bit_diff(a, b):
    """Index of the first bitwise difference of two float64 arrays, or None."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    if a.shape != b.shape:
        return f"shape {a.shape} != {b.shape}"
    neq = np.nonzero(a.view(np.uint64) != b.view(np.uint64))[0]
    if len(neq) == 0:
        return None
    i = neq[0]
    return f"first difference at {i}: {a[i]!r} vs {b[i]!r} ({len(neq)} entries differ)"


def compare_outputs(got, case, n_hist=None):
    """Bit-exact comparison of a run against a golden case.  Returns a list of
    mismatch descriptions (empty when identical)."""
    errors = []
    for key in ("minimum", "mean", "squared_mean"):
        ref = case[key]
        n = len(ref) if n_hist is None else min(n_hist, len(ref))
        d = bit_diff(np.asarray(got[key])[:n], ref[:n])
        if d:
            errors.append(f"{key}: {d}")
    if n_hist is None:
        for key, gkey in (("dsf", "dsf"), ("frequency_out", "frequency")):
            d = bit_diff(got[gkey], case[key])
            if d:
                errors.append(f"{key}: {d}")
        if got["generation"] != case["meta"]["generation"]:
            errors.append(f"generation {got['generation']} != {case['meta']['generation']}")
    return errors
