#!/usr/bin/env python3
"""The uneven batch: 64 wavevectors x 16 seeds = 1024 reconstructions of
different shapes in one batched run.

Per wavevector k = 1..64, q_k = 0.5 k, drawn from numpy default_rng(5) in this
order (one draw of each per k, k ascending):

  w        ~ U[0.2, 2]          width factor of the main peak
  T        from {33, 65, 129, 257} timeslices
  G        from {1024, 2048, 4096, 8192} frequencies per sign
  omega    ~ U[128, 512]        omega_max
  model    standard or normalization, 50/50

then one wavevector, drawn last, whose 16 instances use the zero-temperature
kernel (with the standard model).

Spectrum (a choice; the main peak follows the workload description, the second
is ours): S(w) = 0.7 N(2 q, w q) + 0.3 N(6 q, 1.5 w q), a main excitation and
a broad continuum at three times its frequency, normalised to unit zeroth
moment on [0, omega_max] (16384-point trapezoid grid).

ISF on T uniform tau in [0, beta/2], beta = 1/1.35: F(tau) = int S(w) (e^{-tau w}
+ e^{-(beta-tau) w}) dw, or int S(w) e^{-tau w} dw for the zero-temperature
instances; sigma = 1e-4 F; noise N(0, sigma) from default_rng(1000 + k).
Every instance: population 16, bfull spectra, --normalize, seeds 1..16.

`python make_uneven.py` rewrites uneven_table.json (the committed table);
`instances()` rebuilds the inputs from it.
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "uneven_table.json")
TEMPERATURE = 1.35
SEEDS = list(range(1, 17))
POPULATION = 16


def draw_table():
    rng = np.random.default_rng(5)
    rows = []
    for k in range(1, 65):
        w = float(rng.uniform(0.2, 2.0))
        T = int(rng.choice([33, 65, 129, 257]))
        G = int(rng.choice([1024, 2048, 4096, 8192]))
        omega_max = float(rng.uniform(128.0, 512.0))
        model = "standard" if rng.uniform() < 0.5 else "normalization"
        rows.append({"k": k, "q": 0.5 * k, "w": w, "T": T, "G": G, "omega_max": omega_max, "model": model,
                     "zero_temperature": False})
    zt = int(rng.integers(0, 64))
    rows[zt]["zero_temperature"] = True
    rows[zt]["model"] = "standard"
    return rows


def gaussian(w, mu, sd):
    return np.exp(-0.5 * ((w - mu) / sd) ** 2) / (sd * np.sqrt(2.0 * np.pi))


def isf_of(row, ngrid=16384):
    beta = 1.0 / TEMPERATURE
    q, wf = row["q"], row["w"]
    w = np.linspace(0.0, row["omega_max"], ngrid)
    s = 0.7 * gaussian(w, 2.0 * q, wf * q) + 0.3 * gaussian(w, 6.0 * q, 1.5 * wf * q)
    s /= np.trapezoid(s, w)
    tau = np.linspace(0.0, 0.5 * beta, row["T"])
    if row["zero_temperature"]:
        f = np.array([np.trapezoid(s * np.exp(-t * w), w) for t in tau])
    else:
        f = np.array([np.trapezoid(s * (np.exp(-t * w) + np.exp(-(beta - t) * w)), w) for t in tau])
    sigma = 1e-4 * f
    return tau, f + np.random.default_rng(1000 + row["k"]).normal(0.0, sigma), sigma


def load_table():
    with open(TABLE) as fh:
        return json.load(fh)


def instances(table=None):
    """One dict per reconstruction (wavevector-major, seeds 1..16): inputs and
    run parameters."""
    out = []
    for row in table or load_table():
        tau, f, sigma = isf_of(row)
        for seed in SEEDS:
            out.append(dict(row, seed=seed, tau=tau, isf=f, err=sigma, P=POPULATION))
    return out


if __name__ == "__main__":
    with open(TABLE, "w") as fh:
        json.dump(draw_table(), fh, indent=1)
