#!/usr/bin/env python3
"""Write the synthetic isf.bin inputs of the benchmark workloads.

Each file holds [tau(T) | F(T) | sigma(T)] as little-endian float64, the layout
deac.e reads (reference: deac.cpp:2060-2074, tools/convert_isf_data.py:13-19).

  isf_c1_T33.bin   S = 0.6 N(2,0.5) + 0.4 N(8,1.5) on [0,64], beta = 1/1.35,
                   33 tau on [0, beta/2], sigma = 1e-3 F, seeded N(0,sigma) noise
  isf_c2_T65.bin   three peaks (mu,sd,w) = (3,0.4,.5), (20,4,.3), (60,10,.2)
                   on [0,512], 65 tau, sigma = 1e-4 F
  isf_c3_T257.bin  the same spectrum with 257 tau

The 1024 inputs of the uneven batch are built by make_uneven.py from the
committed uneven_table.json.

F(tau) = int S(w) (e^{-tau w} + e^{-(beta-tau) w}) dw by the trapezoid rule on
a 65536-point grid; S is normalised to unit zeroth moment on the same grid.
Noise uses numpy default_rng(seed) with seed = the run's --seed (1).
"""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TEMPERATURE = 1.35


def gaussian(w, mu, sd):
    return np.exp(-0.5 * ((w - mu) / sd) ** 2) / (sd * np.sqrt(2.0 * np.pi))


def synthetic_isf(peaks, omega_max, ntau, rel_error, seed, temperature=TEMPERATURE, ngrid=65536):
    beta = 1.0 / temperature
    w = np.linspace(0.0, omega_max, ngrid)
    s = sum(wt * gaussian(w, mu, sd) for mu, sd, wt in peaks)
    s /= np.trapezoid(s, w)
    tau = np.linspace(0.0, 0.5 * beta, ntau)
    f = np.array([np.trapezoid(s * (np.exp(-t * w) + np.exp(-(beta - t) * w)), w) for t in tau])
    sigma = rel_error * f
    f_noisy = f + np.random.default_rng(seed).normal(0.0, sigma)
    return tau, f_noisy, sigma


def write_isf(path, tau, f, sigma):
    np.concatenate([tau, f, sigma]).astype("<f8").tofile(path)


CONFIG1_PEAKS = [(2.0, 0.5, 0.6), (8.0, 1.5, 0.4)]
CONFIG2_PEAKS = [(3.0, 0.4, 0.5), (20.0, 4.0, 0.3), (60.0, 10.0, 0.2)]


def main():
    write_isf(os.path.join(HERE, "isf_c1_T33.bin"), *synthetic_isf(CONFIG1_PEAKS, 64.0, 33, 1e-3, 1))
    write_isf(os.path.join(HERE, "isf_c2_T65.bin"), *synthetic_isf(CONFIG2_PEAKS, 512.0, 65, 1e-4, 1))
    write_isf(os.path.join(HERE, "isf_c3_T257.bin"), *synthetic_isf(CONFIG2_PEAKS, 512.0, 257, 1e-4, 1))


if __name__ == "__main__":
    main()
