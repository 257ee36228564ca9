// Host side of the DEAC engine: problem validation, the kernel matrix and
// moment vectors, the initial population, the output transform and the
// xoshiro256+ generator with GF(2) jump-ahead.  All FP64, evaluated on the
// host with glibc libm exactly as the reference evaluates it; the file is
// compiled with -ffp-contract=off and every fused multiply-add the reference
// build performs is an explicit std::fma.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "deac_engine.h"

namespace deac {

// Text of the last error on this thread (validation messages, unsupported
// combinations); deac_last_error_message returns it.
std::string &last_message();

// ---------------------------------------------------------------- RNG
// xoshiro256+ (rng.hpp:21-44) and its seeding (rng.hpp:46-59).
struct Xoshiro {
    uint64_t s[4];
    uint64_t next();
    double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // deac.cpp:437
};
void seed_state(uint64_t seed, uint64_t out[4]);

// 256x256 matrix over GF(2) acting on the xoshiro256+ state, stored by
// columns: col[b] = M * e_b.
struct Gf2Matrix {
    uint64_t col[256][4];
    static Gf2Matrix identity();
    static Gf2Matrix one_step();                  // one xoshiro256+ transition
    void apply(const uint64_t in[4], uint64_t out[4]) const;
    Gf2Matrix operator*(const Gf2Matrix &rhs) const;  // this o rhs
};
// T^n for the xoshiro256+ transition T (square-and-multiply, cached powers).
Gf2Matrix transition_power(uint64_t n);
// rows[r] = row r of M (bit b of rows[r] = bit r of col[b]): 256 x 4 words.
void matrix_rows(const Gf2Matrix &m, uint64_t *rows);
// 4-bit lookup table: tab[n][v] = M * (v << 4n), 64 x 16 x 4 words.
void nibble_table(const Gf2Matrix &m, uint64_t *tab);

// ---------------------------------------------------------------- model
struct Model {
    size_t T = 0, G = 0, P = 0;
    int nsign = 2;          // 2: positive and negative frequencies, 1: zero temperature or bdsf
    size_t Kc = 0;          // nsign * G, genome length across both signs
    int model = DEAC_MODEL_STANDARD, spectra = DEAC_SPFSF;
    bool zero_t = false, allow_negative = false, renormalize = false, normalize = false;
    bool bdsf = false;      // detailed-balance spectra: one sign; the trial contraction is a rounded-product fold
    bool use_first = false, use_third = false;
    bool use_nfm = false;   // -1st-moment term (bdsf above zero temperature)
    double nfm_value = 0, nfm_error = 0;
    std::vector<double> nfm_dt;   // T trapezoid weights over tau
    double temperature = 0, beta = 0, zeroth_moment = 0;
    double first_moment = -1, third_moment = -1, third_moment_error = 0;
    double cr0 = 0.9, tau_cr = 0.1, f0 = 0.9, tau_f = 0.1;
    std::vector<double> tau, isf, err, freq;
    std::vector<double> kernel;   // T x Kc, row-major, [t][sign*G + j]
    std::vector<double> nterm, fterm, tterm;  // Kc each (empty when unused)
};

// Returns DEAC_OK or DEAC_ERR_INVALID_ARGUMENT with the reference's message
// (deac.cpp:2075-2160).
deac_status validate(const deac_problem &p, std::string &message);
// Builds K (deac.cpp:310-401) and the moment vectors (deac.cpp:489-827).
Model build_model(const deac_problem &p);

// Initial population [P][Kc] (deac.cpp:430-458) and its normalization
// (deac.cpp:584-602); advances rng.
void initial_population(const Model &m, Xoshiro &rng, std::vector<double> &pop);

// In-order reductions of the reference build (see host_model.cpp).
double dot_fold4(double acc, const double *a, const double *b, size_t n);
double mean_of(const double *a, size_t n);
double squared_mean_of(const double *a, size_t n);

// Output transform and mirrored frequency grid (deac.cpp:1594-1675).
void output_transform(const Model &m, const double *best, double *dsf, double *frequency);

}  // namespace deac
