/*
 * deac_oracle.h -- CPU restatement of the reference DEAC CPU path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (libdeac_amd.so, deac.e) never links or calls it.
 *
 * Every function restates a piece of src/deac/src/deac.cpp or
 * src/deac/include/rng.hpp of nscottnichols/deac-cpp 2.0.0-rc1 and cites the
 * lines it follows.  Floating-point evaluation order follows the reference as
 * built with g++ 11 -O3 -std=gnu++20 (-ffp-contract=fast): the places where
 * that build fuses a multiply-add are written as explicit fma() calls and the
 * rest of this file is compiled with -ffp-contract=off.
 */
#ifndef DEAC_ORACLE_H
#define DEAC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* spectral models (reference: -DDEAC_MODEL, src/deac/CMakeLists.txt:97-109) */
enum { DOR_MODEL_STANDARD = 0, DOR_MODEL_HYPERBOLIC = 1, DOR_MODEL_NORMALIZATION = 2 };
/* spectra types (reference: --spectra_type, deac.cpp:2040-2058).  DOR_BDSF is
 * the only type of the bosonic detailed-balance build
 * (USE_BOSONIC_DETAILED_BALANCE_CONDITION_DSF, deac.cpp:2016-2018, 2041-2052):
 * one sign, the detailed-balance kernel, G-length output. */
enum { DOR_SPBSF = 0, DOR_SPFSF = 1, DOR_BFULL = 2, DOR_FFULL = 3, DOR_BDSF = 4 };

typedef struct dor_problem {
    size_t number_of_timeslices;          /* T */
    const double *imaginary_time;         /* T */
    const double *isf;                    /* T */
    const double *isf_error;              /* T */
    size_t genome_size;                   /* G */
    const double *frequency;              /* G */
    double temperature;
    int model;
    int spectra_type;
    int zero_temperature;                 /* single-sign kernel e^{-tau w} */
    int allow_negative_spectral_weight;   /* reference ALLOW_NEGATIVE_SPECTRAL_WEIGHT */
    int renormalize_each_generation;      /* reference SINGLE_PARTICLE_FERMIONIC_SPECTRAL_FUNCTION */
    size_t number_of_generations;         /* N */
    size_t population_size;               /* P */
    int normalize;
    double first_moment;                  /* < 0: unused */
    double third_moment;                  /* < 0: unused */
    double third_moment_error;
    double crossover_probability;
    double self_adapting_crossover_probability;
    double differential_weight;
    double self_adapting_differential_weight_probability;
    double stop_minimum_fitness;
    int track_stats;
    int use_negative_first_moment;        /* bdsf at T > 0: the -1st-moment term (deac.cpp:868-920) */
} dor_problem;

typedef struct dor_result {
    size_t generation;
    double minimum_fitness;
    size_t best_index;
    double *dsf;                 /* out: 2G-1 values (G when zero_temperature or bdsf) */
    double *frequency_out;       /* out: same length */
    double *stats_mean;          /* out: N values when track_stats, may be NULL */
    double *stats_minimum;
    double *stats_squared_mean;
    double *final_fitness;       /* out: P values, may be NULL */
    double *final_population;    /* out: P*nsign*G values [i][sign][j], may be NULL */
} dor_result;

/* rng.hpp:10-15 */
uint64_t dor_splitmix64(uint64_t *state);
/* rng.hpp:25-39 */
uint64_t dor_xoshiro256p_next(uint64_t s[4]);
/* rng.hpp:46-59 */
void dor_xoshiro256p_init(uint64_t seed, uint64_t s[4]);
/* rng.hpp:64-87 */
void dor_xoshiro256p_jump(uint64_t s[4]);

/* deac.cpp:2120-2125: frequency[i] = i*(omega_max/(G-1)) */
void dor_frequency_grid(size_t genome_size, double omega_max, double *frequency);
/* deac.cpp:352-358 */
double dor_trapezoid_weight(const double *frequency, size_t genome_size, size_t j);

/* deac.cpp:318-401: K+ and K- (T x G row-major).  km may be NULL for zero T. */
void dor_kernel_terms(const dor_problem *p, double *kp, double *km);
/* deac.cpp:495-538 (which = 0), 628-668 (which = 1), 728-768 (which = 3) */
void dor_moment_terms(const dor_problem *p, int which, double *tp, double *tm);

/* deac.cpp:875-894: the -1st-moment weights dt (T), the moment of the ISF
 * and its error; returns 0 unless the term is used (bdsf, T > 0). */
int dor_negative_first_moment(const dor_problem *p, double *dt, double *value, double *error);

/* deac.cpp:102-108 as inlined into deac() by the reference build */
double dor_dot_fold4(double acc, const double *a, const double *b, size_t n);
/* deac.cpp:110-118 (out-of-line scalar FMA chain) */
double dor_fma_chain(double acc, const double *a, const double *b, size_t n);
/* deac.cpp:120-126 as inlined into deac() */
double dor_chi_square(const double *observed, const double *calculated, const double *error, size_t n);
/* deac.cpp:163-169, 171-177 as inlined into deac() */
double dor_mean(const double *a, size_t n);
double dor_squared_mean(const double *a, size_t n);
/* deac.cpp:200-216 */
void dor_set_mutant_indices(uint64_t s[4], size_t *mutant_indices, size_t mutant_index0, size_t length);

/* deac.cpp:218-1923, CPU branches.  rng is advanced in place.  Returns 0. */
int dor_run(const dor_problem *p, uint64_t rng[4], dor_result *r);

#ifdef __cplusplus
}
#endif
#endif
