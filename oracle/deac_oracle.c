/*
 * deac_oracle.c -- CPU restatement of the reference DEAC CPU path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
 * product.  Compile with -ffp-contract=off: every fused multiply-add the
 * reference build performs is an explicit fma() below, every other product
 * is rounded on its own.
 *
 * Reference: nscottnichols/deac-cpp 2.0.0-rc1, src/deac/src/deac.cpp and
 * src/deac/include/rng.hpp.  The reference build that fixes the evaluation
 * order is g++ 11.4 -O3 -std=gnu++20 -fno-math-errno -march=x86-64-v3:
 *   - matrix_multiply_LxM_by_MxN (deac.cpp:110-118) stays out of line and is
 *     a scalar FMA chain over the genome;
 *   - matrix_multiply_MxN_by_Nx1 (deac.cpp:102-108) and
 *     reduced_chi_square_statistic (deac.cpp:120-126) are inlined into deac()
 *     and vectorised as in-order ("fold-left") reductions: products of 4-wide
 *     blocks are rounded, then added one at a time; a 2-wide block follows;
 *     a last odd element is fused;
 *   - squared_mean (deac.cpp:171-177) is vectorised the same way with a
 *     scalar fused tail of up to three elements.
 * These shapes are pinned by the golden histories under tests/golden/.
 */
#include "deac_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ RNG */

uint64_t dor_splitmix64(uint64_t *state)                 /* rng.hpp:10-15 */
{
    uint64_t z = (*state += 0x9E3779B97f4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k)         /* rng.hpp:17-19 */
{
    return (x << k) | (x >> (64 - k));
}

uint64_t dor_xoshiro256p_next(uint64_t s[4])            /* rng.hpp:25-39 */
{
    const uint64_t out = s[0] + s[3];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return out;
}

void dor_xoshiro256p_init(uint64_t seed, uint64_t s[4]) /* rng.hpp:46-59 */
{
    /* Each state word keeps only 32 bits of a splitmix64 output. */
    uint64_t sm = seed;
    uint64_t a = dor_splitmix64(&sm);
    s[0] = (uint32_t)a;
    s[1] = (uint32_t)(a >> 32);
    uint64_t b = dor_splitmix64(&sm);
    s[2] = (uint32_t)b;
    s[3] = (uint32_t)(b >> 32);
}

void dor_xoshiro256p_jump(uint64_t s[4])                /* rng.hpp:64-87 */
{
    static const uint64_t poly[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL,
                                     0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL};
    uint64_t acc[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; w++)
        for (int bit = 0; bit < 64; bit++) {
            if (poly[w] & (1ULL << bit))
                for (int k = 0; k < 4; k++) acc[k] ^= s[k];
            dor_xoshiro256p_next(s);
        }
    memcpy(s, acc, sizeof acc);
}

/* deac.cpp:437 etc.: (x >> 11) * 0x1.0p-53 */
static inline double u01(uint64_t s[4])
{
    return (double)(dor_xoshiro256p_next(s) >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------ set-up */

void dor_frequency_grid(size_t G, double omega_max, double *f) /* deac.cpp:2120-2125 */
{
    double d = omega_max / (double)(G - 1);
    for (size_t i = 0; i < G; i++) f[i] = (double)i * d;
}

double dor_trapezoid_weight(const double *f, size_t G, size_t j) /* deac.cpp:352-358 */
{
    if (j == 0) return 0.5 * (f[j + 1] - f[j]);
    if (j == G - 1) return 0.5 * (f[j] - f[j - 1]);
    return 0.5 * (f[j + 1] - f[j - 1]);
}

static double periodicity_of(int spectra_type)          /* deac.cpp:238-247 */
{
    return (spectra_type == DOR_SPFSF || spectra_type == DOR_FFULL) ? -1.0 : 1.0;
}

void dor_kernel_terms(const dor_problem *p, double *kp, double *km) /* deac.cpp:318-401 */
{
    const size_t T = p->number_of_timeslices, G = p->genome_size;
    const double *f = p->frequency;
    const double per = periodicity_of(p->spectra_type);
    const double beta = 1.0 / p->temperature;
    for (size_t i = 0; i < T; i++) {
        const double t = p->imaginary_time[i];
        for (size_t j = 0; j < G; j++) {
            const double w = f[j];
            const double df = dor_trapezoid_weight(f, G, j);
            const size_t k = i * G + j;
            if (p->zero_temperature) {                  /* deac.cpp:392-399, positive side */
                kp[k] = df * exp(-t * w);
                continue;
            }
            if (p->spectra_type == DOR_BDSF) {          /* deac.cpp:322, 330, 337, 362, 370, 378 */
                if (p->model == DOR_MODEL_HYPERBOLIC) {
                    const double bo2mt = 0.5 * beta - t;
                    kp[k] = df * cosh(bo2mt * w);
                } else {
                    const double bmt = beta - t;
                    const double num = exp(-bmt * w) + exp(-t * w);
                    if (p->model == DOR_MODEL_NORMALIZATION) {
                        const double den = 1.0 + exp(-beta * w);
                        kp[k] = df * (num / den);
                    } else {
                        kp[k] = df * num;
                    }
                }
                continue;
            }
            switch (p->model) {
            case DOR_MODEL_HYPERBOLIC: {                /* deac.cpp:323-324, 363-364 */
                /* The reference build hoists -0.5*beta and 1.5*beta out of the
                 * loop: both subtractions round the product first. */
                const double bo2mtp = -0.5 * beta - t;
                const double bo2mtn = 1.5 * beta - t;
                kp[k] = per * df * exp(bo2mtp * w) / 2;
                km[k] = per * df * exp(-bo2mtn * w) / 2;
                break;
            }
            case DOR_MODEL_NORMALIZATION: {             /* deac.cpp:376-384 */
                const double tmb = t - beta;
                const double num = exp(-t * w);
                const double den = 1.0 + exp(-beta * w);
                kp[k] = df * (num / den);
                const double e = exp(tmb * w);
                const double den_n = 1.0 + exp(-beta * w);
                km[k] = per * df * (e - e / den_n);
                break;
            }
            default: {                                  /* deac.cpp:367-370 */
                const double tmb = t - beta;
                kp[k] = df * exp(-t * w);
                km[k] = per * df * exp(tmb * w);
                break;
            }
            }
        }
    }
}

/* which = 0: normalization terms (deac.cpp:495-538); 1: first-moment terms
 * (deac.cpp:628-668); 3: third-moment terms (deac.cpp:728-768). */
void dor_moment_terms(const dor_problem *p, int which, double *tp, double *tm)
{
    const size_t G = p->genome_size;
    const double *f = p->frequency;
    const double beta = 1.0 / p->temperature;
    for (size_t j = 0; j < G; j++) {
        const double w = f[j];
        const double df = dor_trapezoid_weight(f, G, j);
        /* wp: power of w carried by the term, written as the reference
         * writes it (df*f, df*pow(f,3)*f). */
        double a, b;
        if (p->zero_temperature) {                      /* deac.cpp:531-535, 662-666, 762-766 */
            if (which == 0) a = df;
            else if (which == 1) a = df * w;
            else a = df * pow(w, 3) * w;
            tp[j] = a;
            continue;
        }
        if (p->spectra_type == DOR_BDSF) {              /* deac.cpp:508-526, 645-661, 745-761 */
            double c;                                   /* the model's factor */
            if (p->model == DOR_MODEL_HYPERBOLIC)
                c = which == 0 ? cosh(0.5 * beta * w) : sinh(0.5 * beta * w);
            else if (p->model == DOR_MODEL_NORMALIZATION)
                c = which == 0 ? 1.0 : tanh(0.5 * beta * w);
            else
                c = which == 0 ? 1.0 + exp(-beta * w) : 1.0 - exp(-beta * w);
            if (which == 0) a = p->model == DOR_MODEL_NORMALIZATION ? df : df * c;
            else if (which == 1) a = df * w * c;
            else a = df * pow(w, 3) * w * c;
            tp[j] = a;
            continue;
        }
        switch (p->model) {
        case DOR_MODEL_HYPERBOLIC: {
            const double e = exp(-0.5 * beta * w);
            if (which == 0) { a = 0.5 * df / e; b = 0.5 * df * e; }
            else if (which == 1) { a = 0.5 * df * w / e; b = -0.5 * df * w * e; }
            else { a = 0.5 * df * pow(w, 3) * w / e; b = -0.5 * df * pow(w, 3) * w * e; }
            break;
        }
        case DOR_MODEL_NORMALIZATION: {
            const double l = 1.0 / (1.0 + exp(-beta * w));
            const double e = exp(-beta * w);
            const double r = e / (1.0 + e);
            if (which == 0) { a = df * l; b = df * r; }
            else if (which == 1) { a = df * w * l; b = -df * w * r; }
            else { a = df * pow(w, 3) * w * l; b = -df * pow(w, 3) * w * r; }
            break;
        }
        default:
            if (which == 0) { a = df; b = df; }
            else if (which == 1) { a = df * w; b = -df * w; }
            else { a = df * pow(w, 3) * w; b = -df * pow(w, 3) * w; }
            break;
        }
        tp[j] = a;
        if (tm) tm[j] = b;
    }
}

int dor_negative_first_moment(const dor_problem *p, double *dt, double *value, double *error)
{
    /* deac.cpp:879-894.  The error keeps only the last timeslice's term (the
     * reference assigns it in the loop rather than accumulating it). */
    if (!(p->spectra_type == DOR_BDSF && !p->zero_temperature && p->use_negative_first_moment)) return 0;
    const size_t T = p->number_of_timeslices;
    double v = 0.0, e = 0.0;
    for (size_t j = 0; j < T; j++) {
        const double d = dor_trapezoid_weight(p->imaginary_time, T, j);
        dt[j] = d;
        v = fma(p->isf[j], d, v);
        e = (p->isf_error[j] * p->isf_error[j]) * (d * d);
    }
    *value = v;
    *error = sqrt(e);
    return 1;
}

/* ------------------------------------------------------------ reductions */

double dor_dot_fold4(double acc, const double *a, const double *b, size_t n)
{
    size_t i = 0;
    if (n >= 4)
        for (; i + 4 <= n; i += 4)
            for (size_t l = 0; l < 4; l++) acc += a[i + l] * b[i + l];
    if (n - i >= 2) {
        acc += a[i] * b[i];
        acc += a[i + 1] * b[i + 1];
        i += 2;
    }
    if (n - i == 1) acc = fma(a[i], b[i], acc);
    return acc;
}

double dor_fma_chain(double acc, const double *a, const double *b, size_t n)
{
    for (size_t k = 0; k < n; k++) acc = fma(a[k], b[k], acc);
    return acc;
}

double dor_chi_square(const double *o, const double *e, const double *s, size_t n)
{
    double acc = 0.0;
    size_t i = 0;
    if (n >= 4)
        for (; i + 4 <= n; i += 4)
            for (size_t l = 0; l < 4; l++) {
                const double d = (o[i + l] - e[i + l]) / s[i + l];
                acc += d * d;
            }
    if (n - i >= 2)
        for (int l = 0; l < 2; l++, i++) {
            const double d = (o[i] - e[i]) / s[i];
            acc += d * d;
        }
    if (n - i == 1) {
        const double d = (o[i] - e[i]) / s[i];
        acc = fma(d, d, acc);
    }
    return acc;
}

double dor_mean(const double *a, size_t n)
{
    double acc = 0.0;
    for (size_t i = 0; i < n; i++) acc += a[i];
    return acc / (double)n;
}

double dor_squared_mean(const double *a, size_t n)
{
    double acc = 0.0;
    size_t i = 0;
    for (; i + 4 <= n; i += 4)
        for (size_t l = 0; l < 4; l++) acc += a[i + l] * a[i + l];
    for (; i < n; i++) acc = fma(a[i], a[i], acc);
    return acc / (double)n;
}

static double minimum_of(const double *a, size_t n)      /* deac.cpp:128-136 */
{
    double m = a[0];
    for (size_t i = 0; i < n; i++)
        if (a[i] < m) m = a[i];
    return m;
}

void dor_set_mutant_indices(uint64_t s[4], size_t *mi, size_t i0, size_t P) /* deac.cpp:200-216 */
{
    mi[0] = mi[1] = mi[2] = i0;
    while (mi[0] == i0) mi[0] = dor_xoshiro256p_next(s) % P;
    while (mi[1] == i0 || mi[1] == mi[0]) mi[1] = dor_xoshiro256p_next(s) % P;
    while (mi[2] == i0 || mi[2] == mi[0] || mi[2] == mi[1]) mi[2] = dor_xoshiro256p_next(s) % P;
}

/* ------------------------------------------------------------ the run */
/* The contraction of the trial population (deac.cpp:1389-1394).  The
 * two-sign builds call matrix_multiply_LxM_by_MxN out of line (a scalar FMA
 * chain).  The detailed-balance builds (bdsf, the zero-temperature one
 * included) have one call in the loop, which the reference build inlines
 * and vectorises like matrix_multiply_MxN_by_Nx1: a fold-left of rounded
 * products with the last odd element fused (dor_dot_fold4).  Their initial
 * population (deac.cpp:860) still goes through the out-of-line FMA chain.
 * Read from the x86 assembly of the reference build; pinned by the bdsf and
 * zero-temperature goldens. */
static int loop_contraction_folds(const dor_problem *p)
{
    return p->spectra_type == DOR_BDSF;
}

/* The -1st-moment term of the bosonic detailed-balance build (deac.cpp:
 * 868-920, 1398-1416): weights, target value and error; used = 0 when off. */
typedef struct nfm_term {
    int used;
    double value, error;
    double *dt;
} nfm_term;

/* fitness of one member, deac.cpp:957-976 (initial) and 1539-1555 (loop) */
static double member_fitness(const dor_problem *p, const double *model_isf, double first_i, double third_i,
                             const nfm_term *nfm)
{
    const size_t T = p->number_of_timeslices;
    double fit = dor_chi_square(p->isf, model_isf, p->isf_error, T) / (double)T;
    if (nfm->used) {                                    /* deac.cpp:961-965, 1543-1547 */
        const double m = dor_dot_fold4(0.0, model_isf, nfm->dt, T);
        const double q = (nfm->value - m) / nfm->error;
        fit = fma(q, q, fit);
    }
    if (p->first_moment >= 0.0) {
        const double d = first_i - p->first_moment;
        fit += (d * d) / p->first_moment;
    }
    if (p->third_moment >= 0.0) {
        const double q = (p->third_moment - third_i) / p->third_moment_error;
        fit = fma(q, q, fit);
    }
    return fit;
}

int dor_run(const dor_problem *p, uint64_t rng[4], dor_result *r)
{
    const size_t T = p->number_of_timeslices, G = p->genome_size, P = p->population_size;
    const size_t N = p->number_of_generations;
    const int nsign = (p->zero_temperature || p->spectra_type == DOR_BDSF) ? 1 : 2;
    const size_t row = (size_t)nsign * G;       /* doubles per member, [sign][gene] */
    nfm_term nfm = {0, 0.0, 0.0, malloc(sizeof(double) * T)};
    nfm.used = dor_negative_first_moment(p, nfm.dt, &nfm.value, &nfm.error);
    const int use_first = p->first_moment >= 0.0, use_third = p->third_moment >= 0.0;
    const double zeroth_moment = p->isf[0];     /* deac.cpp:306 */

    double *kern = malloc(sizeof(double) * T * row);
    double *kp = kern, *km = nsign == 2 ? kern + T * G : NULL;
    dor_kernel_terms(p, kp, km);

    double *pop_old = malloc(sizeof(double) * P * row);
    double *pop_new = malloc(sizeof(double) * P * row);
    double *cr_old = malloc(sizeof(double) * P * nsign), *cr_new = malloc(sizeof(double) * P * nsign);
    double *fw_old = malloc(sizeof(double) * P * nsign), *fw_new = malloc(sizeof(double) * P * nsign);
    double *fitness = malloc(sizeof(double) * P);
    double *model_isf = malloc(sizeof(double) * P * T);
    double *first = calloc(P, sizeof(double)), *third = calloc(P, sizeof(double));
    double *nterm = malloc(sizeof(double) * row), *fterm = malloc(sizeof(double) * row);
    double *tterm = malloc(sizeof(double) * row);
    unsigned char *mask = malloc(row);
    size_t mutant[3];

    /* initial population, deac.cpp:430-458 */
    for (size_t i = 0; i < P; i++)
        for (size_t j = 0; j < G; j++)
            for (int s = 0; s < nsign; s++) {
                double u = u01(rng);
                if (p->allow_negative_spectral_weight) u -= 0.5;
                pop_old[i * row + s * G + j] = u;
            }
    if (nsign == 2)
        for (size_t i = 0; i < P; i++) pop_old[i * row + G] = pop_old[i * row];

    /* normalization terms and initial normalization, deac.cpp:489-603 */
    if (p->normalize || p->renormalize_each_generation)
        dor_moment_terms(p, 0, nterm, nsign == 2 ? nterm + G : NULL);
    if (p->normalize)
        for (size_t i = 0; i < P; i++) {
            double *x = pop_old + i * row;
            double norm = 0.0;
            for (int s = 0; s < nsign; s++) norm = dor_dot_fold4(norm, x + s * G, nterm + s * G, G);
            for (size_t j = 0; j < row; j++) x[j] *= zeroth_moment / norm;
        }
    if (use_first) dor_moment_terms(p, 1, fterm, nsign == 2 ? fterm + G : NULL);
    if (use_third) dor_moment_terms(p, 3, tterm, nsign == 2 ? tterm + G : NULL);

    /* moments, model ISF and fitness of the initial population, deac.cpp:605-976 */
    for (size_t i = 0; i < P; i++) {
        const double *x = pop_old + i * row;
        if (use_first) {
            double m = 0.0;
            for (int s = 0; s < nsign; s++) m = dor_dot_fold4(m, x + s * G, fterm + s * G, G);
            first[i] = m;
        }
        if (use_third) {
            double m = 0.0;
            for (int s = 0; s < nsign; s++) m = dor_dot_fold4(m, x + s * G, tterm + s * G, G);
            third[i] = m;
        }
        for (size_t t = 0; t < T; t++) {
            double e = 0.0;
            for (int s = 0; s < nsign; s++) e = dor_fma_chain(e, kern + s * T * G + t * G, x + s * G, G);
            model_isf[i * T + t] = e;
        }
        fitness[i] = member_fitness(p, model_isf + i * T, first[i], third[i], &nfm);
    }

    /* control parameters, deac.cpp:990-1032 */
    for (size_t i = 0; i < P * nsign; i++) {
        cr_old[i] = p->crossover_probability;
        fw_old[i] = p->differential_weight;
    }

    size_t generation = 0;
    double minimum_fitness;
    for (size_t ii = 0; ii + 1 < N; ii++) {                 /* deac.cpp:1138 */
        generation = ii;
        minimum_fitness = minimum_of(fitness, P);           /* deac.cpp:1146 */
        if (p->track_stats) {                               /* deac.cpp:1149-1161 */
            if (r->stats_mean) r->stats_mean[ii] = dor_mean(fitness, P);
            if (r->stats_minimum) r->stats_minimum[ii] = minimum_fitness;
            if (r->stats_squared_mean) r->stats_squared_mean[ii] = dor_squared_mean(fitness, P);
        }
        if (minimum_fitness <= p->stop_minimum_fitness) break; /* deac.cpp:1163-1166 */

        /* jDE control update, deac.cpp:1180-1213: CR+, F+, CR-, F- per member */
        for (size_t i = 0; i < P; i++)
            for (int s = 0; s < nsign; s++) {
                const size_t k = i * nsign + s;
                if (u01(rng) < p->self_adapting_crossover_probability) cr_new[k] = u01(rng);
                else cr_new[k] = cr_old[k];
                if (u01(rng) < p->self_adapting_differential_weight_probability) fw_new[k] = 2.0 * u01(rng);
                else fw_new[k] = fw_old[k];
            }

        /* mutant indices, crossover masks and trial vectors, deac.cpp:1229-1311.
         * The reference draws all masks of member i before building any
         * trial; building member i's trial right after its masks reads only
         * pop_old, so the result is the same. */
        for (size_t i = 0; i < P; i++) {
            dor_set_mutant_indices(rng, mutant, i, P);
            for (int s = 0; s < nsign; s++) {
                const double cr = cr_new[i * nsign + s];
                for (size_t j = 0; j < G; j++) mask[s * G + j] = u01(rng) < cr;
            }
            const double *x1 = pop_old + mutant[0] * row, *x2 = pop_old + mutant[1] * row;
            const double *x3 = pop_old + mutant[2] * row, *xp = pop_old + i * row;
            double *y = pop_new + i * row;
            for (int s = 0; s < nsign; s++) {
                const double F = fw_new[i * nsign + s];
                for (size_t j = 0; j < G; j++) {
                    const size_t k = s * G + j;
                    if (mask[k]) {
                        const double v = fma(F, x2[k] - x3[k], x1[k]);
                        y[k] = p->allow_negative_spectral_weight ? v : fabs(v);
                    } else {
                        y[k] = xp[k];
                    }
                }
            }
            if (nsign == 2) y[G] = y[0];                    /* deac.cpp:1305-1308 */
        }

        /* per-generation renormalization, deac.cpp:1342-1364 */
        if (p->normalize && p->renormalize_each_generation)
            for (size_t i = 0; i < P; i++) {
                double *y = pop_new + i * row;
                double norm = 0.0;
                for (int s = 0; s < nsign; s++) norm = dor_dot_fold4(norm, y + s * G, nterm + s * G, G);
                for (size_t j = 0; j < row; j++) y[j] *= zeroth_moment / norm;
            }

        /* model ISF, moments, fitness and greedy selection, deac.cpp:1385-1573 */
        for (size_t i = 0; i < P; i++) {
            const double *y = pop_new + i * row;
            double fm = 0.0, tm = 0.0;
            if (use_first)
                for (int s = 0; s < nsign; s++) fm = dor_dot_fold4(fm, y + s * G, fterm + s * G, G);
            if (use_third)
                for (int s = 0; s < nsign; s++) tm = dor_dot_fold4(tm, y + s * G, tterm + s * G, G);
            for (size_t t = 0; t < T; t++) {
                double e = 0.0;
                if (loop_contraction_folds(p))
                    e = dor_dot_fold4(0.0, kern + t * G, y, G);
                else
                    for (int s = 0; s < nsign; s++) e = dor_fma_chain(e, kern + s * T * G + t * G, y + s * G, G);
                model_isf[i * T + t] = e;
            }
            const double fit = member_fitness(p, model_isf + i * T, fm, tm, &nfm);
            if (fit <= fitness[i]) {
                fitness[i] = fit;
                for (int s = 0; s < nsign; s++) {
                    cr_old[i * nsign + s] = cr_new[i * nsign + s];
                    fw_old[i * nsign + s] = fw_new[i * nsign + s];
                }
                memcpy(pop_old + i * row, y, sizeof(double) * row);
            }
        }
    }

    /* best member and output transform, deac.cpp:1592-1675 */
    size_t best = 0;
    minimum_fitness = fitness[0];
    for (size_t i = 0; i < P; i++)
        if (fitness[i] < minimum_fitness) { minimum_fitness = fitness[i]; best = i; }
    const double beta = 1.0 / p->temperature;
    const double *xb = pop_old + best * row;
    for (size_t i = 0; i < G; i++) {
        const double w = p->frequency[i];
        if (nsign == 1) {                               /* deac.cpp:1605-1608, 1620-1671 */
            double v = xb[i];
            if (!p->zero_temperature) {                 /* bdsf */
                if (p->model == DOR_MODEL_HYPERBOLIC) v = 0.5 * xb[i] * exp(0.5 * beta * w);
                else if (p->model == DOR_MODEL_NORMALIZATION) v = xb[i] / (1.0 + exp(-beta * w));
            }
            r->frequency_out[i] = w;
            r->dsf[i] = v;
            continue;
        }
        const size_t ip = G + i - 1, in = G - i - 1;
        r->frequency_out[ip] = w;
        r->frequency_out[in] = -w;
        const double sp = xb[i], sn = xb[G + i];
        const int st = p->spectra_type;
        double vp, vn;
        switch (p->model) {
        case DOR_MODEL_HYPERBOLIC:
            if (st == DOR_SPBSF) {
                vp = -2.0 * M_PI * sp * sinh(beta * w / 2);
                vn = 2.0 * M_PI * sn * sinh(beta * w / 2);
            } else if (st == DOR_SPFSF) {
                vp = -2.0 * M_PI * sp * cosh(beta * w / 2);
                vn = -2.0 * M_PI * sn * cosh(beta * w / 2);
            } else {
                vp = 0.5 * sp * exp(0.5 * beta * w);
                vn = 0.5 * sn * exp(-0.5 * beta * w);
            }
            break;
        case DOR_MODEL_NORMALIZATION:
            if (st == DOR_SPBSF) {
                vp = -2.0 * M_PI * sp * tanh(0.5 * beta * w);
                vn = 2.0 * M_PI * sn * tanh(0.5 * beta * w);
            } else if (st == DOR_SPFSF) {
                vp = -2.0 * M_PI * sp;
                vn = -2.0 * M_PI * sn;
            } else {
                vp = sp / (1.0 + exp(-beta * w));
                vn = sn / (1.0 + exp(beta * w));
            }
            break;
        default:
            if (st == DOR_SPBSF) {
                vp = -2.0 * M_PI * sp * (1.0 - exp(-beta * w));
                vn = -2.0 * M_PI * sn * (1.0 - exp(beta * w));
            } else if (st == DOR_SPFSF) {
                vp = -2.0 * M_PI * sp * (1.0 + exp(-beta * w));
                vn = -2.0 * M_PI * sn * (1.0 + exp(beta * w));
            } else {
                vp = sp;
                vn = sn;
            }
            break;
        }
        r->dsf[ip] = vp;
        r->dsf[in] = vn;
    }

    /* final statistics entry, deac.cpp:1677-1685 */
    if (generation + 2 == N) {
        generation += 1;
        if (p->track_stats) {
            if (r->stats_mean) r->stats_mean[generation] = dor_mean(fitness, P);
            if (r->stats_minimum) r->stats_minimum[generation] = minimum_fitness;
            if (r->stats_squared_mean) r->stats_squared_mean[generation] = dor_squared_mean(fitness, P);
        }
    }
    r->generation = generation;
    r->minimum_fitness = minimum_fitness;
    r->best_index = best;
    if (r->final_fitness) memcpy(r->final_fitness, fitness, sizeof(double) * P);
    if (r->final_population) memcpy(r->final_population, pop_old, sizeof(double) * P * row);

    free(kern); free(pop_old); free(pop_new); free(cr_old); free(cr_new);
    free(fw_old); free(fw_new); free(fitness); free(model_isf); free(first);
    free(third); free(nterm); free(fterm); free(tterm); free(mask); free(nfm.dt);
    return 0;
}
