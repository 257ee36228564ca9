/*
 * deac_engine.h -- C ABI of the MI355X DEAC engine (libdeac_amd.so).
 *
 * The engine runs the DEAC generation loop on one gfx950 GPU.  It replaces
 * two levels of nscottnichols/deac-cpp 2.0.0-rc1 at once:
 *   - the function level: void deac(...) in src/deac/src/deac.cpp:218-229,
 *     whose loop is deac.cpp:1137-1574 (C++ signature, std::string/fs::path
 *     arguments, errors by exit());
 *   - the operator level: the gpu_* launchers of src/deac/include/
 *     deac_gpu.hip.hpp:651-761 (declared for CUDA in deac_gpu.cuh:6-28),
 *     about 25 launches and 12 stream syncs per generation.
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 *
 * Ownership: the caller owns every input and output buffer.  The engine
 * copies what it needs in deac_create and owns all device memory.
 * Errors: every entry point returns a deac_status; none calls exit().
 * Threading: a handle is used by one host thread at a time; distinct handles
 * are independent.
 */
#ifndef DEAC_ENGINE_H
#define DEAC_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEAC_ENGINE_VERSION "0.1.0"

typedef enum deac_status {
    DEAC_OK = 0,
    DEAC_ERR_INVALID_ARGUMENT = 1,  /* a problem field fails the reference's validation */
    DEAC_ERR_NO_DEVICE = 2,         /* no usable gfx950 device */
    DEAC_ERR_OUT_OF_MEMORY = 3,     /* device or host allocation failed */
    DEAC_ERR_DEVICE = 4,            /* a HIP call or kernel launch failed */
    DEAC_ERR_UNSUPPORTED = 5        /* a combination this build does not run */
} deac_status;

/* Spectral model (reference: compile-time -DDEAC_MODEL, CMakeLists.txt:97-109;
 * here a per-problem runtime choice). */
typedef enum deac_model {
    DEAC_MODEL_STANDARD = 0,
    DEAC_MODEL_HYPERBOLIC = 1,
    DEAC_MODEL_NORMALIZATION = 2
} deac_model;

/* Spectra type (reference: --spectra_type, deac.cpp:2040-2058).  DEAC_BDSF
 * is the reference's bosonic detailed-balance build
 * (USE_BOSONIC_DETAILED_BALANCE_CONDITION_DSF, src/deac/CMakeLists.txt:21,
 * 33-34), whose only spectra type is "bdsf" (deac.cpp:2016-2018): one sign
 * of genes, the detailed-balance kernel of the chosen model
 * (deac.cpp:322-394), G-length output (deac.cpp:1596-1606), the optional
 * -1st-moment term (deac.cpp:868-920), and the trial population's model ISF
 * summed as that build sums it: rounded products added in order, the last
 * of an odd count fused.  With zero_temperature it is the reference's ZeroT
 * build. */
typedef enum deac_spectra_type {
    DEAC_SPBSF = 0,
    DEAC_SPFSF = 1,
    DEAC_BFULL = 2,
    DEAC_FFULL = 3,
    DEAC_BDSF = 4
} deac_spectra_type;

/* Everything deac() takes except the RNG, the generation count, the stop
 * threshold, stats tracking and the file names (deac.cpp:218-229). */
typedef struct deac_problem {
    size_t number_of_timeslices;          /* T >= 2 */
    const double *imaginary_time;         /* T values */
    const double *isf;                    /* T values, F(tau) */
    const double *isf_error;              /* T values, > 0 */
    size_t genome_size;                   /* G >= 2 */
    const double *frequency;              /* G values, finite, >= 0, sorted */
    double temperature;                   /* > 0 unless zero_temperature */
    int model;                            /* deac_model */
    int spectra_type;                     /* deac_spectra_type */
    int zero_temperature;                 /* reference ZeroT build: kernel e^{-tau w}, one sign */
    int allow_negative_spectral_weight;   /* reference ALLOW_NEGATIVE_SPECTRAL_WEIGHT */
    int renormalize_each_generation;      /* reference SINGLE_PARTICLE_FERMIONIC_SPECTRAL_FUNCTION */
    size_t population_size;               /* P >= 4 */
    int normalize;                        /* --normalize */
    double first_moment;                  /* --first_moment, < 0 disables */
    double third_moment;                  /* --third_moment, < 0 disables */
    double third_moment_error;            /* > 0 when third_moment is used */
    double crossover_probability;         /* -r, initial CR */
    double self_adapting_crossover_probability;           /* -u, tau_CR */
    double differential_weight;                           /* -F, initial F */
    double self_adapting_differential_weight_probability; /* -v, tau_F */
    int use_negative_first_moment;        /* --use_negative_first_moment: used with DEAC_BDSF above zero
                                             temperature (deac.cpp:230-236, 868-920), ignored otherwise */
} deac_problem;

/* Precision of the K.S contraction of the model ISF (deac.cpp:1385-1395):
 *   DEAC_PRECISION_FP64    exact: bit for bit the reference's FMA chains;
 *   DEAC_PRECISION_FP32    K scaled per tau row by a power of two, K and S in
 *                          FP32 on v_mfma_f32_16x16x4f32, FP32 accumulation;
 *   DEAC_PRECISION_BF16X3  K and S split into bfloat16 pairs, three products
 *                          per element on v_mfma_f32_16x16x32_bf16, FP32
 *                          accumulation. */
typedef enum deac_precision {
    DEAC_PRECISION_FP64 = 0,
    DEAC_PRECISION_FP32 = 1,
    DEAC_PRECISION_BF16X3 = 2
} deac_precision;

typedef struct deac_run_options {
    size_t number_of_generations;         /* N >= 2; at most N-1 evolution steps */
    double stop_minimum_fitness;          /* stop when the best fitness <= this */
    int track_stats;                      /* fill the three stats arrays */
    /* ks_precision, a deac_precision.  FP64 (0) is exact everywhere.
     * deac_run: FP32 and BF16X3 run the split generation loop, one kernel
     * per step, with the model ISF from the mixed-precision contraction.
     * deac_batch_run: FP32 and BF16X3 run the batch kernel in the delta form:
     * each member's model ISF is kept in FP64 and a trial's is that plus
     * K.(trial - parent) at the chosen precision; the population is
     * re-evaluated exactly in FP64 before the run, every 1024 generations
     * and at the end, so the reported fitness is the exact FP64 chi^2 of the
     * reported genome.  Either way, trial vectors, chi^2, selection,
     * statistics and constraints stay FP64 in the reference's order, and the
     * random draws are the reference's.
     * split_loop nonzero (deac_run only): FP64 runs through the split loop
     * as well, with the exact contraction (results identical to the fused
     * kernel). */
    int ks_precision;
    int split_loop;
} deac_run_options;

/* Outputs of one reconstruction.  All arrays are caller-allocated. */
typedef struct deac_result {
    size_t generation;           /* as printed by the reference (deac.cpp:1761) */
    double minimum_fitness;      /* best final fitness (deac.cpp:1592) */
    size_t best_index;           /* member holding it (first on ties) */
    double *dsf;                 /* deac_output_length() values */
    double *frequency;           /* deac_output_length() values, mirrored grid */
    double *stats_mean;          /* N values when track_stats (only generation+1 written), or NULL */
    double *stats_minimum;       /* per-generation best fitness history, or NULL */
    double *stats_squared_mean;  /* or NULL */
    double *final_fitness;       /* P values, or NULL */
} deac_result;

typedef struct deac_handle deac_handle;

/* Validates the problem exactly as the reference's main() does
 * (deac.cpp:2060-2160), builds the kernel matrix K(tau,w), the normalization
 * and moment vectors on the host in FP64 (deac.cpp:303-827) and uploads them
 * to the current HIP device.  On DEAC_ERR_INVALID_ARGUMENT, *message (if not
 * NULL) points to the reference's error text; it stays valid until the next
 * call on this thread. */
deac_status deac_create(const deac_problem *problem, deac_handle **out, const char **message);

/* Runs one reconstruction: initial population from rng_state (in/out, the
 * 4-word xoshiro256+ state of rng.hpp:21-23; the all-zero state, a fixed
 * point of the generator, is DEAC_ERR_INVALID_ARGUMENT), optional
 * normalization, then at
 * most N-1 generations of the fused GPU pipeline.  Plain FP64: reproduces the
 * reference CPU build's trajectory.  Replaces deac.cpp:430-1685. */
deac_status deac_run(deac_handle *handle, uint64_t rng_state[4],
                     const deac_run_options *options, deac_result *result);

void deac_destroy(deac_handle *handle);

/* Length of the dsf/frequency outputs: 2G-1, or G at zero temperature and
 * for DEAC_BDSF (deac.cpp:1583-1606). */
size_t deac_output_length(const deac_problem *problem);

/* xoshiro256+ state for a --seed value: seeded with 1407513600 + seed
 * (deac.cpp:2089-2091, rng.hpp:46-59). */
void deac_seed_state(uint64_t seed, uint64_t state[4]);

/* Advances a xoshiro256+ state by n draws in O(log n) GF(2) steps
 * (generalises xoshiro256p_jump, rng.hpp:64-87, to any distance). */
void deac_rng_advance(uint64_t state[4], uint64_t n);

/* The default frequency grid, frequency[i] = i*(omega_max/(G-1))
 * (deac.cpp:2120-2125). */
deac_status deac_frequency_grid(size_t genome_size, double omega_max, double *frequency);

const char *deac_status_string(deac_status status);

/* Text of the last error reported on the calling thread by any entry point
 * (the reference prints such messages before exit(), deac.cpp:2075-2160);
 * empty if none.  Valid until the next call on this thread. */
const char *deac_last_error_message(void);

/* Device timing of the last deac_run on this handle, from HIP events on the
 * handle's stream (no counterpart in the reference, which has no timers). */
typedef struct deac_run_profile {
    double loop_ms;              /* whole generation loop */
    double walk_ms;              /* RNG walk kernels */
    double mask_ms;              /* crossover-mask kernels */
    double evolve_ms;            /* fused generation kernels */
    size_t generations;          /* generations evolved */
    size_t evolve_launches;      /* launches of the fused generation kernel */
    double eval_ms;              /* batch, mixed K.S: exact FP64 re-evaluation launches */
    size_t eval_launches;
    size_t graph_launches;       /* deac_run's split loop: launches of captured HIP graphs */
    size_t evolve_workgroups;    /* batch: workgroups of the evolution launches, summed (instances of one
                                    problem with small populations may share one) */
} deac_run_profile;
deac_status deac_last_run_profile(const deac_handle *handle, deac_run_profile *profile);

/* ---------------------------------------------------------------- batches
 * Many independent reconstructions in one persistent launch: the reference's
 * multi-seed workflow (tools/generate_commands.sh:18-32, one deac.e process
 * per seed, averaged by tools/plot_simple_average_deac_results.py:8-28) as
 * one call.  At ks_precision FP64, each instance i computes exactly what
 * deac_run would compute for problems[i] from rng_states[4i..4i+3] (the
 * reference's order and random stream); at FP32 or BF16X3 it runs the same
 * random stream with the delta-form mixed contraction (deac_run_options).
 * Each instance has its own generation count, stop test and precision; an
 * instance that stops leaves the launch.  Identical problems share one
 * device copy of the kernel matrix.  Population size at most 64 (33 to 64:
 * T <= 132); moment constraints, per-generation renormalization and
 * DEAC_BDSF (T <= 128) run at every precision (mixed: the trials' moment
 * terms and zeroth moments summed in FP64 in the kernel's own order, every
 * reported fitness from the FP64 evaluation). */
typedef struct deac_batch deac_batch;

deac_status deac_batch_create(const deac_problem *problems, size_t count, deac_batch **out, const char **message);

/* rng_states: 4*count words, in/out as deac_run's rng_state (no instance's
 * state may be all zero); options and results: count entries each, as
 * deac_run's. */
deac_status deac_batch_run(deac_batch *batch, uint64_t *rng_states, const deac_run_options *options,
                           deac_result *results);

/* Continues every instance from where the last deac_batch_run or
 * deac_batch_continue on this batch left it (population, fitness,
 * self-adapted CR and F and random state stay on the device) for
 * options[i].number_of_generations more generations, counted like a run's:
 * N - 1 more evolution steps, stats histories of this segment (entry 0 is
 * the last entry of the previous segment), generation relative to the
 * segment.  At FP64 a run of N1 generations continued by N2 is bit-identical
 * to one run of N1 + N2 - 1.  rng_states is output only.  No counterpart in
 * the reference (whose runs cannot be resumed); used to time generations
 * past a warm-up within one run. */
deac_status deac_batch_continue(deac_batch *batch, uint64_t *rng_states, const deac_run_options *options,
                                deac_result *results);

/* Launch schedule of a batch: how the instances of different kernel shapes
 * share the device (no effect on any result).  deac_batch_tune picks one by
 * timing the candidates on the device; it is kept by the batch and, for
 * batches of the same shapes (T, G, P, signs, precision of every instance),
 * by the process. */
typedef struct deac_batch_schedule {
    int group_streams;         /* 1..3: kernel-shape groups launched side by side on this many streams */
    int group_order;           /* 0: most expensive group first, 1: least expensive first, 2: by kernel shape */
    int instance_order;        /* 0: longest instances first, 1: round robin over distinct problems */
    double ms_per_generation;  /* deac_batch_tune: the chosen schedule's time per generation of the batch */
    int pack;                  /* FP64 instances of one problem (populations of at most 8) stacked in one
                                  workgroup, up to 4 and as many as fit its 16 member rows: 0 automatic
                                  (by the active instances per compute unit: none while they fill under
                                  3/4 of the device), 1 never, 2..4 at most this many */
} deac_batch_schedule;

/* Times every candidate schedule on `generations` generations of this batch
 * (after one run of that length from rng_states with the options' problems
 * and precisions, each candidate continues it twice; outputs discarded,
 * rng_states not modified), keeps the fastest for later runs and returns it.
 * No counterpart in the reference. */
deac_status deac_batch_tune(deac_batch *batch, const uint64_t *rng_states, const deac_run_options *options,
                            size_t generations, deac_batch_schedule *chosen);
deac_status deac_batch_set_schedule(deac_batch *batch, const deac_batch_schedule *schedule);

/* Profiling builds only (make ROLE_TIMERS=1; DEAC_ERR_UNSUPPORTED otherwise):
 * per wavefront w of the first workgroup of every batch launch of the
 * kernel family of `precision` (FP64, or the mixed FP32/BF16X3 kernel),
 * summed since the last reset: cycles[3w] shader-clock cycles waiting in the
 * chunk pipeline's barriers, cycles[3w+1] cycles of the launches,
 * cycles[3w+2] generations evolved.  Wavefronts 0-3 chains, 4-11 workers,
 * 12 walk, 13-15 crossover masks.  No counterpart in the reference. */
deac_status deac_batch_role_cycles(int precision, unsigned long long cycles[48], int reset);
deac_status deac_batch_get_schedule(const deac_batch *batch, deac_batch_schedule *schedule);

/* deac_batch_run without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Results
 * are those of deac_batch_run.  The statistics histories are staged in a
 * device array of the library, 4194304 doubles (32 MiB) per GPU, held until
 * the process ends: calls whose tracked histories (3 N doubles per instance
 * with track_stats) exceed it return DEAC_ERR_INVALID_ARGUMENT before any
 * launch.  Mixed K.S precisions need their kernel tiles, which the first
 * deac_batch_run at that precision builds (DEAC_ERR_UNSUPPORTED otherwise).
 * Calls on one GPU share the array: calls that overlap on one GPU race. */
deac_status deac_batch_run_noalloc(deac_batch *batch, uint64_t *rng_states, const deac_run_options *options,
                                   deac_result *results);

void deac_batch_destroy(deac_batch *batch);

size_t deac_batch_size(const deac_batch *batch);

/* Device timing of the last deac_batch_run: loop_ms covers every launch
 * (initial evaluation included), generations is summed over instances. */
deac_status deac_batch_last_profile(const deac_batch *batch, deac_run_profile *profile);

/* ---- K.S precision sweep ------------------------------------------------
 * The model ISF E = K.S of a stack of genomes, the contraction of the
 * reference's fitness evaluation (deac.cpp:1385-1395, matrix_multiply_LxM_by_MxN
 * at deac.cpp:110-118), at a chosen deac_precision of K.S (above).
 * It has no counterpart among the reference's functions: it exists to report
 * |dchi^2|/chi^2 of the mixed-precision forms against FP64.
 *
 * genomes: count rows of deac_genome_length(problem) doubles (positive
 * frequencies then negative, the population layout of deac.cpp:430-458;
 * positive only at zero temperature and for DEAC_BDSF).  FP64 is the exact
 * FMA chain, also for DEAC_BDSF (whose reference build sums its trials, not
 * its initial population, as a fold).
 * Outputs (each may be NULL): model_isf, count x T; chi2, count reduced
 * chi^2 values (deac.cpp:120-126, 1540-1542; FP64 in the reference's order,
 * no moment terms); kernel_ms, device time of the contraction, from the
 * start of its kernel to the end of the last one (for FP32 and BF16X3, the
 * FP32 sum over genome slices), host launch latency excluded. */
size_t deac_genome_length(const deac_problem *problem);

deac_status deac_model_isf(const deac_problem *problem, const double *genomes, size_t count, int precision,
                           double *model_isf, double *chi2, double *kernel_ms, const char **message);

#ifdef __cplusplus
}
#endif
#endif
