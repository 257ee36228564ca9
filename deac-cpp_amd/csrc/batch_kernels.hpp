// Batched DEAC kernel: many independent reconstructions in one launch, one
// persistent workgroup per reconstruction.  Exact FP64: the reference's
// arithmetic and order (bit-identical to deac_run for every instance); mixed
// K.S (FP32, BF16X3): the delta form on an FP64 parent model ISF, chi^2 and
// selection in FP64.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace deac {

constexpr int kBatchWaves = 16;             // 4 chain, 8 worker, 1 walk, 3 mask
constexpr int kBatchThreads = kBatchWaves * 64;
constexpr int kBatchMaskLanes = 192;        // 3 mask wavefronts
constexpr int kBatchChunkDoubles = 2048;    // members (padded) x genome positions per chunk
constexpr int kBatchMaxP = 64;             // members per workgroup (33 to 64: T <= 132)
constexpr int kBatchRegionDoubles = 6 * kBatchChunkDoubles;  // 4 population + 2 trial buffers
constexpr int kBatchMaxRem = 4;             // kernel rows past the last full tile run on 4x4x4 MFMAs
constexpr int kBatchPackMax = 4;            // instances of one model stacked in one workgroup (FP64)

// Everything one workgroup needs about its reconstruction.  Model arrays may
// be shared between instances; state arrays are the instance's own.
struct BatchDesc {
    // model, laid out for the kernel by the host
    const double *ktile;       // [ntt][Lpad/8][64 lanes][2]: K tiles in MFMA A order
    const double *krem;        // [Lpad/8][4 nrem][2]: rows past the last full 16-row tile, in
                               // v_mfma_f64_4x4x4_4b A order (row r + nrem * position k)
    const void *kmix;          // mixed K.S: [ntm][Lpad/16][64][4] floats (FP32) or [ntm][Lpad/32][64][8 hi, 8 lo]
                               // bfloat16 (BF16X3), rows scaled by 2^-ex[t], zero past T and L
    const double *kscale;      // mixed K.S: [T] 2^ex[t]
    const double *obs, *sig;   // [T] ISF and its error
    const double *mterm;       // [3][Lpad] first-moment, third-moment and normalization weights (zero past L),
                               // or null when the model uses none of them
    const double *kfold;       // bdsf fold kernel: [Lpad][fold_rows], entry (k, t) = K row t at position k
                               // (zero past T and L)
    const double *nfm_dt;      // [T] trapezoid weights of the -1st-moment term (use_nfm)
    const uint64_t *walk_rows; // [256][4] rows of T^L, T the xoshiro transition
    const uint64_t *sub_tab;   // [S-1][64][16][4] nibble tables of T^(k*Lsub), k = 1..S-1
    // state of Q instances of one model stacked along the members: member m
    // of the workgroup is member m % Pi of instance m / Pi (Q = 1 unless the
    // host packs small populations, FP64 kernel only)
    double *pop;               // genomes, [Pi][Lpad] per instance (positive sign then negative sign);
                               // instance q's row 0 lies pop_row0[q] bytes past pop
    uint32_t pop_row0[kBatchPackMax];
    double *fitness[kBatchPackMax];                          // [Pi] per instance
    double *cr_old[kBatchPackMax], *f_old[kBatchPackMax];    // [Pi][2] self-adapted CR and F per sign
    uint32_t *masks;           // [3][Ppad][2][Lw] crossover bits, two candidates per generation
    uint64_t *rng[kBatchPackMax];  // [4] in: state at the top of gen0; out: at the first generation not evolved
    double *epar;              // [T][Ppad] model ISF of each member's genome, FP64 (mixed K.S; written by
                               // evaluation launches), or null
    double *stats[kBatchPackMax];  // [3][stats_len] mean, minimum, squared mean per generation, or null
    long *status[kBatchPackMax];   // [0] stopped, [1] generation reached
    int *accp[kBatchPackMax];      // [Pi] acceptances of the last generation a launch evolved (written at
                                   // its end; read at the start of a resume launch)
    long stats_len[kBatchPackMax], stats_base[kBatchPackMax];  // entry for generation g: stats[g - stats_base]
    long gen0, gen_end;        // evolve generations gen0 .. gen_end-1 unless the stop test fires
    double stop[kBatchPackMax];    // stop_minimum_fitness per instance
    double tau_cr, tau_f;
    double first_moment, third_moment, third_moment_error;  // used when use_first / use_third
    double zeroth;             // zeroth moment the trials are renormalized to (renorm)
    double nfm_value, nfm_error;  // -1st-moment term (use_nfm): target and error
    uint64_t mod_magic;        // floor((2^64-1)/Pi)
    int init;                  // evaluate the initial population first (gen0 == 0; FP64 kernel only)
    int resume;                // the previous launch of this instance (Q = 1) ended at gen0 and left the
                               // masks of gen0 and the acceptances of gen0 - 1: walk gen0 and gen0 + 1 and
                               // skip the mask pass
    int mode;                  // 0: evolve; 1: evaluation only (FP64 kernel, init set): fitness and epar of the
                               // stored population, nothing else written
    int T, G, nsign, L, Lpad, Lw, nchunk, ntt, nrem, ntm, S, Lsub, allow_neg;
    int P;                     // members in the workgroup, Q Pi
    int Q, Pi;                 // instances and members per instance
    int use_first, use_third;  // moment constraints (deac.cpp:1420-1477, 1550-1555); mixed: the REM = 1 kernels
    int renorm;                // renormalize each generation's trials (deac.cpp:1342-1364); mixed: REM = 1
    int use_nfm;               // -1st-moment term in the fitness (bdsf above zero temperature, deac.cpp:961-965)
    int fold_rows;             // bdsf fold kernel: row stride of kfold, R (+ 4 with extra timeslices)
};

// Exact FP64 kernel.  NT: 16-row kernel tiles per chain wavefront (1, 2 or 4); NM: 16-member
// tiles (1, 2, or 4 with NT <= 2); rem: 1 when the kernel has 1..4 rows past its last full tile.
hipError_t launch_batch(const BatchDesc *descs, int count, int nt, int nm, int rem, hipStream_t s);

// The detailed-balance (bdsf) kernel: the trials' model ISF as a fold of rounded products
// on the FP64 VALU (the reference build's loop contraction), everything else as the FP64
// kernel; never evaluates a population from scratch (the FP64 kernel's evaluation launch
// gives the initial fitness).  shape = 2 RL + XR: a lane folds RL consecutive timeslices of
// one of 64 / nm row blocks (R = 64 RL / nm timeslices) for 4 members; XR = 1 folds timeslices
// R .. T - 1 (at most 4), one per chain wave.  RL: 1 or 2 (nm 1), 1, 2 or 4 (nm 2), 1, 2, 4
// or 8 (nm 4); nm as above.
hipError_t launch_batch_fold(const BatchDesc *descs, int count, int shape, int nm, hipStream_t s);

// Device scratch of deac_batch_run_noalloc: a file-scope array of kBatchScratchDoubles
// doubles per GPU (32 MiB), its address looked up once per device with that device
// current (hipGetSymbolAddress).  Calls that overlap on one GPU share it.
constexpr size_t kBatchScratchDoubles = size_t(1) << 22;
constexpr int kBatchMaxDevices = 64;
hipError_t batch_scratch(int device, double **ptr);

// Profiling builds (DEAC_ROLE_TIMERS): per wavefront of workgroup 0, summed over launches,
// [wave][0] barrier-wait cycles, [1] launch cycles, [2] generations; hipErrorNotSupported
// otherwise.  One counter set per kernel family.
hipError_t batch_role_cycles_fp64(unsigned long long *out, int reset);
hipError_t batch_role_cycles_mixed(unsigned long long *out, int reset);

// Mixed K.S kernel (precision DEAC_PRECISION_FP32 or DEAC_PRECISION_BF16X3).  nq: (kernel
// tile, member tile) pairs per chain wavefront, one of kBatchMixedNq; nm as above; moments:
// 1 for problems with moment constraints (the trials' moments summed by the workers).
constexpr int kBatchMixedNq[] = {1, 2, 3, 5, 9};
int batch_mixed_nq(int ntm, int nm);  // smallest instantiated nq for ntm kernel tiles (0: none)
hipError_t launch_batch_mixed(const BatchDesc *descs, int count, int precision, int nq, int nm, int moments,
                              hipStream_t s);

}  // namespace deac
