// Batched DEAC kernel: many independent reconstructions in one launch, one
// persistent workgroup per reconstruction, plain FP64 arithmetic in the
// reference's order (bit-identical to deac_run for every instance).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace deac {

constexpr int kBatchWaves = 16;             // 4 chain, 8 worker, 1 walk, 3 mask
constexpr int kBatchThreads = kBatchWaves * 64;
constexpr int kBatchMaskLanes = 192;        // 3 mask wavefronts
constexpr int kBatchChunkDoubles = 2048;    // members (padded) x genome positions per chunk
constexpr int kBatchMaxP = 32;
constexpr int kBatchRegionDoubles = 6 * kBatchChunkDoubles;  // 4 population + 2 trial buffers
constexpr int kBatchMaxRem = 4;             // kernel rows past the last full tile run on 4x4x4 MFMAs

// Everything one workgroup needs about its reconstruction.  Model arrays may
// be shared between instances; state arrays are the instance's own.
struct BatchDesc {
    // model, laid out for the kernel by the host
    const double *ktile;       // [ntt][Lpad/8][64 lanes][2]: K tiles in MFMA A order
    const double *krem;        // [Lpad/8][4 nrem][2]: rows past the last full 16-row tile, in
                               // v_mfma_f64_4x4x4_4b A order (row r + nrem * position k)
    const double *obs, *sig;   // [T] ISF and its error
    const uint64_t *walk_rows; // [256][4] rows of T^L, T the xoshiro transition
    const uint64_t *sub_tab;   // [S-1][64][16][4] nibble tables of T^(k*Lsub), k = 1..S-1
    // state
    double *pop;               // [P][Lpad] genomes, positive sign then negative sign
    double *fitness;           // [P]
    double *cr_old, *f_old;    // [P][2] self-adapted CR and F per sign
    uint32_t *masks;           // [3][Ppad][2][Lw] crossover bits, two candidates per generation
    uint64_t *rng;             // [4] in: state at the top of gen0; out: at the first generation not evolved
    double *stats;             // [3][stats_len] mean, minimum, squared mean per generation, or null
    long *status;              // [0] stopped, [1] generation reached
    long stats_len;
    long gen0, gen_end;        // evolve generations gen0 .. gen_end-1 unless the stop test fires
    double stop, tau_cr, tau_f;
    uint64_t mod_magic;        // floor((2^64-1)/P)
    int init;                  // evaluate the initial population first (gen0 == 0)
    int T, G, nsign, L, Lpad, Lw, nchunk, ntt, nrem, P, S, Lsub, allow_neg;
};

// NT: 16-row kernel tiles per chain wavefront (1, 2 or 4); NM: 16-member tiles (1 or 2);
// rem: 1 when the kernel has 1..4 rows past its last full tile.
hipError_t launch_batch(const BatchDesc *descs, int count, int nt, int nm, int rem, hipStream_t s);

}  // namespace deac
