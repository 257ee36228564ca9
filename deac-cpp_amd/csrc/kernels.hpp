// Device-side interface of the DEAC engine: argument blocks and launchers of
// the gfx950 kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>

namespace deac {

// Exact FP64 model-ISF contraction: one compute wavefront holds 4 members x
// 16 timeslices, one sequential FMA chain per output on
// v_mfma_f64_4x4x4_4b_f64; the kernel-matrix and genome chunks of kChunk
// positions come from LDS, kStages buffers, filled by the helper waves from a
// ring of registers a few chunks ahead.
constexpr int kChainRows = 4;
constexpr int kChainTau = 16;
constexpr int kChunk = 128;
constexpr int kStages = 2;

// Random choices of one member in one generation, produced by walk_kernel.
struct MemberCtl {
    uint64_t mask_state[4];  // xoshiro256+ state at the member's first crossover-mask draw
    uint32_t mutant[3];      // donor indices r1, r2, r3 (deac.cpp:200-216)
    uint32_t flags;          // bit 2s: CR of sign s redrawn; bit 2s+1: F of sign s redrawn
    double cr_value[2];      // redrawn CR per sign
    double f_value[2];       // redrawn F per sign
};

struct WalkArgs {
    uint64_t *state;             // in/out, 4 words
    uint64_t *gen_start;         // [ngen][4] state at the top of each generation
    MemberCtl *ctl;              // [ngen][P]
    const uint64_t *skip_rows;   // rows of T^(nsign*G), 256 x 4 words
    int ngen, P, nsign;
    uint64_t mod_magic;          // floor((2^64-1)/P)
    double tau_cr, tau_f;
};

struct MaskArgs {
    const MemberCtl *ctl;        // [ngen][P]
    const uint64_t *lane_tab;    // [64][64][16][4] nibble tables of T^(lane*chunk)
    uint64_t *raw;               // [ngen][P][Kc] 53-bit draws (x >> 11)
    int ngen, P, Kc, chunk;
};

struct EvoArgs {
    int P, T, G, nsign, Kc;
    int stride;                  // row stride of K and of the populations: Kc padded to kChunk
    int MG, TB;                  // member groups of kChainRows, tau blocks of kChainTau
    int use_first, use_third, renorm, allow_neg, init_mode, track_stats;
    double stop, zeroth, first_moment, third_moment, third_moment_error;
    int use_nfm;                 // -1st-moment term of the fitness (bdsf): weights nfm_dt [T]
    double nfm_value, nfm_error;
    const double *nfm_dt;
    int fold_ks;                 // the trials' model ISF as a rounded-product fold (bdsf; split loop only)
    const double *kmat;          // [TB*16][stride], zero padded
    const double *nterm, *fterm, *tterm;   // [Kc] or null
    const double *obs, *sig;     // [T]
    double *pop0, *pop1;         // [MG*4][stride] each, member-major rows, zero padded
    int *cur;                    // [P] buffer holding each member's current genome
    double *fitness;             // [P]
    double *cr_old, *f_old;      // [P][nsign]
    double *emod;                // [P][T]
    double *mom;                 // [P][4]: normalization, first, third moment
    const MemberCtl *ctl;        // [ngen][P]
    const uint64_t *raw;         // [ngen][P][Kc]
    double *stats;               // [3][stats_len]: mean, minimum, squared mean
    size_t stats_len;
    unsigned *bar;               // grid barrier counter, zero at launch
    int *status;                 // [0] stopped, [1] generation index within the batch, [2] barrier timeout
    long gen0;                   // generation index of the batch's first generation
    int ngen;
    long *iter;                  // split loop under graph capture: [0] gen0, [1] generation within the batch
                                 // (read in place of gen0 and the kernels' g), or null
};

hipError_t launch_walk(const WalkArgs &a, hipStream_t s);
hipError_t launch_mask(const MaskArgs &a, hipStream_t s);
hipError_t launch_evolve(const EvoArgs &a, int nwg, hipStream_t s);
hipError_t launch_zero_u32(unsigned *p, size_t n, hipStream_t s);
// Workgroups of the evolution kernel that are co-resident on the device: one
// per CU.
int evolve_max_workgroups(int device);

}  // namespace deac
