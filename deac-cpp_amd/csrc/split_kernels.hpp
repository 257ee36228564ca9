// Launchers of the split generation loop (split_kernels.hip): the steps of
// one generation of the fused evolution kernel as separate kernels, with the
// model ISF from the K.S contraction kernels (precision_kernels.hpp).  Every
// kernel returns at once when status[0] (a stop) is set.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

#include "kernels.hpp"

namespace deac {

// Top of generation g of the batch (best fitness, statistics, stop test) and
// its trial vectors into the spare buffers, in one launch.
hipError_t launch_split_trial(const EvoArgs &a, int g, hipStream_t s);
// Per-generation renormalization of the trials (call only when a.renorm).
hipError_t launch_split_renorm(const EvoArgs &a, hipStream_t s);
// Genome operand of the contraction from the trials (init = 0) or the
// current genomes (init = 1): Nt tiles of 16 members, Kb k-blocks.
hipError_t launch_split_pack_fp64(const EvoArgs &a, int init, double *B, int Nt, int Kb, hipStream_t s);
hipError_t launch_split_pack_fp32(const EvoArgs &a, int init, float *B, int Nt, int Kb, hipStream_t s);
hipError_t launch_split_pack_bf16x3(const EvoArgs &a, int init, void *Bh, void *Bl, int Nt, int Kb, hipStream_t s);
// Model ISF a.emod from FP32 slices part[slices][Mpad][Npad] (mn = Mpad*Npad)
// and the per-row exponents ex[T].
hipError_t launch_split_model(const EvoArgs &a, const float *part, int slices, int Npad, size_t mn, const int *ex,
                              hipStream_t s);
// First / third moment rows (call only when a moment term is used).
// bdsf: the trials' model ISF as a fold of rounded products (FP64).
hipError_t launch_split_fold(const EvoArgs &a, hipStream_t s);
hipError_t launch_split_moments(const EvoArgs &a, int init, hipStream_t s);
// Fitness from a.emod and a.mom; greedy selection (init = 0) or the initial
// fitness (init = 1).
hipError_t launch_split_select(const EvoArgs &a, int g, int init, hipStream_t s);
// Generation counter of a batch run through captured graphs (EvoArgs::iter):
// set = 1: iter = {gen0, 0}; set = 0: the next generation of the batch.
hipError_t launch_split_iter(long *iter, long gen0, int set, hipStream_t s);

}  // namespace deac
