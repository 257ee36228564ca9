// Launchers of the K.S precision-sweep kernels (precision_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace deac {

// Operand prefetch depth of the kernels, in k-blocks: the k-block counts
// passed in (Kb, and the slice length) are multiples of these.
constexpr int kSweepDepth = 8;
constexpr int kSweepDepth16 = 4;

// Each launcher takes optional events (null: none) that the dispatch itself
// records when it starts and when it ends, so that they time the kernel
// alone and not the host's launch latency.  The contraction kernels also take
// an optional flag (null: none): when it is nonzero they return at once.

// E[n][t] (FP64, exact chains); operands in 16x16x4 lane order, Kb = genome / 4.
hipError_t launch_ks_fp64(const double *A, const double *B, double *E, int T, int N, int Mt, int Nt, int Kb,
                          hipStream_t s, hipEvent_t start, hipEvent_t stop, const int *halt);
// Partial FP32 tiles [slices][16 Mt][16 Nt]; Kb = genome / 4.
hipError_t launch_ks_fp32(const float *A, const float *B, float *part, int Mt, int Nt, int Kb, int kb_per_slice,
                          int slices, hipStream_t s, hipEvent_t start, hipEvent_t stop, const int *halt);
// Partial FP32 tiles from bfloat16 hi/lo operands in 16x16x32 lane order; Kb = genome / 32.
hipError_t launch_ks_bf16x3(const void *Ah, const void *Al, const void *Bh, const void *Bl, float *part, int Mt,
                            int Nt, int Kb, int kb_per_slice, int slices, hipStream_t s, hipEvent_t start,
                            hipEvent_t stop, const int *halt);
// out[i] = sum over slices of part[slice][i], in slice order, FP32.
hipError_t launch_reduce_slices(const float *part, float *out, int slices, size_t mn, hipStream_t s, hipEvent_t start,
                                hipEvent_t stop);

}  // namespace deac
