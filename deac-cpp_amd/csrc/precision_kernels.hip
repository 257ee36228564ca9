// Model ISF E = K.S of a stack of genomes at three K.S precisions, on the
// gfx950 matrix cores (the precision sweep of the K.S contraction):
//
//   FP64    v_mfma_f64_16x16x4f64, one accumulator per output tile over the
//           whole genome in ascending position order: each output element is
//           the reference's sequential FMA chain (deac.cpp:110-118), bit for
//           bit, as in the batch kernel;
//   FP32    K (scaled per tau row by a power of two) and S rounded to FP32,
//           v_mfma_f32_16x16x4f32 with FP32 accumulation over slices of the
//           genome, slices summed in FP32;
//   BF16x3  K = Kh + Kl and S = Sh + Sl split into bfloat16 pairs,
//           Kl.Sh + Kh.Sl + Kh.Sh per 32 positions on v_mfma_f32_16x16x32_bf16,
//           FP32 accumulation as above.
//
// Operands arrive in MFMA lane order (prepared on the host), one coalesced
// load per wave and k-block:
//   16x16x4  (FP64, FP32): A[mt][kb][lane] = K[16 mt + lane%16][4 kb + lane/16],
//                          B[nt][kb][lane] = S[16 nt + lane%16][4 kb + lane/16];
//   16x16x32 (BF16):       A[mt][kb][lane][e] = K[16 mt + lane%16][32 kb + 8 (lane/16) + e],
//                          B likewise with S.
// Outputs: FP64 E[n][t] directly; FP32/BF16 partial tiles [slice][Mpad][Npad]
// reduced in slice order by reduce_slices.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "precision_kernels.hpp"

namespace deac {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kDepth = kSweepDepth;      // k-blocks in flight, 16x16x4 operands
constexpr int kDepth16 = kSweepDepth16;  // k-blocks in flight, 16x16x32 operands

// One wave per (tau tile, genome tile); the four waves of a workgroup take
// four genome tiles.
__global__ __launch_bounds__(256) void ks_fp64_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                                      double *__restrict__ E, int T, int N, int Mt, int Nt, int Kb,
                                                      const int *halt)
{
    if (halt && *halt) return;
    const int lane = threadIdx.x & 63;
    const int mt = blockIdx.x, nt = blockIdx.y * 4 + static_cast<int>(threadIdx.x >> 6);
    if (mt >= Mt || nt >= Nt) return;
    const double *a = A + static_cast<size_t>(mt) * Kb * 64 + lane;
    const double *b = B + static_cast<size_t>(nt) * Kb * 64 + lane;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    // operands of the next kDepth k-blocks in flight (Kb is a multiple of
    // kDepth; the refill past the end re-reads the last block, unused)
    double ra[kDepth], rb[kDepth];
#pragma unroll
    for (int i = 0; i < kDepth; i++) {
        ra[i] = a[static_cast<size_t>(i) * 64];
        rb[i] = b[static_cast<size_t>(i) * 64];
    }
    for (int kb = 0; kb < Kb; kb += kDepth) {
#pragma unroll
        for (int i = 0; i < kDepth; i++) {
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(ra[i], rb[i], acc, 0, 0, 0);
            const size_t nk = static_cast<size_t>(min(kb + i + kDepth, Kb - 1)) * 64;
            ra[i] = a[nk];
            rb[i] = b[nk];
        }
    }
    // lane l holds rows 4v + l/16 of the tile, column l%16
    const int n = nt * 16 + (lane & 15);
    for (int v = 0; v < 4; v++) {
        const int t = mt * 16 + 4 * v + (lane >> 4);
        if (t < T && n < N) E[static_cast<size_t>(n) * T + t] = acc[v];
    }
}

__device__ __forceinline__ void store_partial(float *part, int slice, int Mpad, int Npad, int mt, int nt, int lane,
                                              f4 acc)
{
    // lane l holds rows 4 (l/16) + v of the tile, column l%16
    float *p = part + (static_cast<size_t>(slice) * Mpad + mt * 16 + 4 * (lane >> 4)) * Npad + nt * 16 + (lane & 15);
    for (int v = 0; v < 4; v++) p[static_cast<size_t>(v) * Npad] = acc[v];
}

// grid (Mt, ceil(Nt/4), slices): one output tile per wave over one slice
__global__ __launch_bounds__(256) void ks_fp32_kernel(const float *__restrict__ A, const float *__restrict__ B,
                                                      float *__restrict__ part, int Mt, int Nt, int Kb, int kb_per_slice,
                                                      const int *halt)
{
    if (halt && *halt) return;
    const int lane = threadIdx.x & 63;
    const int mt = blockIdx.x, nt = blockIdx.y * 4 + static_cast<int>(threadIdx.x >> 6), slice = blockIdx.z;
    if (nt >= Nt) return;
    const int kb0 = slice * kb_per_slice, kb1 = min(kb0 + kb_per_slice, Kb);
    const float *a = A + static_cast<size_t>(mt) * Kb * 64 + lane;
    const float *b = B + static_cast<size_t>(nt) * Kb * 64 + lane;
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    float ra[kDepth], rb[kDepth];
#pragma unroll
    for (int i = 0; i < kDepth; i++) {
        ra[i] = a[static_cast<size_t>(kb0 + i) * 64];
        rb[i] = b[static_cast<size_t>(kb0 + i) * 64];
    }
    for (int kb = kb0; kb < kb1; kb += kDepth) {
#pragma unroll
        for (int i = 0; i < kDepth; i++) {
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(ra[i], rb[i], acc, 0, 0, 0);
            const size_t nk = static_cast<size_t>(min(kb + i + kDepth, kb1 - 1)) * 64;
            ra[i] = a[nk];
            rb[i] = b[nk];
        }
    }
    store_partial(part, slice, Mt * 16, Nt * 16, mt, nt, lane, acc);
}

__global__ __launch_bounds__(256) void ks_bf16x3_kernel(const bf8 *__restrict__ Ah, const bf8 *__restrict__ Al,
                                                        const bf8 *__restrict__ Bh, const bf8 *__restrict__ Bl,
                                                        float *__restrict__ part, int Mt, int Nt, int Kb,
                                                        int kb_per_slice, const int *halt)
{
    if (halt && *halt) return;
    const int lane = threadIdx.x & 63;
    const int mt = blockIdx.x, nt = blockIdx.y * 4 + static_cast<int>(threadIdx.x >> 6), slice = blockIdx.z;
    if (nt >= Nt) return;
    const int kb0 = slice * kb_per_slice, kb1 = min(kb0 + kb_per_slice, Kb);
    const size_t aoff = static_cast<size_t>(mt) * Kb * 64 + lane, boff = static_cast<size_t>(nt) * Kb * 64 + lane;
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    bf8 rah[kDepth16], ral[kDepth16], rbh[kDepth16], rbl[kDepth16];
#pragma unroll
    for (int i = 0; i < kDepth16; i++) {
        const size_t k = static_cast<size_t>(kb0 + i) * 64;
        rah[i] = Ah[aoff + k];
        ral[i] = Al[aoff + k];
        rbh[i] = Bh[boff + k];
        rbl[i] = Bl[boff + k];
    }
    for (int kb = kb0; kb < kb1; kb += kDepth16) {
#pragma unroll
        for (int i = 0; i < kDepth16; i++) {
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ral[i], rbh[i], acc, 0, 0, 0);  // small terms first
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rah[i], rbl[i], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(rah[i], rbh[i], acc, 0, 0, 0);
            const size_t nk = static_cast<size_t>(min(kb + i + kDepth16, kb1 - 1)) * 64;
            rah[i] = Ah[aoff + nk];
            ral[i] = Al[aoff + nk];
            rbh[i] = Bh[boff + nk];
            rbl[i] = Bl[boff + nk];
        }
    }
    store_partial(part, slice, Mt * 16, Nt * 16, mt, nt, lane, acc);
}

// E32[m][n] = sum over slices (in order, FP32) of part[slice][m][n]
__global__ __launch_bounds__(256) void reduce_slices_kernel(const float *__restrict__ part, float *__restrict__ out,
                                                            int slices, size_t mn)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= mn) return;
    float s = 0.0f;
#pragma unroll 8
    for (int k = 0; k < slices; k++) s += part[static_cast<size_t>(k) * mn + i];
    out[i] = s;
}

}  // namespace

hipError_t launch_ks_fp64(const double *A, const double *B, double *E, int T, int N, int Mt, int Nt, int Kb,
                          hipStream_t s, hipEvent_t start, hipEvent_t stop, const int *halt)
{
    hipExtLaunchKernelGGL(ks_fp64_kernel, dim3(Mt, (Nt + 3) / 4), dim3(256), 0, s, start, stop, 0, A, B, E, T, N, Mt,
                          Nt, Kb, halt);
    return hipGetLastError();
}

hipError_t launch_ks_fp32(const float *A, const float *B, float *part, int Mt, int Nt, int Kb, int kb_per_slice,
                          int slices, hipStream_t s, hipEvent_t start, hipEvent_t stop, const int *halt)
{
    hipExtLaunchKernelGGL(ks_fp32_kernel, dim3(Mt, (Nt + 3) / 4, slices), dim3(256), 0, s, start, stop, 0, A, B, part,
                          Mt, Nt, Kb, kb_per_slice, halt);
    return hipGetLastError();
}

hipError_t launch_ks_bf16x3(const void *Ah, const void *Al, const void *Bh, const void *Bl, float *part, int Mt,
                            int Nt, int Kb, int kb_per_slice, int slices, hipStream_t s, hipEvent_t start,
                            hipEvent_t stop, const int *halt)
{
    hipExtLaunchKernelGGL(ks_bf16x3_kernel, dim3(Mt, (Nt + 3) / 4, slices), dim3(256), 0, s, start, stop, 0,
                          static_cast<const bf8 *>(Ah), static_cast<const bf8 *>(Al), static_cast<const bf8 *>(Bh),
                          static_cast<const bf8 *>(Bl), part, Mt, Nt, Kb, kb_per_slice, halt);
    return hipGetLastError();
}

hipError_t launch_reduce_slices(const float *part, float *out, int slices, size_t mn, hipStream_t s, hipEvent_t start,
                                hipEvent_t stop)
{
    const unsigned blocks = static_cast<unsigned>((mn + 255) / 256);
    hipExtLaunchKernelGGL(reduce_slices_kernel, dim3(blocks), dim3(256), 0, s, start, stop, 0, part, out, slices, mn);
    return hipGetLastError();
}

}  // namespace deac
