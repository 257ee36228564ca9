// Device helpers shared by the fused evolution kernel (kernels.hip) and the
// split generation loop (split_kernels.hip): member rows, jDE controls, trial
// values and the reference's non-fused dot product.
#pragma once

#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace deac {

namespace {

// Chain of one (member, row) pair in the order of the reference's inlined
// matrix_multiply_MxN_by_Nx1 (deac.cpp:102-108): rounded products added one at
// a time, except the last element of each sign when G is odd, which is fused.
__device__ double fold_row(const double *x, const double *row, int Kc, int G)
{
    constexpr int kBatch = 16;
    double acc = 0.0;
    const bool godd = (G & 1) != 0;
    for (int k0 = 0; k0 < Kc; k0 += kBatch) {
        double xv[kBatch], rv[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int k = k0 + u;
            xv[u] = k < Kc ? x[k] : 0.0;
            rv[u] = k < Kc ? row[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int k = k0 + u;
            if (k >= Kc) break;
            if (godd && (k == G - 1 || k == Kc - 1)) acc = fma(xv[u], rv[u], acc);
            else acc = acc + xv[u] * rv[u];
        }
    }
    return acc;
}

// Genome row of member i that the chain phase reads: the current genome for
// the initial fitness, the trial (in the spare buffer) otherwise.
__device__ __forceinline__ const double *chain_row(const EvoArgs &a, int i, bool init)
{
    const int c = i < a.P ? a.cur[i] : 0;
    const bool second = init ? c != 0 : c == 0;
    return (second ? a.pop1 : a.pop0) + static_cast<size_t>(i) * a.stride;
}

// CR and F of member i, sign sg in the current generation.
__device__ __forceinline__ void control_of(const EvoArgs &a, const MemberCtl &c, int i, int sg, double &cr, double &f)
{
    cr = (c.flags >> (2 * sg)) & 1 ? c.cr_value[sg] : a.cr_old[i * a.nsign + sg];
    f = (c.flags >> (2 * sg + 1)) & 1 ? c.f_value[sg] : a.f_old[i * a.nsign + sg];
}

__device__ __forceinline__ const double *current_row(const EvoArgs &a, int i)
{
    return (a.cur[i] ? a.pop1 : a.pop0) + static_cast<size_t>(i) * a.stride;
}

// Trial value at genome position k of member i (deac.cpp:1261-1304).  The
// parent and donor values are all loaded, whichever the crossover keeps, so
// the loads issue together instead of behind the comparison.
__device__ __forceinline__ double trial_value(const EvoArgs &a, const MemberCtl &c, const uint64_t *raw, int i, int k)
{
    const int sg = k >= a.G ? 1 : 0;
    double cr, f;
    control_of(a, c, i, sg, cr, f);
    const double u = static_cast<double>(raw[k]) * 0x1.0p-53;
    const double x1 = current_row(a, c.mutant[0])[k];
    const double x2 = current_row(a, c.mutant[1])[k];
    const double x3 = current_row(a, c.mutant[2])[k];
    const double xp = current_row(a, i)[k];
    const double v = fma(f, x2 - x3, x1);
    const double m = a.allow_neg ? v : fabs(v);
    return u < cr ? m : xp;
}

}  // namespace

}  // namespace deac
