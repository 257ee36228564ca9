// Operand layouts of the K.S contraction kernels (precision_kernels.hip),
// built on the host: the kernel matrix once per problem and, for
// deac_model_isf, the stack of genomes.  The split generation loop packs its
// trial genomes into the same layout on the device (split_kernels.hip).
//
// Lane order, one coalesced load per wave and k-block:
//   FP64, FP32  slot (tile, kb, lane)     = X[16 tile + lane%16][4 kb + lane/16]
//   BF16X3      slot (tile, kb, lane, e)  = X[16 tile + lane%16][32 kb + 8 (lane/16) + e]
// with X = K (tiles of 16 timeslices) or S (tiles of 16 genomes).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "deac_engine.h"
#include "host_model.hpp"
#include "precision_kernels.hpp"

namespace deac {

// Round-to-nearest-even FP32 -> bfloat16 bits (finite inputs).
inline uint16_t bf16_bits(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, sizeof u);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return static_cast<uint16_t>(u >> 16);
}

inline float bf16_value(uint16_t b)
{
    const uint32_t u = static_cast<uint32_t>(b) << 16;
    float x;
    std::memcpy(&x, &u, sizeof x);
    return x;
}

// x = hi + lo: hi = bf16(x), lo = bf16(x - hi), the remainder taken in FP32.
inline void bf16_split(double x, uint16_t &hi, uint16_t &lo)
{
    const float f = static_cast<float>(x);
    hi = bf16_bits(f);
    lo = bf16_bits(f - bf16_value(hi));
}

struct KsLayout {
    int precision = DEAC_PRECISION_FP64;
    size_t T = 0, N = 0, L = 0;        // timeslices, genomes, genome length
    int Mt = 0, Nt = 0;                // tiles of 16 timeslices and of 16 genomes
    size_t kblock = 4;                 // positions per k-block: 4, or 32 for BF16X3
    size_t per = 1;                    // values per lane and k-block: 1, or 8 for BF16X3
    size_t Kb = 0;                     // k-blocks, a multiple of the kernels' prefetch depth
    int kb_per_slice = 0, slices = 1;  // genome slices (FP32, BF16X3)
    std::vector<int> ex;               // K row t is scaled by 2^-ex[t] (FP32, BF16X3)

    size_t mpad() const { return static_cast<size_t>(Mt) * 16; }
    size_t npad() const { return static_cast<size_t>(Nt) * 16; }
    size_t a_count() const { return static_cast<size_t>(Mt) * Kb * 64 * per; }
    size_t b_count() const { return static_cast<size_t>(Nt) * Kb * 64 * per; }
    // operand slot (tile, kb, lane, e) -> row within the operand, position
    void at(int tile, size_t kb, int l, size_t e, size_t &row, size_t &k) const
    {
        row = static_cast<size_t>(tile) * 16 + (l & 15);
        k = kblock * kb + (per == 8 ? 8 * static_cast<size_t>(l >> 4) + e : static_cast<size_t>(l >> 4));
    }
    size_t slot(int tile, size_t kb, int l, size_t e) const
    {
        return ((static_cast<size_t>(tile) * Kb + kb) * 64 + l) * per + e;
    }
};

inline KsLayout ks_layout(const Model &m, int precision, size_t count)
{
    KsLayout y;
    y.precision = precision;
    y.T = m.T;
    y.N = count;
    y.L = m.Kc;
    y.Mt = static_cast<int>((m.T + 15) / 16);
    y.Nt = static_cast<int>((count + 15) / 16);
    const bool bf = precision == DEAC_PRECISION_BF16X3;
    y.kblock = bf ? 32 : 4;
    y.per = bf ? 8 : 1;
    const size_t depth = bf ? kSweepDepth16 : kSweepDepth;
    y.Kb = (y.L + y.kblock * depth - 1) / (y.kblock * depth) * depth;  // zero padded
    if (precision != DEAC_PRECISION_FP64) {
        // K row t scaled by 2^-e_t (exact), e_t the exponent of its largest
        // magnitude, so every row's largest entry lies in [1, 2)
        y.ex.assign(m.T, 0);
        for (size_t t = 0; t < m.T; t++) {
            double mx = 0.0;
            for (size_t k = 0; k < y.L; k++) mx = std::fmax(mx, std::fabs(m.kernel[t * y.L + k]));
            y.ex[t] = mx > 0.0 ? std::ilogb(mx) : 0;
        }
        // genome slices: enough workgroups (about 2048) to fill the device,
        // at least 8 k-blocks per slice.  (Fewer, longer slices are cheaper
        // to sum but longer FP32 accumulations: capping them at 32 at
        // T=257, G=16384, P=64 raised the model-ISF error 100-fold.)
        const size_t tiles = static_cast<size_t>(y.Mt) * ((y.Nt + 3) / 4);
        const size_t want =
            std::max<size_t>(1, std::min<size_t>((2048 + tiles - 1) / tiles, std::max<size_t>(y.Kb / 8, 1)));
        y.kb_per_slice = static_cast<int>(((y.Kb + want - 1) / want + depth - 1) / depth * depth);
        y.slices = static_cast<int>((y.Kb + y.kb_per_slice - 1) / y.kb_per_slice);
    }
    return y;
}

// K operand, FP64: K as it is, padded with -0.0 (with genomes padded with
// +0.0 the padded products are -0.0 and leave every chain exactly as it is).
inline void ks_kernel_fp64(const Model &m, const KsLayout &y, std::vector<double> &A)
{
    A.assign(y.a_count(), -0.0);
    for (int mt = 0; mt < y.Mt; mt++)
        for (size_t kb = 0; kb < y.Kb; kb++)
            for (int l = 0; l < 64; l++) {
                size_t t, k;
                y.at(mt, kb, l, 0, t, k);
                if (t < y.T && k < y.L) A[y.slot(mt, kb, l, 0)] = m.kernel[t * y.L + k];
            }
}

// K operand, FP32: rows scaled by 2^-ex[t], rounded to FP32, padded with 0.
inline void ks_kernel_fp32(const Model &m, const KsLayout &y, std::vector<float> &A)
{
    A.assign(y.a_count(), 0.0f);
    for (int mt = 0; mt < y.Mt; mt++)
        for (size_t kb = 0; kb < y.Kb; kb++)
            for (int l = 0; l < 64; l++) {
                size_t t, k;
                y.at(mt, kb, l, 0, t, k);
                if (t < y.T && k < y.L)
                    A[y.slot(mt, kb, l, 0)] = static_cast<float>(std::ldexp(m.kernel[t * y.L + k], -y.ex[t]));
            }
}

// K operand, BF16X3: rows scaled by 2^-ex[t], split into bfloat16 hi and lo.
inline void ks_kernel_bf16x3(const Model &m, const KsLayout &y, std::vector<uint16_t> &Ah, std::vector<uint16_t> &Al)
{
    Ah.assign(y.a_count(), 0);
    Al.assign(y.a_count(), 0);
    for (int mt = 0; mt < y.Mt; mt++)
        for (size_t kb = 0; kb < y.Kb; kb++)
            for (int l = 0; l < 64; l++)
                for (size_t e = 0; e < 8; e++) {
                    size_t t, k;
                    y.at(mt, kb, l, e, t, k);
                    if (t < y.T && k < y.L) {
                        const size_t i = y.slot(mt, kb, l, e);
                        bf16_split(std::ldexp(m.kernel[t * y.L + k], -y.ex[t]), Ah[i], Al[i]);
                    }
                }
}

// Genome operand from count rows of L doubles in host memory, padded with +0.0.
inline void ks_genomes_fp64(const KsLayout &y, const double *g, std::vector<double> &B)
{
    B.assign(y.b_count(), 0.0);
    for (int nt = 0; nt < y.Nt; nt++)
        for (size_t kb = 0; kb < y.Kb; kb++)
            for (int l = 0; l < 64; l++) {
                size_t n, k;
                y.at(nt, kb, l, 0, n, k);
                if (n < y.N && k < y.L) B[y.slot(nt, kb, l, 0)] = g[n * y.L + k];
            }
}

inline void ks_genomes_fp32(const KsLayout &y, const double *g, std::vector<float> &B)
{
    B.assign(y.b_count(), 0.0f);
    for (int nt = 0; nt < y.Nt; nt++)
        for (size_t kb = 0; kb < y.Kb; kb++)
            for (int l = 0; l < 64; l++) {
                size_t n, k;
                y.at(nt, kb, l, 0, n, k);
                if (n < y.N && k < y.L) B[y.slot(nt, kb, l, 0)] = static_cast<float>(g[n * y.L + k]);
            }
}

inline void ks_genomes_bf16x3(const KsLayout &y, const double *g, std::vector<uint16_t> &Bh, std::vector<uint16_t> &Bl)
{
    Bh.assign(y.b_count(), 0);
    Bl.assign(y.b_count(), 0);
    for (int nt = 0; nt < y.Nt; nt++)
        for (size_t kb = 0; kb < y.Kb; kb++)
            for (int l = 0; l < 64; l++)
                for (size_t e = 0; e < 8; e++) {
                    size_t n, k;
                    y.at(nt, kb, l, e, n, k);
                    if (n < y.N && k < y.L) {
                        const size_t i = y.slot(nt, kb, l, e);
                        bf16_split(g[n * y.L + k], Bh[i], Bl[i]);
                    }
                }
}

}  // namespace deac
