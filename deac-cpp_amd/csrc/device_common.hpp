// Device helpers shared by the gfx950 kernels: the xoshiro256+ step, its
// exact uniform, the modulo of the donor draws, GF(2) parity and the
// s_waitcnt forms the pipelines use.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace deac {

__device__ __forceinline__ uint64_t xnext(uint64_t (&s)[4])  // rng.hpp:25-39
{
    // The reference's five in-place xors, written as three-input xors of the
    // old state (v_xor3_b32) and a rotate by two funnel shifts.
    const uint64_t s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3];
    const uint64_t out = s0 + s3;
    const uint64_t t = s1 << 17;
    const uint64_t n3 = s3 ^ s1;
    s[1] = s1 ^ s2 ^ s0;
    s[0] = s0 ^ n3;
    s[2] = s2 ^ s0 ^ t;
    const uint32_t lo = static_cast<uint32_t>(n3), hi = static_cast<uint32_t>(n3 >> 32);
    // rotl(n3, 45) = rotr(n3, 19)
    const uint32_t rlo = __builtin_amdgcn_alignbit(hi, lo, 19), rhi = __builtin_amdgcn_alignbit(lo, hi, 19);
    s[3] = (static_cast<uint64_t>(rhi) << 32) | rlo;
    return out;
}

// (x >> 11) * 2^-53, exact (deac.cpp:437)
__device__ __forceinline__ double to_unit(uint64_t x) { return static_cast<double>(x >> 11) * 0x1.0p-53; }

// x % P for P < 2^32 with magic = floor((2^64-1)/P): the quotient estimate is
// low by at most one, fixed by one conditional subtraction.
__device__ __forceinline__ uint32_t mod_p(uint64_t x, uint64_t magic, uint32_t P)
{
    const uint64_t q = __umul64hi(x, magic);
    uint64_t r = x - q * P;
    if (r >= P) r -= P;
    return static_cast<uint32_t>(r);
}

__device__ __forceinline__ int parity64(uint64_t x)
{
    return __popc(static_cast<uint32_t>(x) ^ static_cast<uint32_t>(x >> 32)) & 1;
}

}  // namespace deac

// s_waitcnt with vmcnt <= n (n < 64) and the other counters unconstrained.
#define DEAC_WAIT_VM(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (((n) >> 4) << 14) | (7 << 4) | (15 << 8))
// s_waitcnt lgkmcnt(0) (LDS operations retired), other counters unconstrained.
#define DEAC_WAIT_LDS() __builtin_amdgcn_s_waitcnt(15 | (3 << 14) | (7 << 4))
