// Device helpers shared by the gfx950 kernels: the xoshiro256+ step, its
// exact uniform, the modulo of the donor draws, GF(2) parity and the
// s_waitcnt forms the pipelines use.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace deac {

__device__ __forceinline__ uint64_t xnext(uint64_t (&s)[4])  // rng.hpp:25-39
{
    const uint64_t out = s[0] + s[3];
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 45) | (s[3] >> 19);
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
