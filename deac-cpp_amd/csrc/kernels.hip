// gfx950 kernels of the DEAC engine, plain FP64 form.
//
// One generation of the reference loop (deac.cpp:1138-1573) is split in
// two independent parts:
//   * the random choices, which depend only on the xoshiro256+ stream:
//       walk_kernel  - one wavefront walks the sequential, data-dependent
//                      part of the stream (jDE F/CR redraws, rejection-
//                      sampled donor indices) and skips each member's 2G
//                      crossover-mask draws with one GF(2) jump;
//       mask_kernel  - one wavefront per member regenerates the mask draws:
//                      lane l jumps to draw l*chunk with a nibble table and
//                      steps chunk draws;
//   * evolve_kernel - a persistent kernel (all workgroups co-resident) that runs a batch of
//     generations: best-fitness/stats/stop test, trial vectors (DE/rand/1,
//     binomial crossover, |.|, zero-frequency match), the model ISF
//     E = K.S as exact FP64 FMA chains (4x4x4 FP64 MFMAs), chi^2 in the reference's summation
//     order, greedy selection.  Grid barriers separate the phases.
// The file is compiled with -ffp-contract=off: every fused multiply-add is
// written as fma().
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "evolve_common.hpp"
#include "kernels.hpp"

namespace deac {

namespace {

constexpr int kWave = 64;
constexpr int kEvoThreads = 256;
constexpr int kEvoWaves = kEvoThreads / kWave;

// Grid-wide barrier of the persistent evolution kernel.  The wait is
// bounded: after kBarrierSpins polls the workgroup records a failure and every
// caller leaves the kernel, so a broken launch cannot hang the device.
constexpr unsigned kBarrierSpins = 1u << 24;

__device__ __forceinline__ bool grid_barrier(unsigned *bar, unsigned &target, int *status)
{
    // Each wave drains its own memory operations; then one thread per
    // workgroup publishes the workgroup's stores at agent scope (L2 write-back,
    // visible to the other XCDs), arrives, waits, and invalidates this CU's
    // caches for what the other workgroups wrote.
    __shared__ int timed_out;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    target += gridDim.x;
    if (threadIdx.x == 0) {
        timed_out = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins == kBarrierSpins) {
                timed_out = 1;
                __hip_atomic_store(status + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return timed_out == 0;
}

// The walk's stream is wave-uniform: its step as scalar (SALU) operations,
// the rotate as two 64-bit shifts (the funnel-shift form of xnext would put
// it on the VALU and bring it back with v_readfirstlane on every draw).
__device__ __forceinline__ uint64_t xnext_uniform(uint64_t (&s)[4])  // rng.hpp:25-39
{
    const uint64_t s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3];
    const uint64_t out = s0 + s3;
    const uint64_t t = s1 << 17;
    const uint64_t n3 = s3 ^ s1;
    s[1] = s1 ^ s2 ^ s0;
    s[0] = s0 ^ n3;
    s[2] = s2 ^ s0 ^ t;
    s[3] = (n3 << 45) | (n3 >> 19);
    return out;
}

// mod_p for a wave-uniform draw: the remainder before the correction is
// below 2P < 2^32, so the correction compares 32-bit words (SALU)
__device__ __forceinline__ uint32_t mod_p_uniform(uint64_t x, uint64_t magic, uint32_t P)
{
    const uint64_t q = __umul64hi(x, magic);
    uint32_t r = static_cast<uint32_t>(x - q * P);
    if (r >= P) r -= P;
    return r;
}

}  // namespace

// ---------------------------------------------------------------- walk

__global__ __launch_bounds__(64) void walk_kernel(WalkArgs a)
{
    __shared__ uint64_t rows[256 * 4];
    const int lane = threadIdx.x;
    for (int k = lane; k < 256 * 4; k += kWave) rows[k] = a.skip_rows[k];
    __syncthreads();

    // Every lane carries the same state; lane 0 writes the records.
    uint64_t s[4] = {a.state[0], a.state[1], a.state[2], a.state[3]};
    const uint32_t P = static_cast<uint32_t>(a.P);
    for (int g = 0; g < a.ngen; g++) {
        if (lane == 0)
            for (int w = 0; w < 4; w++) a.gen_start[static_cast<size_t>(g) * 4 + w] = s[w];
        MemberCtl *ctl = a.ctl + static_cast<size_t>(g) * a.P;
        // jDE control redraws, deac.cpp:1180-1213: CR+, F+, CR-, F- per member.
        for (int i = 0; i < a.P; i++) {
            uint32_t flags = 0;
            double crv[2] = {0.0, 0.0}, fv[2] = {0.0, 0.0};
            for (int sg = 0; sg < a.nsign; sg++) {
                if (to_unit(xnext_uniform(s)) < a.tau_cr) {
                    flags |= 1u << (2 * sg);
                    crv[sg] = to_unit(xnext_uniform(s));
                }
                if (to_unit(xnext_uniform(s)) < a.tau_f) {
                    flags |= 2u << (2 * sg);
                    fv[sg] = 2.0 * to_unit(xnext_uniform(s));
                }
            }
            if (lane == 0) {
                ctl[i].flags = flags;
                ctl[i].cr_value[0] = crv[0];
                ctl[i].cr_value[1] = crv[1];
                ctl[i].f_value[0] = fv[0];
                ctl[i].f_value[1] = fv[1];
            }
        }
        // Donor indices (deac.cpp:200-216), then skip the member's masks.
        for (int i = 0; i < a.P; i++) {
            const uint32_t i0 = static_cast<uint32_t>(i);
            uint32_t m0 = i0, m1 = i0, m2 = i0;
            while (m0 == i0) m0 = mod_p_uniform(xnext_uniform(s), a.mod_magic, P);
            while (m1 == i0 || m1 == m0) m1 = mod_p_uniform(xnext_uniform(s), a.mod_magic, P);
            while (m2 == i0 || m2 == m0 || m2 == m1) m2 = mod_p_uniform(xnext_uniform(s), a.mod_magic, P);
            if (lane == 0) {
                ctl[i].mutant[0] = m0;
                ctl[i].mutant[1] = m1;
                ctl[i].mutant[2] = m2;
                for (int w = 0; w < 4; w++) ctl[i].mask_state[w] = s[w];
            }
            // s <- T^(nsign*G) s: lane l forms output bits l, 64+l, 128+l,
            // 192+l as parities of (row & s); ballots assemble the words.
            // (r & s) ^ acc is one v_bitop3 per 32-bit half (truth table
            // 0x6a over the inputs 0xf0, 0xcc, 0xaa).
            uint64_t out[4];
            for (int w = 0; w < 4; w++) {
                const uint64_t *r = rows + (w * 64 + lane) * 4;
                uint32_t lo = static_cast<uint32_t>(r[0]) & static_cast<uint32_t>(s[0]);
                uint32_t hi = static_cast<uint32_t>(r[0] >> 32) & static_cast<uint32_t>(s[0] >> 32);
                for (int k = 1; k < 4; k++) {
                    lo = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(r[k]), static_cast<uint32_t>(s[k]), lo, 0x6a);
                    hi = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(r[k] >> 32), static_cast<uint32_t>(s[k] >> 32),
                                                     hi, 0x6a);
                }
                out[w] = __ballot(__popc(lo ^ hi) & 1);
            }
            for (int w = 0; w < 4; w++) s[w] = out[w];
        }
    }
    if (lane == 0)
        for (int w = 0; w < 4; w++) a.state[w] = s[w];
}

// ---------------------------------------------------------------- masks

__global__ __launch_bounds__(256) void mask_kernel(MaskArgs a)
{
    const int lane = threadIdx.x & (kWave - 1);
    const size_t wave = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) / kWave;
    if (wave >= static_cast<size_t>(a.ngen) * a.P) return;
    const MemberCtl &c = a.ctl[wave];
    uint64_t s[4] = {c.mask_state[0], c.mask_state[1], c.mask_state[2], c.mask_state[3]};
    // s <- T^(lane*chunk) s from the lane's nibble table.
    const uint64_t *tab = a.lane_tab + static_cast<size_t>(lane) * (64 * 16 * 4);
    uint64_t o[4] = {0, 0, 0, 0};
#pragma unroll 4
    for (int n = 0; n < 64; n++) {
        const int v = static_cast<int>((s[n >> 4] >> ((n & 15) * 4)) & 15);
        const uint64_t *e = tab + (n * 16 + v) * 4;
        o[0] ^= e[0];
        o[1] ^= e[1];
        o[2] ^= e[2];
        o[3] ^= e[3];
    }
    for (int w = 0; w < 4; w++) s[w] = o[w];
    uint64_t *raw = a.raw + wave * a.Kc;
    const int k0 = lane * a.chunk;
    for (int c2 = 0; c2 < a.chunk; c2++) {
        const uint64_t x = xnext(s);
        const int k = k0 + c2;
        if (k < a.Kc) raw[k] = x >> 11;
    }
}

// ---------------------------------------------------------------- evolution

namespace {

// LDS of the chain phase: kStages buffers of the kernel-matrix chunk
// [16 tau][kChunk] and of the genome chunk [4 members][kChunk].  Rows are
// padded by one pair so the 16 tau rows start on different banks.
constexpr int kRowPad = kChunk + 2;
constexpr int kKBuf = kChainTau * kRowPad;
constexpr int kSBuf = kChainRows * kRowPad;
constexpr int kChainLds = kStages * (kKBuf + kSBuf);

// Register-staged copies of one chunk: 20 rows of 1 KB (kChunk doubles), 16
// kernel-matrix rows then 4 genome rows, dealt round-robin to the helper
// waves 1-3 (7, 7 and 6 rows); lane l moves bytes [16l, 16l+16) of a row.
// Plain 16-byte loads into a ring of registers, kAhead chunks in flight per
// wave, then ds_write_b128 into the chunk's LDS stage: LDS-DMA copies
// (global_load_lds) stall at issue after a few are in flight per CU.
constexpr int kCopiesPerChunk = kChainTau + kChainRows;
constexpr int kHelperCopies = (kCopiesPerChunk + 2) / 3;
constexpr int kAhead = 4;
static_assert(kChunk == 128, "one copy per row: 64 lanes x 16 bytes");
static_assert(kStages >= 2, "the chunk written at step ch must not be the one read at step ch");

// Two doubles as a native vector: plain 16-byte loads and stores that stay
// in registers (HIP's double2 struct is copied through memory).
typedef double dvec2 __attribute__((ext_vector_type(2)));

// A helper wave's copy rows (the same for the whole task, scalar): row j of
// the wave is copy q = wid-1 + 3j of a chunk -- kernel row q (q < 16), genome
// row q-16 (q < 20) or, for the third wave's seventh slot, none.  base: the
// row's start in global memory (rows past the last timeslice, and the empty
// slot, read the first kernel row; their values are discarded); off: the
// row's offset in stage 0 of the LDS buffers, step: the stage stride.
typedef __attribute__((address_space(1))) const dvec2 global_dvec2;

struct HelperRows {
    const double *b0, *b1, *b2, *b3, *b4, *b5, *b6;
    int o0, o1, o2, o3, o4, o5, o6;
    int k0, k1, k2, k3, k4, k5, k6;  // stage stride of the row's buffer
};

__device__ __forceinline__ void helper_row(const EvoArgs &a, const double *const *srows, int tb, int q,
                                           const double *&b, int &o, int &k)
{
    if (q < kChainTau) {
        const int t = tb * kChainTau + q;
        b = a.kmat + (t < a.T ? static_cast<size_t>(t) * a.stride : 0);
        o = q * kRowPad;
        k = kKBuf;
    } else if (q < kCopiesPerChunk) {
        b = srows[q - kChainTau];
        o = kStages * kKBuf + (q - kChainTau) * kRowPad;
        k = kSBuf;
    } else {
        b = a.kmat;
        o = -1;
        k = 0;
    }
}

__device__ __forceinline__ HelperRows helper_rows(const EvoArgs &a, const double *const *srows, int tb, int wid)
{
    HelperRows h;
    helper_row(a, srows, tb, wid - 1, h.b0, h.o0, h.k0);
    helper_row(a, srows, tb, wid + 2, h.b1, h.o1, h.k1);
    helper_row(a, srows, tb, wid + 5, h.b2, h.o2, h.k2);
    helper_row(a, srows, tb, wid + 8, h.b3, h.o3, h.k3);
    helper_row(a, srows, tb, wid + 11, h.b4, h.o4, h.k4);
    helper_row(a, srows, tb, wid + 14, h.b5, h.o5, h.k5);
    helper_row(a, srows, tb, wid + 17, h.b6, h.o6, h.k6);
    return h;
}

// One ring slot: a helper wave's 7 rows of one chunk, in named registers (no
// array, so nothing is indexed at run time).
struct RowRing {
    dvec2 v0, v1, v2, v3, v4, v5, v6;
};
static_assert(kHelperCopies == 7, "RowRing holds 7 rows");

__device__ __forceinline__ dvec2 load_row(const double *base, int chunk, int lane)
{
    // (a global-address-space pointer: a flat load would also count against
    // the LDS counter, which the helpers drain before every barrier)
    return ((global_dvec2 *)(base + static_cast<size_t>(chunk) * kChunk))[lane];
}

__device__ __forceinline__ void store_row(double *lds, int stage, int off, int step, int lane, const dvec2 &v)
{
    if (off >= 0) reinterpret_cast<dvec2 *>(lds + off + stage * step)[lane] = v;
}

__device__ __forceinline__ void load_chunk(const HelperRows &h, int chunk, int lane, RowRing &r)
{
    r.v0 = load_row(h.b0, chunk, lane);
    r.v1 = load_row(h.b1, chunk, lane);
    r.v2 = load_row(h.b2, chunk, lane);
    r.v3 = load_row(h.b3, chunk, lane);
    r.v4 = load_row(h.b4, chunk, lane);
    r.v5 = load_row(h.b5, chunk, lane);
    r.v6 = load_row(h.b6, chunk, lane);
}

__device__ __forceinline__ void store_chunk(double *lds, int stage, const HelperRows &h, int lane, const RowRing &r)
{
    store_row(lds, stage, h.o0, h.k0, lane, r.v0);
    store_row(lds, stage, h.o1, h.k1, lane, r.v1);
    store_row(lds, stage, h.o2, h.k2, lane, r.v2);
    store_row(lds, stage, h.o3, h.k3, lane, r.v3);
    store_row(lds, stage, h.o4, h.k4, lane, r.v4);
    store_row(lds, stage, h.o5, h.k5, lane, r.v5);
    store_row(lds, stage, h.o6, h.k6, lane, r.v6);
}

// Workgroup barrier for the chunk pipeline.  Every wave retires what the
// others depend on before arriving: the helpers their LDS writes, the
// compute wave its LDS reads (DEAC_WAIT_LDS).
__device__ __forceinline__ void chunk_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Model ISF E[i][t] of one (tau block, member group) task; the whole
// workgroup takes part: wave 0 runs the chains, waves 1-3 stage chunks.
// E = sum over the concatenated genome (positive sign, then negative sign)
// as one sequential FMA chain per (member, tau) starting from zero
// (matrix_multiply_LxM_by_MxN called twice, deac.cpp:1386-1394).
__device__ void chain_task(const EvoArgs &a, double *lds, int tb, int mg, bool init)
{
    // the wave index as a scalar: every per-row choice below is then uniform
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
    __shared__ const double *srows[kChainRows];
    if (threadIdx.x < kChainRows) srows[threadIdx.x] = chain_row(a, mg * kChainRows + threadIdx.x, init);
    __syncthreads();
    const int nchunk = a.stride / kChunk;
    // Chunk c travels in ring slot c % kAhead (r0..r3); at step ch the
    // helpers write chunk ch+1 into LDS stage (ch+1) % kStages and load chunk
    // ch+1+kAhead into the slot it left.  Loads past the last chunk re-read
    // it; the steps run to a multiple of kAhead (the extra ones only meet at
    // the barrier), so every load is unconditional.
    static_assert(kAhead == 4, "four ring slots r0..r3");
    // The compute wave and the helpers run loops of their own, each meeting
    // the other at one barrier per step (the same count: the steps run to a
    // multiple of kAhead).
    const int nsteps = (nchunk + kAhead - 1) / kAhead * kAhead;
    double acc = 0.0;
    if (wid == 0) {
        chunk_barrier();  // chunk 0 staged
        for (int ch = 0; ch < nsteps; ch++) {
            if (ch < nchunk) {
                // 64 chains (16 tau x 4 members) advanced 4 positions per
                // v_mfma_f64_4x4x4_4b_f64, which is bit for bit four fused
                // multiply-adds in ascending position order: block b holds
                // tau rows 4b..4b+3 of the task, every block the same 4
                // members.  A: lane l reads tau row l&15 at position
                // j + (l>>4); B: member l&3 at the same position (a broadcast
                // across the blocks).
                const int st = ch % kStages;
                const double *kb = lds + st * kKBuf + (lane & 15) * kRowPad + (lane >> 4);
                const double *sb = lds + kStages * kKBuf + st * kSBuf + (lane & 3) * kRowPad + (lane >> 4);
                double ka[kChunk / 4], sa[kChunk / 4];
#pragma unroll
                for (int j = 0; j < kChunk / 4; j++) {
                    ka[j] = kb[4 * j];
                    sa[j] = sb[4 * j];
                }
#pragma unroll
                for (int j = 0; j < kChunk / 4; j++) acc = __builtin_amdgcn_mfma_f64_4x4x4f64(ka[j], sa[j], acc, 0, 0, 0);
                DEAC_WAIT_LDS();  // the stage may be refilled after the barrier
            }
            chunk_barrier();
        }
    } else {
        RowRing r0, r1, r2, r3;
        const HelperRows hr = helper_rows(a, srows, tb, wid);
        const int last = nchunk - 1;
        load_chunk(hr, 0, lane, r0);
        load_chunk(hr, min(1, last), lane, r1);
        load_chunk(hr, min(2, last), lane, r2);
        load_chunk(hr, min(3, last), lane, r3);
        store_chunk(lds, 0, hr, lane, r0);
        load_chunk(hr, min(kAhead, last), lane, r0);
        DEAC_WAIT_LDS();
        chunk_barrier();
        // step ch: chunk ch+1 into LDS stage (ch+1) % kStages, then chunk
        // ch+1+kAhead into the ring slot it left
        auto step = [&](int ch, RowRing &rs) __attribute__((always_inline)) {
            if (ch + 1 < nchunk) store_chunk(lds, (ch + 1) % kStages, hr, lane, rs);
            load_chunk(hr, min(ch + 1 + kAhead, last), lane, rs);
            DEAC_WAIT_LDS();  // chunk ch+1 is in LDS before the barrier
            chunk_barrier();
        };
        for (int ch0 = 0; ch0 < nsteps; ch0 += kAhead) {
            step(ch0, r1);
            step(ch0 + 1, r2);
            step(ch0 + 2, r3);
            step(ch0 + 3, r0);
        }
    }
    if (wid == 0) {
        // 4x4x4_4b output: lane l holds row l>>4 of block (l>>2)&3, column l&3
        const int i = mg * kChainRows + (lane & 3), t = tb * kChainTau + 4 * ((lane >> 2) & 3) + (lane >> 4);
        if (i < a.P && t < a.T) a.emod[static_cast<size_t>(i) * a.T + t] = acc;
    }
    __syncthreads();
}

}  // namespace

// Phase timestamps (s_memtime) of workgroup 0, compiled in only for
// profiling builds: slot 0 top of generation, 1 after the trial barrier,
// 2 after the chain barrier, 3 after the selection barrier.
#ifdef DEAC_PHASE_TIMERS
__device__ unsigned long long deac_phase_clock[4];
#define DEAC_PHASE(slot)                                                                      \
    do {                                                                                      \
        if (blockIdx.x == 0 && threadIdx.x == 0 && !init) {                                   \
            const unsigned long long now = clock64();                                         \
            if (slot == 0) phase_t0 = now;                                                     \
            else atomicAdd(&deac_phase_clock[slot], now - phase_t0), phase_t0 = now;           \
        }                                                                                     \
    } while (0)
#else
#define DEAC_PHASE(slot) \
    do {                 \
    } while (0)
#endif

// One workgroup per CU (its LDS fills the CU), so one wave per SIMD: the
// register file is the wave's, and the helpers' row ring fits beside the rest.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void evolve_kernel(EvoArgs a)
{
#ifdef DEAC_PHASE_TIMERS
    unsigned long long phase_t0 = 0;
#endif
    extern __shared__ __attribute__((aligned(16))) double lds[];  // 16-byte aligned: ds_write_b128
    __shared__ double dsh[kEvoWaves][kWave];
    __shared__ int stop_flag;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nwg = gridDim.x;
    const int total_waves = nwg * kEvoWaves;
    // Waves are numbered wave-slot-major so that consecutive tasks land on
    // different workgroups (and CUs).
    const int gwave = wid * nwg + blockIdx.x;
    const size_t gthread = static_cast<size_t>(blockIdx.x) * kEvoThreads + threadIdx.x;
    const size_t nthreads = static_cast<size_t>(nwg) * kEvoThreads;
    unsigned target = 0;
    const int nchain = a.MG * a.TB;
    const int mgrp = (a.P + 7) / 8;  // moment tasks: 8 members x {first, third} per wave
    const int nmom = (a.use_first || a.use_third) ? mgrp : 0;
    const int half = a.stride / 2;

    for (int g = 0; g <= a.ngen; g++) {
        const bool init = a.init_mode != 0;
        if (!init) {
            if (g == a.ngen) break;
            DEAC_PHASE(0);
            // Top of the generation: best fitness, statistics, stop test
            // (deac.cpp:1145-1166).  The fitnesses are staged in LDS with one
            // coalesced load (the chain buffers are idle here) and scanned
            // by thread 0 in the reference's order.
            auto top = [&](const double *fit) __attribute__((always_inline)) {
                double mn = fit[0];
                for (int i = 0; i < a.P; i++)
                    if (fit[i] < mn) mn = fit[i];
                const long ii = a.gen0 + g;
                if (blockIdx.x == 0 && a.track_stats) {
                    double s = 0.0;
                    for (int i = 0; i < a.P; i++) s += fit[i];
                    double q = 0.0;
                    const int p4 = a.P & ~3;
                    for (int i = 0; i < a.P; i++) {
                        const double v = fit[i];
                        if (i < p4) q = q + v * v;
                        else q = fma(v, v, q);
                    }
                    a.stats[ii] = s / static_cast<double>(a.P);
                    a.stats[a.stats_len + ii] = mn;
                    a.stats[2 * a.stats_len + ii] = q / static_cast<double>(a.P);
                }
                const bool stop = mn <= a.stop;
                if (stop && blockIdx.x == 0) {
                    a.status[0] = 1;
                    a.status[1] = g;
                }
                stop_flag = stop ? 1 : 0;
            };
            if (a.P <= kChainLds) {
                for (int i = threadIdx.x; i < a.P; i += kEvoThreads) lds[i] = a.fitness[i];
                __syncthreads();
                if (threadIdx.x == 0) top(lds);
            } else if (threadIdx.x == 0) {
                top(a.fitness);
            }
            __syncthreads();
            if (stop_flag) break;

            // Trial vectors into each member's spare buffer, two genome
            // positions per work item; the padding stays zero.
            const MemberCtl *ctlg = a.ctl + static_cast<size_t>(g) * a.P;
            const uint64_t *rawg = a.raw + static_cast<size_t>(g) * a.P * a.Kc;
            const size_t items = static_cast<size_t>(a.P) * half;
            for (size_t it = gthread; it < items; it += nthreads) {
                const int i = static_cast<int>(it / half);
                const int kp = static_cast<int>(it % half);
                const MemberCtl &c = ctlg[i];
                const uint64_t *raw = rawg + static_cast<size_t>(i) * a.Kc;
                double v[2];
                for (int e = 0; e < 2; e++) {
                    const int k = 2 * kp + e;
                    if (k >= a.Kc) v[e] = 0.0;
                    else if (a.nsign == 2 && k == a.G) v[e] = trial_value(a, c, raw, i, 0);  // deac.cpp:1305-1308
                    else v[e] = trial_value(a, c, raw, i, k);
                }
                double *dst = (a.cur[i] ? a.pop0 : a.pop1) + static_cast<size_t>(i) * a.stride;
                *reinterpret_cast<double2 *>(dst + 2 * kp) = make_double2(v[0], v[1]);
            }
            if (!grid_barrier(a.bar, target, a.status)) return;
            DEAC_PHASE(1);

            if (a.renorm) {
                // Per-generation renormalization (deac.cpp:1342-1364).
                for (int i = gwave; i < a.P; i += total_waves)
                    if (lane == 0) {
                        const double *src = (a.cur[i] ? a.pop0 : a.pop1) + static_cast<size_t>(i) * a.stride;
                        a.mom[i * 4 + 0] = fold_row(src, a.nterm, a.Kc, a.G);
                    }
                if (!grid_barrier(a.bar, target, a.status)) return;
                for (size_t it = gthread; it < items; it += nthreads) {
                    const int i = static_cast<int>(it / half);
                    const int kp = static_cast<int>(it % half);
                    double *dst = (a.cur[i] ? a.pop0 : a.pop1) + static_cast<size_t>(i) * a.stride;
                    double2 *p2 = reinterpret_cast<double2 *>(dst + 2 * kp);
                    double2 v = *p2;
                    const double scale = a.zeroth / a.mom[i * 4 + 0];
                    v.x *= scale;
                    v.y *= scale;
                    *p2 = v;
                }
                if (!grid_barrier(a.bar, target, a.status)) return;
            }
        }

        // Model ISF chains (one workgroup per task) and moment rows.
        for (int task = blockIdx.x; task < nchain; task += nwg) chain_task(a, lds, task % a.TB, task / a.TB, init);
        for (int task = gwave; task < nmom; task += total_waves) {
            const int il = lane & 7, rr = lane >> 3;
            const int i = task * 8 + il;
            if (i < a.P && rr < 2) {
                const double *row = rr == 0 ? a.fterm : a.tterm;
                const bool used = rr == 0 ? a.use_first : a.use_third;
                if (used) a.mom[i * 4 + 1 + rr] = fold_row(chain_row(a, i, init), row, a.Kc, a.G);
            }
        }
        if (!grid_barrier(a.bar, target, a.status)) return;
        DEAC_PHASE(2);

        // Fitness and greedy selection, one wavefront per member
        // (deac.cpp:1539-1572; initial fitness deac.cpp:957-976).
        for (int i = gwave; i < a.P; i += total_waves) {
            double chi = 0.0;
            const int tlast = (a.T & 1) ? a.T - 1 : -1;  // fused element of the reduction
            for (int t0 = 0; t0 < a.T; t0 += kWave) {
                const int t = t0 + lane;
                double d = 0.0;
                if (t < a.T) d = (a.obs[t] - a.emod[static_cast<size_t>(i) * a.T + t]) / a.sig[t];
                dsh[wid][lane] = d;
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (lane == 0) {
                    const int n = min(kWave, a.T - t0);
                    for (int u = 0; u < n; u++) {
                        const double du = dsh[wid][u];
                        if (t0 + u == tlast) chi = fma(du, du, chi);
                        else chi = chi + du * du;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
            if (lane == 0) {
                double fit = chi / static_cast<double>(a.T);
                if (a.use_first) {
                    const double d = a.mom[i * 4 + 1] - a.first_moment;
                    fit += (d * d) / a.first_moment;
                }
                if (a.use_third) {
                    const double q = (a.third_moment - a.mom[i * 4 + 2]) / a.third_moment_error;
                    fit = fma(q, q, fit);
                }
                if (init) {
                    a.fitness[i] = fit;
                } else if (fit <= a.fitness[i]) {
                    const MemberCtl &c = a.ctl[static_cast<size_t>(g) * a.P + i];
                    a.fitness[i] = fit;
                    a.cur[i] ^= 1;
                    for (int sg = 0; sg < a.nsign; sg++) {
                        double cr, f;
                        control_of(a, c, i, sg, cr, f);
                        a.cr_old[i * a.nsign + sg] = cr;
                        a.f_old[i * a.nsign + sg] = f;
                    }
                }
            }
        }
        if (!grid_barrier(a.bar, target, a.status)) return;
        DEAC_PHASE(3);
        if (init) break;
    }
}

__global__ void zero_u32_kernel(unsigned *p, size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- launchers

hipError_t launch_walk(const WalkArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(walk_kernel, dim3(1), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_mask(const MaskArgs &a, hipStream_t s)
{
    const size_t waves = static_cast<size_t>(a.ngen) * a.P;
    const size_t blocks = (waves * kWave + 255) / 256;
    hipLaunchKernelGGL(mask_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// A regular launch: the device holds one cooperative queue, which one
// process keeps once it has used it, so cooperative launches from a second
// process fail.  The grid never exceeds the co-resident limit that
// evolve_max_workgroups reports (one workgroup per CU, bounded by LDS), so on
// an otherwise idle device every workgroup is resident; if one is not, the
// bounded grid barrier reports a timeout instead of hanging.
hipError_t launch_evolve(const EvoArgs &a, int nwg, hipStream_t s)
{
    void *args[] = {const_cast<EvoArgs *>(&a)};
    return hipLaunchKernel(reinterpret_cast<const void *>(evolve_kernel), dim3(nwg), dim3(kEvoThreads), args,
                           kChainLds * sizeof(double), s);
}

hipError_t launch_zero_u32(unsigned *p, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_u32_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

int evolve_max_workgroups(int device)
{
    int per_cu = 0, cus = 0;
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(evolve_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kChainLds * sizeof(double)) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(evolve_kernel),
                                                     kEvoThreads, kChainLds * sizeof(double)) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    // One workgroup per CU: the grid barrier needs every workgroup resident,
    // and a second one per CU is not guaranteed to be even where the
    // occupancy estimate allows it.
    return std::min(per_cu, 1) * cus;
}

}  // namespace deac
