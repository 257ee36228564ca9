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
//   * evolve_kernel - a cooperative persistent kernel that runs a batch of
//     generations: best-fitness/stats/stop test, trial vectors (DE/rand/1,
//     binomial crossover, |.|, zero-frequency match), the model ISF
//     E = K.S as exact FP64 FMA chains, chi^2 in the reference's summation
//     order, greedy selection.  Grid barriers separate the phases.
// The file is compiled with -ffp-contract=off: every fused multiply-add is
// written as fma().
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace deac {

namespace {

constexpr int kWave = 64;
constexpr int kEvoThreads = 256;
constexpr int kEvoWaves = kEvoThreads / kWave;

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

// Element index of (member i, genome position k) in the interleaved
// population layout [member-block][pair][8 members][2].
__device__ __forceinline__ size_t pidx(int i, int k, int Kcp)
{
    return ((static_cast<size_t>(i >> 3) * Kcp + (k >> 1)) * 8 + (i & 7)) * 2 + (k & 1);
}

// Grid-wide barrier of the cooperative evolution kernel.  The wait is
// bounded: after kBarrierSpins polls the workgroup records a failure and every
// caller leaves the kernel, so a broken launch cannot hang the device.
constexpr unsigned kBarrierSpins = 1u << 24;

__device__ __forceinline__ bool grid_barrier(unsigned *bar, unsigned &target, int *status)
{
    // Every wave publishes its stores at agent scope before the arrival, so
    // the data is visible to workgroups on other XCDs after the wait.
    __shared__ int timed_out;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    target += gridDim.x;
    if (threadIdx.x == 0) {
        timed_out = 0;
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
            __builtin_amdgcn_s_sleep(1);
            if (++spins == kBarrierSpins) {
                timed_out = 1;
                __hip_atomic_store(status + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return timed_out == 0;
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
                if (to_unit(xnext(s)) < a.tau_cr) {
                    flags |= 1u << (2 * sg);
                    crv[sg] = to_unit(xnext(s));
                }
                if (to_unit(xnext(s)) < a.tau_f) {
                    flags |= 2u << (2 * sg);
                    fv[sg] = 2.0 * to_unit(xnext(s));
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
            while (m0 == i0) m0 = mod_p(xnext(s), a.mod_magic, P);
            while (m1 == i0 || m1 == m0) m1 = mod_p(xnext(s), a.mod_magic, P);
            while (m2 == i0 || m2 == m0 || m2 == m1) m2 = mod_p(xnext(s), a.mod_magic, P);
            if (lane == 0) {
                ctl[i].mutant[0] = m0;
                ctl[i].mutant[1] = m1;
                ctl[i].mutant[2] = m2;
                for (int w = 0; w < 4; w++) ctl[i].mask_state[w] = s[w];
            }
            // s <- T^(nsign*G) s: lane l forms output bits l, 64+l, 128+l,
            // 192+l as parities of (row & s); ballots assemble the words.
            uint64_t out[4];
            for (int w = 0; w < 4; w++) {
                const uint64_t *r = rows + (w * 64 + lane) * 4;
                const uint64_t x = (r[0] & s[0]) ^ (r[1] & s[1]) ^ (r[2] & s[2]) ^ (r[3] & s[3]);
                out[w] = __ballot(parity64(x));
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

// Chain of one (member, row) pair in the order of the reference's inlined
// matrix_multiply_MxN_by_Nx1 (deac.cpp:102-108): rounded products added one at
// a time, except the last element of each sign when G is odd, which is fused.
__device__ double fold_row(const double *trial, int i, const double *row, int Kc, int Kcp, int G)
{
    constexpr int kBatch = 16;
    double acc = 0.0;
    const bool godd = (G & 1) != 0;
    for (int k0 = 0; k0 < Kc; k0 += kBatch) {
        double x[kBatch], r[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int k = k0 + u;
            x[u] = k < Kc ? trial[pidx(i, k, Kcp)] : 0.0;
            r[u] = k < Kc ? row[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kBatch; u++) {
            const int k = k0 + u;
            if (k >= Kc) break;
            if (godd && (k == G - 1 || k == Kc - 1)) acc = fma(x[u], r[u], acc);
            else acc = acc + x[u] * r[u];
        }
    }
    return acc;
}

// E[i][t] for the 8x8 (member, tau) block of this wavefront: one sequential
// FMA chain per lane over the concatenated genome, positive sign then
// negative sign, starting from zero (matrix_multiply_LxM_by_MxN,
// deac.cpp:110-118, called twice at deac.cpp:1386-1394).
__device__ void chain_block(const EvoArgs &a, int mb, int tb, int lane)
{
    const int il = lane >> 3, tl = lane & 7;
    const int i = mb * kMemberBlock + il, t = tb * kTauBlock + tl;
    const double *src = a.pop0;
    if (i < a.P) {
        const int c = a.cur[i];
        src = (a.init_mode ? c : 1 - c) ? a.pop1 : a.pop0;
    }
    const double2 *kp = reinterpret_cast<const double2 *>(a.kmat) + static_cast<size_t>(tb) * a.Kcp * 8 + tl;
    const double2 *sp = reinterpret_cast<const double2 *>(src) + static_cast<size_t>(mb) * a.Kcp * 8 + il;
    double acc = 0.0;
    double2 kb[kPairBatch], sb[kPairBatch];
#pragma unroll
    for (int u = 0; u < kPairBatch; u++) {
        kb[u] = kp[u * 8];
        sb[u] = sp[u * 8];
    }
    for (int jp = kPairBatch; jp < a.Kcp; jp += kPairBatch) {
        double2 kn[kPairBatch], sn[kPairBatch];
#pragma unroll
        for (int u = 0; u < kPairBatch; u++) {
            kn[u] = kp[(jp + u) * 8];
            sn[u] = sp[(jp + u) * 8];
        }
#pragma unroll
        for (int u = 0; u < kPairBatch; u++) {
            acc = fma(kb[u].x, sb[u].x, acc);
            acc = fma(kb[u].y, sb[u].y, acc);
        }
#pragma unroll
        for (int u = 0; u < kPairBatch; u++) {
            kb[u] = kn[u];
            sb[u] = sn[u];
        }
    }
#pragma unroll
    for (int u = 0; u < kPairBatch; u++) {
        acc = fma(kb[u].x, sb[u].x, acc);
        acc = fma(kb[u].y, sb[u].y, acc);
    }
    if (i < a.P && t < a.T) a.emod[static_cast<size_t>(i) * a.T + t] = acc;
}

// CR and F of member i, sign sg in the current generation.
__device__ __forceinline__ void control_of(const EvoArgs &a, const MemberCtl &c, int i, int sg, double &cr, double &f)
{
    cr = (c.flags >> (2 * sg)) & 1 ? c.cr_value[sg] : a.cr_old[i * a.nsign + sg];
    f = (c.flags >> (2 * sg + 1)) & 1 ? c.f_value[sg] : a.f_old[i * a.nsign + sg];
}

// Trial value at genome position k of member i (deac.cpp:1261-1304).
__device__ __forceinline__ double trial_value(const EvoArgs &a, const MemberCtl &c, const uint64_t *raw, int i, int k)
{
    const int sg = k >= a.G ? 1 : 0;
    double cr, f;
    control_of(a, c, i, sg, cr, f);
    const double u = static_cast<double>(raw[k]) * 0x1.0p-53;
    const double *xp = a.cur[i] ? a.pop1 : a.pop0;
    if (u < cr) {
        const uint32_t r1 = c.mutant[0], r2 = c.mutant[1], r3 = c.mutant[2];
        const double x1 = (a.cur[r1] ? a.pop1 : a.pop0)[pidx(r1, k, a.Kcp)];
        const double x2 = (a.cur[r2] ? a.pop1 : a.pop0)[pidx(r2, k, a.Kcp)];
        const double x3 = (a.cur[r3] ? a.pop1 : a.pop0)[pidx(r3, k, a.Kcp)];
        const double v = fma(f, x2 - x3, x1);
        return a.allow_neg ? v : fabs(v);
    }
    return xp[pidx(i, k, a.Kcp)];
}

}  // namespace

__global__ __launch_bounds__(256) void evolve_kernel(EvoArgs a)
{
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
    const int nchain = a.MB * a.TB;
    const int nmom = (a.use_first || a.use_third) ? a.MB : 0;

    for (int g = 0; g <= a.ngen; g++) {
        const bool init = a.init_mode != 0;
        if (!init) {
            if (g == a.ngen) break;
            // Top of the generation: best fitness, statistics, stop test
            // (deac.cpp:1145-1166).
            if (threadIdx.x == 0) {
                double mn = a.fitness[0];
                for (int i = 0; i < a.P; i++)
                    if (a.fitness[i] < mn) mn = a.fitness[i];
                const long ii = a.gen0 + g;
                if (blockIdx.x == 0 && a.track_stats) {
                    double s = 0.0;
                    for (int i = 0; i < a.P; i++) s += a.fitness[i];
                    double q = 0.0;
                    const int p4 = a.P & ~3;
                    for (int i = 0; i < a.P; i++) {
                        const double v = a.fitness[i];
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
            }
            __syncthreads();
            if (stop_flag) break;

            // Trial vectors into the member's spare buffer, two genome
            // positions per work item.
            const MemberCtl *ctlg = a.ctl + static_cast<size_t>(g) * a.P;
            const uint64_t *rawg = a.raw + static_cast<size_t>(g) * a.P * a.Kc;
            const size_t items = static_cast<size_t>(a.P) * a.Kcp;
            for (size_t it = gthread; it < items; it += nthreads) {
                const int i = static_cast<int>(it / a.Kcp);
                const int kp = static_cast<int>(it % a.Kcp);
                const MemberCtl &c = ctlg[i];
                const uint64_t *raw = rawg + static_cast<size_t>(i) * a.Kc;
                double v[2];
                for (int e = 0; e < 2; e++) {
                    const int k = 2 * kp + e;
                    if (k >= a.Kc) v[e] = 0.0;
                    else if (a.nsign == 2 && k == a.G) v[e] = trial_value(a, c, raw, i, 0);  // deac.cpp:1305-1308
                    else v[e] = trial_value(a, c, raw, i, k);
                }
                double *dst = a.cur[i] ? a.pop0 : a.pop1;
                *reinterpret_cast<double2 *>(dst + pidx(i, 2 * kp, a.Kcp)) = make_double2(v[0], v[1]);
            }
            if (!grid_barrier(a.bar, target, a.status)) return;

            if (a.renorm) {
                // Per-generation renormalization (deac.cpp:1342-1364).
                for (int task = gwave; task < a.MB; task += total_waves) {
                    const int i = task * kMemberBlock + lane;
                    if (lane < kMemberBlock && i < a.P) {
                        const double *src = a.cur[i] ? a.pop0 : a.pop1;
                        a.mom[i * 4 + 0] = fold_row(src, i, a.nterm, a.Kc, a.Kcp, a.G);
                    }
                }
                if (!grid_barrier(a.bar, target, a.status)) return;
                for (size_t it = gthread; it < items; it += nthreads) {
                    const int i = static_cast<int>(it / a.Kcp);
                    const int kp = static_cast<int>(it % a.Kcp);
                    double *dst = a.cur[i] ? a.pop0 : a.pop1;
                    double2 *p2 = reinterpret_cast<double2 *>(dst + pidx(i, 2 * kp, a.Kcp));
                    double2 v = *p2;
                    const double scale = a.zeroth / a.mom[i * 4 + 0];
                    v.x *= scale;
                    v.y *= scale;
                    *p2 = v;
                }
                if (!grid_barrier(a.bar, target, a.status)) return;
            }
        }

        // Model ISF chains and moment rows.
        for (int task = gwave; task < nchain + nmom; task += total_waves) {
            if (task < nchain) {
                chain_block(a, task / a.TB, task % a.TB, lane);
            } else {
                const int mb = task - nchain;
                const int il = lane & 7, r = lane >> 3;
                const int i = mb * kMemberBlock + il;
                if (i < a.P && r < 2) {
                    const double *row = r == 0 ? a.fterm : a.tterm;
                    const bool used = r == 0 ? a.use_first : a.use_third;
                    if (used) {
                        const int c = a.cur[i];
                        const double *src = (init ? c : 1 - c) ? a.pop1 : a.pop0;
                        a.mom[i * 4 + 1 + r] = fold_row(src, i, row, a.Kc, a.Kcp, a.G);
                    }
                }
            }
        }
        if (!grid_barrier(a.bar, target, a.status)) return;

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

hipError_t launch_evolve(const EvoArgs &a, int nwg, hipStream_t s)
{
    void *args[] = {const_cast<EvoArgs *>(&a)};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(evolve_kernel), dim3(nwg), dim3(kEvoThreads),
                                      args, 0, s);
}

hipError_t launch_zero_u32(unsigned *p, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_u32_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

int evolve_max_workgroups(int device)
{
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(evolve_kernel),
                                                     kEvoThreads, 0) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    return per_cu * cus;
}

}  // namespace deac
