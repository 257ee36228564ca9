// Split generation loop of the plain engine: one generation of the fused
// evolution kernel (kernels.hip) as a sequence of kernels on one stream, so
// that the model ISF can come from the K.S contraction kernels at any of
// their precisions (precision_kernels.hip) while the trial vectors, chi^2,
// the selection, the statistics and every constraint stay FP64 in the
// reference's order.  Per generation g of a batch (deac.cpp:1145-1572):
//
//   split_trial_kernel    best fitness, statistics, stop test (workgroup 0),
//                         trial vectors into each member's spare buffer
//   split_renorm_*        per-generation renormalization, when enabled
//   split_pack_*          trial rows into the contraction's operand layout
//   ks_*_kernel           the contraction: FP64 exact chains straight into
//                         the model ISF, or FP32 / BF16x3 genome slices
//   split_fold_kernel     (bdsf at FP64, in place of the two above) the
//                         reference build's rounded-product fold
//   split_model_kernel    FP32 slices summed in slice order and scaled back
//                         per tau row into the FP64 model ISF
//   split_moment_kernel   first / third moment rows, when used
//   split_select_kernel   chi^2 and greedy selection
//
// A stop found at the top of generation g sets status[0]; every later kernel
// (of this batch and of any batch queued behind it) reads it first and
// returns, leaving the state of the stop.
// The file is compiled with -ffp-contract=off: every fused multiply-add is
// written as fma().
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.hpp"
#include "evolve_common.hpp"
#include "split_kernels.hpp"

namespace deac {

namespace {

constexpr int kThreads = 256;
constexpr int kWave = 64;

__device__ __forceinline__ bool stopped(const EvoArgs &a) { return a.status[0] != 0; }

__device__ __forceinline__ size_t grid_thread() { return static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; }
__device__ __forceinline__ size_t grid_threads() { return static_cast<size_t>(gridDim.x) * blockDim.x; }

// Top of generation g: best fitness, statistics, stop test (deac.cpp:1145-1166),
// scanned by thread 0 in the reference's order from fit(i).
template <class Fit>
__device__ void top_scan(const EvoArgs &a, int g, Fit fit)
{
    double mn = fit(0);
    for (int i = 0; i < a.P; i++)
        if (fit(i) < mn) mn = fit(i);
    const long ii = (a.iter ? a.iter[0] : a.gen0) + g;
    if (a.track_stats) {
        double s = 0.0;
        for (int i = 0; i < a.P; i++) s += fit(i);
        double q = 0.0;
        const int p4 = a.P & ~3;
        for (int i = 0; i < a.P; i++) {
            const double v = fit(i);
            if (i < p4) q = q + v * v;
            else q = fma(v, v, q);
        }
        a.stats[ii] = s / static_cast<double>(a.P);
        a.stats[a.stats_len + ii] = mn;
        a.stats[2 * a.stats_len + ii] = q / static_cast<double>(a.P);
    }
    if (mn <= a.stop) {
        a.status[0] = 1;
        a.status[1] = g;
    }
}

// Run by every thread of one workgroup: the fitnesses are staged in LDS by
// one coalesced load (up to kTopLds of them), so the sequential scans do not
// wait on a load each.
constexpr int kTopLds = 512;
constexpr int kSelectLds = 4096;  // residuals staged per member (32 KB)

__device__ void top_of_generation(const EvoArgs &a, int g)
{
    __shared__ double fs[kTopLds];
    if (a.P <= kTopLds) {
        for (int i = threadIdx.x; i < a.P; i += blockDim.x) fs[i] = a.fitness[i];
        __syncthreads();
        if (threadIdx.x == 0) top_scan(a, g, [&](int i) { return fs[i]; });
    } else if (threadIdx.x == 0) {
        top_scan(a, g, [&](int i) { return a.fitness[i]; });
    }
}

// Trial vectors of generation g into each member's spare buffer, two genome
// positions per work item (deac.cpp:1261-1308); the padding stays zero.
// Workgroup 0 first runs the top of generation g: it reads only the
// fitnesses, which no trial writes, and if its stop test fires the trials
// are never selected (every later kernel of the generation returns).
__global__ __launch_bounds__(kThreads) void split_trial_kernel(EvoArgs a, int g)
{
    if (stopped(a)) return;
    if (a.iter) g = static_cast<int>(a.iter[1]);
    if (blockIdx.x == 0) top_of_generation(a, g);
    const int half = a.stride / 2;
    const MemberCtl *ctlg = a.ctl + static_cast<size_t>(g) * a.P;
    const uint64_t *rawg = a.raw + static_cast<size_t>(g) * a.P * a.Kc;
    const size_t items = static_cast<size_t>(a.P) * half;
    for (size_t it = grid_thread(); it < items; it += grid_threads()) {
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
}

// Per-generation renormalization (deac.cpp:1342-1364): zeroth moment of each
// trial, one thread per member, then every position scaled.
__global__ __launch_bounds__(kWave) void split_renorm_sum_kernel(EvoArgs a)
{
    if (stopped(a)) return;
    const int i = static_cast<int>(grid_thread());
    if (i >= a.P) return;
    const double *src = (a.cur[i] ? a.pop0 : a.pop1) + static_cast<size_t>(i) * a.stride;
    a.mom[i * 4 + 0] = fold_row(src, a.nterm, a.Kc, a.G);
}

__global__ __launch_bounds__(kThreads) void split_renorm_scale_kernel(EvoArgs a)
{
    if (stopped(a)) return;
    const int half = a.stride / 2;
    const size_t items = static_cast<size_t>(a.P) * half;
    for (size_t it = grid_thread(); it < items; it += grid_threads()) {
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
}

// Genome operand of the contraction, 16x16x4 lane order:
// B[nt][kb][lane] = row(16 nt + lane%16)[4 kb + lane/16], padded with +0.0.
// The rows are the trials, or the current genomes for the initial fitness.
template <class Out>
__global__ __launch_bounds__(kThreads) void split_pack_kernel(EvoArgs a, int init, Out *B, int Nt, int Kb)
{
    if (stopped(a)) return;
    const size_t slots = static_cast<size_t>(Nt) * Kb * 64;
    for (size_t s = grid_thread(); s < slots; s += grid_threads()) {
        const int l = static_cast<int>(s & 63);
        const size_t kb = (s >> 6) % Kb;
        const int n = static_cast<int>((s >> 6) / Kb) * 16 + (l & 15);
        const size_t k = 4 * kb + (l >> 4);
        double v = 0.0;
        if (n < a.P && k < static_cast<size_t>(a.Kc)) v = chain_row(a, n, init != 0)[k];
        B[s] = static_cast<Out>(v);
    }
}

// Round-to-nearest-even FP32 -> bfloat16 bits (finite inputs).
__device__ __forceinline__ uint32_t bf16_rne(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}

// Genome operand, 16x16x32 lane order, each value split x = hi + lo with
// hi = bf16(x) and lo = bf16(x - hi), the remainder taken in FP32:
// slot (nt, kb, lane) holds positions 32 kb + 8 (lane/16) + e, e = 0..7.
__global__ __launch_bounds__(kThreads) void split_pack_bf16x3_kernel(EvoArgs a, int init, uint4 *Bh, uint4 *Bl,
                                                                      int Nt, int Kb)
{
    if (stopped(a)) return;
    const size_t slots = static_cast<size_t>(Nt) * Kb * 64;
    for (size_t s = grid_thread(); s < slots; s += grid_threads()) {
        const int l = static_cast<int>(s & 63);
        const size_t kb = (s >> 6) % Kb;
        const int n = static_cast<int>((s >> 6) / Kb) * 16 + (l & 15);
        const size_t k0 = 32 * kb + 8 * static_cast<size_t>(l >> 4);
        const double *row = n < a.P ? chain_row(a, n, init != 0) : nullptr;
        uint32_t hi[4] = {0, 0, 0, 0}, lo[4] = {0, 0, 0, 0};
        for (int e = 0; e < 8; e++) {
            const size_t k = k0 + e;
            if (row && k < static_cast<size_t>(a.Kc)) {
                const float f = static_cast<float>(row[k]);
                const uint32_t h = bf16_rne(f);
                const uint32_t r = bf16_rne(f - __uint_as_float(h << 16));
                hi[e >> 1] |= h << (16 * (e & 1));
                lo[e >> 1] |= r << (16 * (e & 1));
            }
        }
        Bh[s] = make_uint4(hi[0], hi[1], hi[2], hi[3]);
        Bl[s] = make_uint4(lo[0], lo[1], lo[2], lo[3]);
    }
}

// Model ISF from the FP32 genome slices: emod[n][t] = 2^ex[t] * (sum over
// slices, in slice order, in FP32, of part[slice][t][n]).
__global__ __launch_bounds__(kThreads) void split_model_kernel(EvoArgs a, const float *part, int slices, int Npad,
                                                               size_t mn, const int *ex)
{
    if (stopped(a)) return;
    // consecutive threads take consecutive genomes: coalesced slice reads
    const size_t total = static_cast<size_t>(a.P) * a.T;
    for (size_t idx = grid_thread(); idx < total; idx += grid_threads()) {
        const int t = static_cast<int>(idx / a.P), n = static_cast<int>(idx % a.P);
        const size_t i = static_cast<size_t>(t) * Npad + n;
        float s = 0.0f;
#pragma unroll 8
        for (int k = 0; k < slices; k++) s += part[static_cast<size_t>(k) * mn + i];
        a.emod[static_cast<size_t>(n) * a.T + t] = ldexp(static_cast<double>(s), ex[t]);
    }
}

// The trials' model ISF as the detailed-balance (bdsf) reference build sums
// it (its loop inlines matrix_multiply_LxM_by_MxN, deac.cpp:110-118,
// 1389-1390, as a fold of rounded products, the last of an odd count fused):
// one thread per (member, tau), straight into emod.
__global__ __launch_bounds__(kThreads) void split_fold_kernel(EvoArgs a)
{
    if (stopped(a)) return;
    const size_t total = static_cast<size_t>(a.P) * a.T;
    for (size_t idx = grid_thread(); idx < total; idx += grid_threads()) {
        const int n = static_cast<int>(idx / a.T), t = static_cast<int>(idx % a.T);
        a.emod[idx] = fold_row(chain_row(a, n, false), a.kmat + static_cast<size_t>(t) * a.stride, a.Kc, a.G);
    }
}

// First and third moment rows of the contraction's genomes, one thread per
// (member, moment) (deac.cpp:1420-1476).
__global__ __launch_bounds__(kWave) void split_moment_kernel(EvoArgs a, int init)
{
    if (stopped(a)) return;
    const int idx = static_cast<int>(grid_thread());
    const int i = idx >> 1, rr = idx & 1;
    if (i >= a.P) return;
    const bool used = rr == 0 ? a.use_first : a.use_third;
    if (!used) return;
    const double *row = rr == 0 ? a.fterm : a.tterm;
    a.mom[i * 4 + 1 + rr] = fold_row(chain_row(a, i, init != 0), row, a.Kc, a.G);
}

// Fitness and greedy selection, one wavefront per member (deac.cpp:1539-1572;
// the initial fitness, deac.cpp:957-976).  The wavefront forms the weighted
// residuals; lane 0 adds their squares in tau order, the last one fused when
// T is odd.
__global__ __launch_bounds__(kWave) void split_select_kernel(EvoArgs a, int g, int init)
{
    if (stopped(a)) return;
    if (a.iter && !init) g = static_cast<int>(a.iter[1]);
    extern __shared__ double dsh[];  // T residuals (T <= kSelectLds), else one window of kWave
    const int i = blockIdx.x, lane = threadIdx.x;
    double chi = 0.0;
    const int tlast = (a.T & 1) ? a.T - 1 : -1;
    auto residual = [&](int t) { return (a.obs[t] - a.emod[static_cast<size_t>(i) * a.T + t]) / a.sig[t]; };
    auto add = [&](int t, double du) { chi = t == tlast ? fma(du, du, chi) : chi + du * du; };
    if (a.T <= kSelectLds) {
        // every residual at once (their loads all in flight), then the sum
        for (int t = lane; t < a.T; t += kWave) dsh[t] = residual(t);
        __syncthreads();
        if (lane == 0)
            for (int t = 0; t < a.T; t++) add(t, dsh[t]);
    } else {
        for (int t0 = 0; t0 < a.T; t0 += kWave) {
            const int t = t0 + lane;
            dsh[lane] = t < a.T ? residual(t) : 0.0;
            __syncthreads();
            if (lane == 0)
                for (int u = 0; u < min(kWave, a.T - t0); u++) add(t0 + u, dsh[u]);
            __syncthreads();
        }
    }
    if (lane != 0) return;
    double fit = chi / static_cast<double>(a.T);
    if (a.use_nfm) {  // deac.cpp:961-965, 1398-1416, 1543-1547
        const double mi = fold_row(a.emod + static_cast<size_t>(i) * a.T, a.nfm_dt, a.T, a.T);
        const double q = (a.nfm_value - mi) / a.nfm_error;
        fit = fma(q, q, fit);
    }
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

// The generation counter of a captured batch: set = 1 starts a batch at
// gen0 (generation 0 of the batch), set = 0 moves to the next generation.
__global__ void split_iter_kernel(long *iter, long gen0, int set)
{
    if (set) {
        iter[0] = gen0;
        iter[1] = 0;
    } else {
        iter[1] += 1;
    }
}

unsigned grid_for(size_t work, unsigned threads)
{
    const size_t blocks = (work + threads - 1) / threads;
    return static_cast<unsigned>(std::max<size_t>(1, std::min<size_t>(blocks, 4096)));
}

}  // namespace

hipError_t launch_split_trial(const EvoArgs &a, int g, hipStream_t s)
{
    const size_t items = static_cast<size_t>(a.P) * (a.stride / 2);
    hipLaunchKernelGGL(split_trial_kernel, dim3(grid_for(items, kThreads)), dim3(kThreads), 0, s, a, g);
    return hipGetLastError();
}

hipError_t launch_split_renorm(const EvoArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(split_renorm_sum_kernel, dim3(grid_for(a.P, kWave)), dim3(kWave), 0, s, a);
    const size_t items = static_cast<size_t>(a.P) * (a.stride / 2);
    hipLaunchKernelGGL(split_renorm_scale_kernel, dim3(grid_for(items, kThreads)), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_split_pack_fp64(const EvoArgs &a, int init, double *B, int Nt, int Kb, hipStream_t s)
{
    const size_t slots = static_cast<size_t>(Nt) * Kb * 64;
    hipLaunchKernelGGL(split_pack_kernel<double>, dim3(grid_for(slots, kThreads)), dim3(kThreads), 0, s, a, init, B,
                       Nt, Kb);
    return hipGetLastError();
}

hipError_t launch_split_pack_fp32(const EvoArgs &a, int init, float *B, int Nt, int Kb, hipStream_t s)
{
    const size_t slots = static_cast<size_t>(Nt) * Kb * 64;
    hipLaunchKernelGGL(split_pack_kernel<float>, dim3(grid_for(slots, kThreads)), dim3(kThreads), 0, s, a, init, B,
                       Nt, Kb);
    return hipGetLastError();
}

hipError_t launch_split_pack_bf16x3(const EvoArgs &a, int init, void *Bh, void *Bl, int Nt, int Kb, hipStream_t s)
{
    const size_t slots = static_cast<size_t>(Nt) * Kb * 64;
    hipLaunchKernelGGL(split_pack_bf16x3_kernel, dim3(grid_for(slots, kThreads)), dim3(kThreads), 0, s, a, init,
                       static_cast<uint4 *>(Bh), static_cast<uint4 *>(Bl), Nt, Kb);
    return hipGetLastError();
}

hipError_t launch_split_model(const EvoArgs &a, const float *part, int slices, int Npad, size_t mn, const int *ex,
                              hipStream_t s)
{
    const size_t total = static_cast<size_t>(a.P) * a.T;
    hipLaunchKernelGGL(split_model_kernel, dim3(grid_for(total, kThreads)), dim3(kThreads), 0, s, a, part, slices,
                       Npad, mn, ex);
    return hipGetLastError();
}

hipError_t launch_split_fold(const EvoArgs &a, hipStream_t s)
{
    const size_t total = static_cast<size_t>(a.P) * a.T;
    const int blocks = static_cast<int>(std::min<size_t>((total + kThreads - 1) / kThreads, 4096));
    hipLaunchKernelGGL(split_fold_kernel, dim3(blocks), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_split_moments(const EvoArgs &a, int init, hipStream_t s)
{
    hipLaunchKernelGGL(split_moment_kernel, dim3(grid_for(static_cast<size_t>(a.P) * 2, kWave)), dim3(kWave), 0, s, a,
                       init);
    return hipGetLastError();
}

hipError_t launch_split_select(const EvoArgs &a, int g, int init, hipStream_t s)
{
    const size_t lds = static_cast<size_t>(a.T <= kSelectLds ? a.T : kWave) * sizeof(double);
    hipLaunchKernelGGL(split_select_kernel, dim3(a.P), dim3(kWave), lds, s, a, g, init);
    return hipGetLastError();
}

hipError_t launch_split_iter(long *iter, long gen0, int set, hipStream_t s)
{
    hipLaunchKernelGGL(split_iter_kernel, dim3(1), dim3(1), 0, s, iter, gen0, set);
    return hipGetLastError();
}

}  // namespace deac
