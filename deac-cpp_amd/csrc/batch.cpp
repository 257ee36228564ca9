// Batched C ABI (deac_batch_* in include/deac_engine.h): many independent
// reconstructions run by one persistent launch, one workgroup each.
//
// Host work: validation and model building per distinct problem (identical
// problems share one device copy of K), the kernel-side layouts (K in MFMA
// tile order, the remainder rows, the walk and sub-stream jump tables), the
// initial populations (on host threads, with each instance's own stream),
// launch slicing with per-instance stop and compaction, and the same
// epilogue as deac_run.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "batch_kernels.hpp"
#include "deac_engine.h"
#include "host_model.hpp"
#include "ks_layout.hpp"

namespace {


#define BATCH_HIP(expr)                        \
    do {                                       \
        if ((expr) != hipSuccess) {            \
            (void)hipGetLastError();           \
            return DEAC_ERR_DEVICE;            \
        }                                      \
    } while (0)

// Device copy of one distinct problem, laid out for the batch kernel.
struct BatchModel {
    deac::Model m;
    int Ppad = 0, NM = 0, NT = 0, C = 0, nchunk = 0, Lpad = 0, Lw = 0, ntt = 0, nrem = 0, S = 0, Lsub = 0;
    int ntm = 0, NQ = 0;                   // mixed K.S: kernel tiles (rows past T zero), pairs per chain wave
    double *ktile = nullptr, *krem = nullptr, *obs = nullptr, *sig = nullptr;
    double *kfold = nullptr;               // bdsf: K transposed [Lpad][fold_rows] for the fold kernel
    double *nfm_dt = nullptr;              // bdsf: -1st-moment weights [T], when used
    int fold_rows = 0, MF = 0;             // bdsf fold kernel: kfold row stride, shape 2 RL + XR
    double *mterm = nullptr;               // [3][Lpad] first-moment, third-moment and normalization weights,
                                           // when moments or per-generation renormalization are used
    uint64_t *walk_rows = nullptr, *sub_tab = nullptr;
    void *kmix[3] = {nullptr, nullptr, nullptr};  // mixed K tiles by precision, built on first use
    double *kscale = nullptr;                     // 2^ex[t] of the mixed K rows
};

// Byte image of everything a problem's results depend on.
std::string problem_key(const deac_problem &p)
{
    std::string k;
    auto put = [&](const void *x, size_t n) { k.append(static_cast<const char *>(x), n); };
    const size_t T = p.number_of_timeslices, G = p.genome_size;
    put(&T, sizeof T);
    put(p.imaginary_time, T * sizeof(double));
    put(p.isf, T * sizeof(double));
    put(p.isf_error, T * sizeof(double));
    put(&G, sizeof G);
    put(p.frequency, G * sizeof(double));
    const deac_problem *q = &p;
    put(&q->temperature, sizeof q->temperature);
    const int ints[7] = {p.model, p.spectra_type, p.zero_temperature, p.allow_negative_spectral_weight,
                         p.renormalize_each_generation, p.normalize, p.use_negative_first_moment};
    put(ints, sizeof ints);
    put(&p.population_size, sizeof p.population_size);
    const double dbl[7] = {p.first_moment, p.third_moment, p.third_moment_error, p.crossover_probability,
                           p.self_adapting_crossover_probability, p.differential_weight,
                           p.self_adapting_differential_weight_probability};
    put(dbl, sizeof dbl);
    return k;
}

std::mutex g_sched_lock;
std::map<std::string, deac_batch_schedule> g_sched_cache;  // tuned schedules by batch signature

}  // namespace

struct deac_batch {
    int device = 0;
    int cu_count = 256;                        // compute units of the device
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, kev0 = nullptr, kev1 = nullptr, kev2 = nullptr, kev3 = nullptr;
    hipEvent_t fork = nullptr;
    // Instances of different kernel shapes run side by side: one stream per
    // shape group (default priority), joined to `stream` by events.
    static constexpr int kGroupStreams = 3;
    hipStream_t gstream[kGroupStreams] = {};
    hipEvent_t gdone[kGroupStreams] = {};
    std::vector<std::unique_ptr<BatchModel>> models;
    std::vector<int> model_of;                 // instance -> model
    std::vector<void *> allocations;
    // per-instance state, packed
    std::vector<size_t> pop_off, mask_off, epar_off;  // offsets into the pools
    double *pop = nullptr, *fitness = nullptr, *cr_old = nullptr, *f_old = nullptr;
    double *epar = nullptr;                    // per instance [T][Ppad]: the members' model ISF (mixed K.S)
    uint32_t *masks = nullptr;
    uint64_t *rng = nullptr;
    long *status = nullptr;                    // per instance: stopped, generation reached
    int *accp = nullptr;                       // per instance [kBatchMaxP]: acceptances of the last generation
                                               // a launch evolved
    std::vector<long> masks_at;                // per instance: generation whose masks (and the acceptances
                                               // before it) its last launch left on the device, or -1
    deac::BatchDesc *descs = nullptr;          // device copy of the launch descriptors
    deac::BatchDesc *hdescs = nullptr;         // their host side, pinned (3 per instance)
    long *hstatus = nullptr;                   // pinned host copy of status
    double *scratch = nullptr;                 // this device's statistics scratch (deac_batch_run_noalloc)
    std::vector<long> reached;                 // per instance: generation reached by the last run (absolute)
    std::vector<char> epar_valid;              // per instance: epar holds the current population's model ISF
    deac_batch_schedule sched{3, 2, 0, 0.0, 0};  // launch schedule (deac_batch_tune)
    std::string signature;                     // shapes of the instances, the process-wide schedule key
    deac_run_profile last{};

    ~deac_batch()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        for (hipStream_t g : gstream)
            if (g) (void)hipStreamSynchronize(g);
        for (void *p : allocations) (void)hipFree(p);
        if (hdescs) (void)hipHostFree(hdescs);
        if (hstatus) (void)hipHostFree(hstatus);
        for (hipEvent_t e : {ev0, ev1, kev0, kev1, kev2, kev3, fork})
            if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : gdone)
            if (e) (void)hipEventDestroy(e);
        for (hipStream_t g : gstream)
            if (g) (void)hipStreamDestroy(g);
        if (stream) (void)hipStreamDestroy(stream);
    }

    template <class T>
    deac_status alloc(T *&out, size_t count)
    {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            return DEAC_ERR_OUT_OF_MEMORY;
        }
        allocations.push_back(p);
        out = static_cast<T *>(p);
        return DEAC_OK;
    }
    void release(void *p)
    {
        auto it = std::find(allocations.begin(), allocations.end(), p);
        if (it != allocations.end()) {
            allocations.erase(it);
            (void)hipFree(p);
        }
    }
};

namespace {

// 1 when a model has moment terms (first, third or -1st moment in the
// fitness, or the zeroth moment of per-generation renormalization), which
// the mixed kernels carry in an instantiation of their own
int mixed_moments(const deac::Model &m)
{
    return (m.use_first || m.use_third || m.use_nfm || (m.normalize && m.renormalize)) ? 1 : 0;
}

// Kernel-side layout of one model; DEAC_ERR_UNSUPPORTED for shapes the batch
// kernel does not run (the single-instance engine runs them).
deac_status plan_model(BatchModel &b, std::string &why)
{
    const deac::Model &m = b.m;
    if (m.P > static_cast<size_t>(deac::kBatchMaxP)) {
        why = "batched runs take at most 64 members per population";
        return DEAC_ERR_UNSUPPORTED;
    }
    const int T = static_cast<int>(m.T), L = static_cast<int>(m.Kc);
    b.Ppad = m.P <= 16 ? 16 : m.P <= 32 ? 32 : 64;
    b.NM = b.Ppad / 16;
    if (m.bdsf) {
        // the detailed-balance build sums the trials' model ISF as a fold of
        // rounded products, which the fold kernel forms on the VALU in
        // register tiles: RL timeslices of one of 64 / NM row blocks per
        // lane (R = 64 RL / NM), the smallest RL whose blocks hold T, or
        // all but at most 4 timeslices, which the four chain waves then
        // fold one each
        b.MF = 0;
        for (int rl = 1; rl <= 8 && b.MF == 0 && T <= 128; rl *= 2) {
            const int R = 64 / b.NM * rl;
            if (R > 128) break;
            if (T <= R) {
                b.MF = 2 * rl;
                b.fold_rows = R;
            } else if (T - R <= 4) {
                b.MF = 2 * rl + 1;
                b.fold_rows = R + 4;
            }
        }
        if (b.MF == 0) {
            why = "batched detailed-balance (bdsf) runs take at most 128 timeslices";
            return DEAC_ERR_UNSUPPORTED;
        }
    }
    b.C = deac::kBatchChunkDoubles / b.Ppad;
    b.nchunk = (L + b.C - 1) / b.C;
    b.Lpad = b.nchunk * b.C;
    // the kernel reads one K tile (Lpad x 16 doubles) through a buffer
    // resource with 32-bit byte offsets
    if (static_cast<long>(b.Lpad) * 128 > INT32_MAX) {
        why = "batched runs take genomes of at most 16 million positions";
        return DEAC_ERR_UNSUPPORTED;
    }
    // the fold kernel reads kfold ([Lpad][fold_rows] doubles) the same way
    if (m.bdsf && static_cast<long>(b.Lpad) * b.fold_rows * 8 > INT32_MAX) {
        why = "batched detailed-balance (bdsf) runs take genomes of at most " +
              std::to_string(INT32_MAX / (8L * b.fold_rows)) + " positions at this number of timeslices";
        return DEAC_ERR_UNSUPPORTED;
    }
    b.Lw = b.Lpad / 32;
    b.ntt = T / 16;
    b.nrem = T % 16;
    // Up to four rows past the last full tile run on 4x4x4 MFMAs next to the
    // 16x16x4 tiles; more become one more 16-row tile padded with zero rows.
    if (b.nrem > deac::kBatchMaxRem) {
        b.ntt += 1;
        b.nrem = 0;
    }
    const int per_wave = (b.ntt + 3) / 4;
    b.NT = per_wave <= 1 ? 1 : per_wave <= 2 ? 2 : per_wave <= 4 ? 4 : 0;
    if (b.NT == 0 || T * b.Ppad > deac::kBatchRegionDoubles) {
        why = "batched runs take at most 260 timeslices";
        return DEAC_ERR_UNSUPPORTED;
    }
    if (b.NM == 4 && b.NT > 2) {
        // 4 member tiles x NT kernel tiles of accumulators per chain wave
        why = "batched runs with more than 32 members take at most 132 timeslices";
        return DEAC_ERR_UNSUPPORTED;
    }
    b.S = deac::kBatchMaskLanes / b.Ppad;
    b.Lsub = (L + 32 * b.S - 1) / (32 * b.S) * 32;
    b.ntm = (T + 15) / 16;
    b.NQ = deac::batch_mixed_nq(b.ntm, b.NM);
    return DEAC_OK;
}

// K in the mixed kernel's operand order (batch_kernels.hpp, BatchDesc::kmix),
// each row scaled by 2^-ex[t] with ex[t] the exponent of its largest
// magnitude (the split loop's scaling, ks_layout.hpp), and 2^ex[t] itself.
deac_status upload_mixed(deac_batch *h, BatchModel &b, int precision)
{
    if (b.kmix[precision]) return DEAC_OK;
    const deac::Model &m = b.m;
    const size_t T = m.T, L = m.Kc;
    std::vector<int> ex(T, 0);
    for (size_t t = 0; t < T; t++) {
        double mx = 0.0;
        for (size_t k = 0; k < L; k++) mx = std::fmax(mx, std::fabs(m.kernel[t * L + k]));
        ex[t] = mx > 0.0 ? std::ilogb(mx) : 0;
    }
    auto scaled = [&](size_t row, size_t pos) {
        return row < T && pos < L ? std::ldexp(m.kernel[row * L + pos], -ex[row]) : 0.0;
    };
    const size_t tile = static_cast<size_t>(b.Lpad) * 16;  // values per 16-row tile
    std::vector<uint32_t> img(static_cast<size_t>(b.ntm) * tile, 0);  // 4 bytes per value either way
    if (precision == DEAC_PRECISION_FP32) {
        // [tile][unit of 16 positions][lane][4]: row 16 tile + l % 16, position 16 u + 4 e + l / 16
        float *f = reinterpret_cast<float *>(img.data());
        for (int tt = 0; tt < b.ntm; tt++)
            for (int u = 0; u < b.Lpad / 16; u++)
                for (int l = 0; l < 64; l++)
                    for (int e = 0; e < 4; e++)
                        f[((static_cast<size_t>(tt) * (b.Lpad / 16) + u) * 64 + l) * 4 + e] = static_cast<float>(
                            scaled(16 * tt + (l & 15), 16 * static_cast<size_t>(u) + 4 * e + (l >> 4)));
    } else {
        // [tile][32-position block][lane][8 hi, 8 lo]: row 16 tile + l % 16, position 32 u + 8 (l / 16) + e
        uint16_t *q = reinterpret_cast<uint16_t *>(img.data());
        for (int tt = 0; tt < b.ntm; tt++)
            for (int u = 0; u < b.Lpad / 32; u++)
                for (int l = 0; l < 64; l++)
                    for (int e = 0; e < 8; e++) {
                        uint16_t hi, lo;
                        deac::bf16_split(scaled(16 * tt + (l & 15), 32 * static_cast<size_t>(u) + 8 * (l >> 4) + e),
                                         hi, lo);
                        const size_t o = ((static_cast<size_t>(tt) * (b.Lpad / 32) + u) * 64 + l) * 16 + e;
                        q[o] = hi;
                        q[o + 8] = lo;
                    }
    }
    uint32_t *dk = nullptr;
    deac_status st = h->alloc(dk, img.size());
    if (st != DEAC_OK) return st;
    if (hipMemcpy(dk, img.data(), img.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
        return DEAC_ERR_DEVICE;
    if (!b.kscale) {
        std::vector<double> sc(T);
        for (size_t t = 0; t < T; t++) sc[t] = std::ldexp(1.0, ex[t]);
        if ((st = h->alloc(b.kscale, T)) != DEAC_OK) return st;
        if (hipMemcpy(b.kscale, sc.data(), T * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
            return DEAC_ERR_DEVICE;
    }
    b.kmix[precision] = dk;
    return DEAC_OK;
}

deac_status upload_model(deac_batch *h, BatchModel &b)
{
    const deac::Model &m = b.m;
    const int T = static_cast<int>(m.T), L = static_cast<int>(m.Kc);
    const int nkb2 = b.Lpad / 8;
    std::vector<double> kt(static_cast<size_t>(std::max(b.ntt, 1)) * nkb2 * 64 * 2, -0.0);
    for (int t = 0; t < b.ntt; t++)
        for (int kb2 = 0; kb2 < nkb2; kb2++)
            for (int l = 0; l < 64; l++)
                for (int e = 0; e < 2; e++) {
                    const int tau = t * 16 + (l & 15), slot = (2 * kb2 + e) * 4 + (l >> 4);
                    double v = -0.0;
                    if (tau < T && slot < L) v = m.kernel[static_cast<size_t>(tau) * L + slot];
                    kt[((static_cast<size_t>(t) * nkb2 + kb2) * 64 + l) * 2 + e] = v;
                }
    // Rows ntt*16 .. T-1 for v_mfma_f64_4x4x4_4b: per 8 positions, 4 nrem
    // pairs (row r + nrem * position k of a 4-position block; the pair holds
    // the two blocks).  Lane l with l & 3 < nrem reads entry
    // (l & 3) + nrem (l >> 4): the A operand, the same for the four 4x4
    // blocks; the rows past nrem read as zeros (their chains are discarded).
    const int nr = std::max(b.nrem, 1);
    std::vector<double> kr(static_cast<size_t>(nkb2) * 4 * nr * 2, -0.0);
    for (int kb2 = 0; kb2 < nkb2; kb2++)
        for (int q = 0; q < 4; q++)
            for (int r = 0; r < b.nrem; r++)
                for (int e = 0; e < 2; e++) {
                    const int slot = (2 * kb2 + e) * 4 + q;
                    const double v = slot < L ? m.kernel[static_cast<size_t>(b.ntt * 16 + r) * L + slot] : -0.0;
                    kr[(static_cast<size_t>(kb2) * 4 * nr + r + nr * q) * 2 + e] = v;
                }
    // Jump tables: T^L by rows for the walk, T^(k*Lsub) by nibbles for the
    // mask sub-streams k = 1..S-1.
    std::vector<uint64_t> rows(256 * 4), tabs(static_cast<size_t>(std::max(b.S - 1, 1)) * 64 * 16 * 4);
    deac::matrix_rows(deac::transition_power(static_cast<uint64_t>(L)), rows.data());
    const deac::Gf2Matrix step = deac::transition_power(static_cast<uint64_t>(b.Lsub));
    deac::Gf2Matrix acc = step;
    for (int k = 1; k < b.S; k++) {
        deac::nibble_table(acc, tabs.data() + static_cast<size_t>(k - 1) * 64 * 16 * 4);
        acc = step * acc;
    }
    deac_status st = DEAC_OK;
    auto A = [&](auto *&p, size_t n) {
        if (st == DEAC_OK) st = h->alloc(p, n);
    };
    A(b.ktile, kt.size());
    A(b.krem, kr.size());
    A(b.obs, m.T);
    A(b.sig, m.T);
    A(b.walk_rows, rows.size());
    A(b.sub_tab, tabs.size());
    // moment weights (deac.cpp:628-668, 728-768) and the normalization
    // weights of per-generation renormalization (deac.cpp:1342-1352), zero
    // past L
    std::vector<double> mt;
    const bool renorm = m.normalize && m.renormalize;
    if (m.use_first || m.use_third || renorm) {
        mt.assign(3 * static_cast<size_t>(b.Lpad), 0.0);
        for (size_t k = 0; k < m.Kc; k++) {
            if (m.use_first) mt[k] = m.fterm[k];
            if (m.use_third) mt[b.Lpad + k] = m.tterm[k];
            if (renorm) mt[2 * static_cast<size_t>(b.Lpad) + k] = m.nterm[k];
        }
        A(b.mterm, mt.size());
    }
    if (st != DEAC_OK) return st;
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (st == DEAC_OK && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) st = DEAC_ERR_DEVICE;
    };
    up(b.ktile, kt.data(), kt.size() * sizeof(double));
    up(b.krem, kr.data(), kr.size() * sizeof(double));
    up(b.obs, m.isf.data(), m.T * sizeof(double));
    up(b.sig, m.err.data(), m.T * sizeof(double));
    up(b.walk_rows, rows.data(), rows.size() * sizeof(uint64_t));
    up(b.sub_tab, tabs.data(), tabs.size() * sizeof(uint64_t));
    if (b.mterm) up(b.mterm, mt.data(), mt.size() * sizeof(double));
    if (m.bdsf && st == DEAC_OK) {
        // [Lpad][fold_rows]: the timeslices of a position contiguous
        const int rw = b.fold_rows;
        std::vector<double> kf(static_cast<size_t>(b.Lpad) * rw, 0.0);
        for (int t = 0; t < T; t++)
            for (int k = 0; k < L; k++)
                kf[static_cast<size_t>(k) * rw + t] = m.kernel[static_cast<size_t>(t) * L + k];
        A(b.kfold, kf.size());
        if (m.use_nfm) A(b.nfm_dt, m.T);
        up(b.kfold, kf.data(), kf.size() * sizeof(double));
        if (b.nfm_dt) up(b.nfm_dt, m.nfm_dt.data(), m.T * sizeof(double));
    }
    return st;
}

// Runs fn(i) for i < n on up to 16 host threads.
template <class F>
void parallel_for(size_t n, F fn)
{
    const size_t nt = std::min<size_t>(n, std::clamp<unsigned>(std::thread::hardware_concurrency(), 1u, 16u));
    if (nt <= 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (size_t i = t; i < n; i += nt) fn(i);
        });
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

deac_status deac_batch_create(const deac_problem *problems, size_t count, deac_batch **out, const char **message)
{
    if (message) *message = nullptr;
    if (!out) return DEAC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!problems || count == 0 || count > (size_t(1) << 20)) {
        deac::last_message() = "a batch holds 1 to 1048576 problems";
        if (message) *message = deac::last_message().c_str();
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    for (size_t i = 0; i < count; i++) {
        std::string why;
        const deac_status v = deac::validate(problems[i], why);
        if (v != DEAC_OK) {
            deac::last_message() = "problem " + std::to_string(i) + ": " + why;
            if (message) *message = deac::last_message().c_str();
            return v;
        }
    }
    // Distinct problems, their models and kernel layouts (host only).
    std::unique_ptr<deac_batch> h(new deac_batch());
    std::map<std::string, int> seen;
    h->model_of.resize(count);
    for (size_t i = 0; i < count; i++) {
        const std::string key = problem_key(problems[i]);
        auto it = seen.find(key);
        if (it != seen.end()) {
            h->model_of[i] = it->second;
            continue;
        }
        auto b = std::make_unique<BatchModel>();
        b->m = deac::build_model(problems[i]);
        std::string why;
        const deac_status st = plan_model(*b, why);
        if (st != DEAC_OK) {
            deac::last_message() = "problem " + std::to_string(i) + ": " + why;
            if (message) *message = deac::last_message().c_str();
            return st;
        }
        seen.emplace(key, static_cast<int>(h->models.size()));
        h->model_of[i] = static_cast<int>(h->models.size());
        h->models.push_back(std::move(b));
    }

    {
        std::vector<std::string> shapes;
        for (size_t i = 0; i < count; i++) {
            const deac::Model &m = h->models[h->model_of[i]]->m;
            shapes.push_back(std::to_string(m.T) + "x" + std::to_string(m.G) + "x" + std::to_string(m.P) + "x" +
                             std::to_string(m.nsign));
        }
        std::sort(shapes.begin(), shapes.end());
        for (auto &x : shapes) h->signature += x + ";";
        std::lock_guard<std::mutex> lock(g_sched_lock);
        auto it = g_sched_cache.find(h->signature);
        if (it != g_sched_cache.end()) h->sched = it->second;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return DEAC_ERR_NO_DEVICE;
    }
    BATCH_HIP(hipGetDevice(&h->device));
    hipDeviceProp_t prop;
    BATCH_HIP(hipGetDeviceProperties(&prop, h->device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return DEAC_ERR_NO_DEVICE;
    h->cu_count = prop.multiProcessorCount;
    BATCH_HIP(deac::batch_scratch(h->device, &h->scratch));  // once per device, this device current
    BATCH_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    for (hipEvent_t *e : {&h->ev0, &h->ev1, &h->kev0, &h->kev1, &h->kev2, &h->kev3}) BATCH_HIP(hipEventCreate(e));
    BATCH_HIP(hipEventCreateWithFlags(&h->fork, hipEventDisableTiming));
    for (int g = 0; g < deac_batch::kGroupStreams; g++) {
        BATCH_HIP(hipStreamCreateWithFlags(&h->gstream[g], hipStreamNonBlocking));
        BATCH_HIP(hipEventCreateWithFlags(&h->gdone[g], hipEventDisableTiming));
    }
    for (auto &b : h->models) {
        const deac_status su = upload_model(h.get(), *b);
        if (su != DEAC_OK) return su;
    }

    // Per-instance state pools.
    size_t npop = 0, nmask = 0, nepar = 0;
    h->pop_off.resize(count);
    h->mask_off.resize(count);
    h->epar_off.resize(count);
    for (size_t i = 0; i < count; i++) {
        const BatchModel &b = *h->models[h->model_of[i]];
        h->pop_off[i] = npop;
        h->mask_off[i] = nmask;
        h->epar_off[i] = nepar;
        npop += b.m.P * static_cast<size_t>(b.Lpad);
        nmask += static_cast<size_t>(3) * b.Ppad * 2 * b.Lw;
        nepar += b.m.T * static_cast<size_t>(b.Ppad);
    }
    deac_status st = DEAC_OK;
    auto A = [&](auto *&p, size_t n) {
        if (st == DEAC_OK) st = h->alloc(p, n);
    };
    A(h->pop, npop);
    A(h->masks, nmask);
    A(h->fitness, count * deac::kBatchMaxP);
    A(h->cr_old, count * deac::kBatchMaxP * 2);
    A(h->f_old, count * deac::kBatchMaxP * 2);
    A(h->rng, count * 4);
    A(h->status, count * 2);
    A(h->accp, count * deac::kBatchMaxP);
    A(h->epar, nepar);
    A(h->descs, 3 * count);  // evolution launches, and the evaluation launches of mixed K.S before and after
    if (st != DEAC_OK) return st;
    // pinned host sides of the per-round descriptor upload and status read
    if (hipHostMalloc(reinterpret_cast<void **>(&h->hdescs), 3 * count * sizeof(deac::BatchDesc)) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&h->hstatus), 2 * count * sizeof(long)) != hipSuccess) {
        (void)hipGetLastError();
        return DEAC_ERR_OUT_OF_MEMORY;
    }
    *out = h.release();
    return DEAC_OK;
}

void deac_batch_destroy(deac_batch *h) { delete h; }

size_t deac_batch_size(const deac_batch *h) { return h ? h->model_of.size() : 0; }

deac_status deac_batch_last_profile(const deac_batch *h, deac_run_profile *profile)
{
    if (!h || !profile) return DEAC_ERR_INVALID_ARGUMENT;
    *profile = h->last;
    return DEAC_OK;
}

}  // extern "C"

namespace {

// deac_batch_run (resume false) and deac_batch_continue (resume true).
// stats_pool: device memory for the statistics histories (3 N doubles per
// tracking instance, packed), or null to allocate them here.
deac_status batch_run_impl(deac_batch *h, uint64_t *rng_states, const deac_run_options *options,
                           deac_result *results, bool resume, double *stats_pool)
{
    if (!h || !rng_states || !options || !results) return DEAC_ERR_INVALID_ARGUMENT;
    const size_t n = h->model_of.size();
    if (resume && h->reached.size() != n) {
        deac::last_message() = "no completed deac_batch_run to continue";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    for (size_t i = 0; i < n; i++) {
        if (options[i].number_of_generations < 2 || !results[i].dsf || !results[i].frequency) {
            deac::last_message() = "instance " + std::to_string(i) + ": number_of_generations must be at least 2 and the "
                              "dsf and frequency outputs are required";
            return DEAC_ERR_INVALID_ARGUMENT;
        }
        if (!resume && (rng_states[4 * i] | rng_states[4 * i + 1] | rng_states[4 * i + 2] | rng_states[4 * i + 3]) == 0) {
            // a fixed point of xoshiro256+: the donor draws would never end
            deac::last_message() = "instance " + std::to_string(i) + ": the rng state must not be all zero";
            return DEAC_ERR_INVALID_ARGUMENT;
        }
        const int prec = options[i].ks_precision;
        if (prec != DEAC_PRECISION_FP64 && prec != DEAC_PRECISION_FP32 && prec != DEAC_PRECISION_BF16X3) {
            deac::last_message() = "instance " + std::to_string(i) + ": ks_precision must be DEAC_PRECISION_FP64, "
                              "DEAC_PRECISION_FP32 or DEAC_PRECISION_BF16X3";
            return DEAC_ERR_INVALID_ARGUMENT;
        }
        if (options[i].split_loop) {
            deac::last_message() = "instance " + std::to_string(i) + ": the batch kernel has no split loop (run it "
                              "through deac_run)";
            return DEAC_ERR_UNSUPPORTED;
        }
        if (prec != DEAC_PRECISION_FP64 && h->models[h->model_of[i]]->NQ == 0) {
            deac::last_message() = "instance " + std::to_string(i) + ": no mixed K.S kernel for this shape";
            return DEAC_ERR_UNSUPPORTED;
        }
    }
    BATCH_HIP(hipSetDevice(h->device));
    hipStream_t s = h->stream;
    for (size_t i = 0; i < n; i++)
        if (options[i].ks_precision != DEAC_PRECISION_FP64) {
            const deac_status su = upload_mixed(h, *h->models[h->model_of[i]], options[i].ks_precision);
            if (su != DEAC_OK) return su;
        }

    // Generation base of each instance: 0 for a new run; where the last run
    // stopped for a continuation (its population, fitness, self-adapted CR
    // and F and random state are still on the device).
    std::vector<long> base(n, 0);
    if (resume) base = h->reached;
    std::vector<char> epar_ok = h->epar_valid;
    epar_ok.resize(n, 0);
    if (!resume) h->masks_at.assign(n, -1);
    h->masks_at.resize(n, -1);
    // a run may fail part way: until it completes, nothing it left counts
    std::vector<long> masks_at = h->masks_at;
    h->masks_at.assign(n, -1);
    h->reached.clear();  // set again when this call completes
    std::vector<uint64_t> rng_host(4 * n);
    if (!resume) {
        // Initial populations on the host (deac.cpp:430-603), each from its own
        // stream; the state after them is where the device walk starts.
        std::vector<double> pop_host(h->pop_off.back() + h->models[h->model_of[n - 1]]->m.P *
                                                             static_cast<size_t>(h->models[h->model_of[n - 1]]->Lpad),
                                     0.0);
        std::vector<double> cr_host(n * deac::kBatchMaxP * 2, 0.0), f_host(n * deac::kBatchMaxP * 2, 0.0);
        parallel_for(n, [&](size_t i) {
            const BatchModel &b = *h->models[h->model_of[i]];
            deac::Xoshiro rng{{rng_states[4 * i], rng_states[4 * i + 1], rng_states[4 * i + 2], rng_states[4 * i + 3]}};
            std::vector<double> pop;
            deac::initial_population(b.m, rng, pop);
            for (size_t m = 0; m < b.m.P; m++)
                std::memcpy(&pop_host[h->pop_off[i] + m * b.Lpad], &pop[m * b.m.Kc], b.m.Kc * sizeof(double));
            std::memcpy(&rng_host[4 * i], rng.s, sizeof rng.s);
            for (size_t m = 0; m < b.m.P; m++)
                for (int sg = 0; sg < 2; sg++) {
                    cr_host[(i * deac::kBatchMaxP + m) * 2 + sg] = b.m.cr0;  // deac.cpp:990-1032
                    f_host[(i * deac::kBatchMaxP + m) * 2 + sg] = b.m.f0;
                }
        });
        BATCH_HIP(hipMemcpyAsync(h->pop, pop_host.data(), pop_host.size() * sizeof(double), hipMemcpyHostToDevice, s));
        BATCH_HIP(hipMemcpyAsync(h->rng, rng_host.data(), rng_host.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
        BATCH_HIP(hipMemcpyAsync(h->cr_old, cr_host.data(), cr_host.size() * sizeof(double), hipMemcpyHostToDevice, s));
        BATCH_HIP(hipMemcpyAsync(h->f_old, f_host.data(), f_host.size() * sizeof(double), hipMemcpyHostToDevice, s));
        // the uploads read host memory that goes out of scope here
        BATCH_HIP(hipStreamSynchronize(s));
    }

    // Statistics buffers for the instances that track them.
    std::vector<double *> stats(n, nullptr);
    struct Freer {
        deac_batch *h;
        std::vector<double *> &v;
        bool own;
        ~Freer()
        {
            if (own)
                for (double *p : v)
                    if (p) h->release(p);
        }
    } freer{h, stats, stats_pool == nullptr};
    if (stats_pool) {
        size_t off = 0;
        for (size_t i = 0; i < n; i++)
            if (options[i].track_stats) {
                stats[i] = stats_pool + off;
                off += 3 * options[i].number_of_generations;
            }
    } else {
        for (size_t i = 0; i < n; i++)
            if (options[i].track_stats && h->alloc(stats[i], 3 * options[i].number_of_generations) != DEAC_OK)
                return DEAC_ERR_OUT_OF_MEMORY;
    }

    // Launch slices: every active instance evolves up to `slice` generations
    // per launch; stopped or finished instances leave the next launch.
    // Mixed K.S instances run slices of at most `refresh` generations, each
    // followed by an exact FP64 evaluation of the population, which resets
    // the parent model ISF the delta form builds on (and makes the final
    // fitness exact); one more such evaluation precedes their first slice.
    long slice = 4096, refresh = 1024;
    if (const char *e = std::getenv("DEAC_BATCH_SLICE")) slice = std::max(1L, std::atol(e));
    if (const char *e = std::getenv("DEAC_BATCH_REFRESH")) refresh = std::max(1L, std::atol(e));
    std::vector<long> next_gen(base), stopped_at(n, -1), n_abs(n);
    for (size_t i = 0; i < n; i++) n_abs[i] = base[i] + static_cast<long>(options[i].number_of_generations);
    std::vector<char> active(n, 1);
    deac::BatchDesc *const descs = h->hdescs;  // pinned: the upload of one round completes before the next
    deac_run_profile prof{};
    BATCH_HIP(hipEventRecord(h->ev0, s));
    BATCH_HIP(hipEventRecord(h->ev1, s));  // re-recorded after every round's status copy
    bool first = true;
    // Launch key: (precision, NT or NQ, NM, remainder rows (FP64) or moment constraints (mixed)).
    typedef std::tuple<int, int, int, int> Key;
    // bdsf instances evolve on the fold kernel (precision kFoldKey); the FP64
    // kernel evaluates their initial population (the reference's initial
    // contraction is the FMA chain, deac.cpp:860)
    constexpr int kFoldKey = 3;
    auto key_of = [&](size_t i, bool eval) -> Key {
        const BatchModel &b = *h->models[h->model_of[i]];
        const int prec = options[i].ks_precision;
        if (b.m.bdsf && !eval && prec == DEAC_PRECISION_FP64) return Key{kFoldKey, b.MF, b.NM, 0};
        if (prec == DEAC_PRECISION_FP64 || eval) return Key{DEAC_PRECISION_FP64, b.NT, b.NM, b.nrem > 0 ? 1 : 0};
        return Key{prec, b.NQ, b.NM, mixed_moments(b.m)};  // last: moment terms in the fitness
    };
    auto launch = [&](const Key &k, const deac::BatchDesc *dd, int cnt, hipStream_t st) {
        if (std::get<0>(k) == DEAC_PRECISION_FP64)
            return deac::launch_batch(dd, cnt, std::get<1>(k), std::get<2>(k), std::get<3>(k), st);
        if (std::get<0>(k) == kFoldKey) return deac::launch_batch_fold(dd, cnt, std::get<1>(k), std::get<2>(k), st);
        return deac::launch_batch_mixed(dd, cnt, std::get<0>(k), std::get<1>(k), std::get<2>(k), std::get<3>(k), st);
    };
    // Instances of instance i's problem that may share its workgroup (FP64;
    // populations that leave room in the 16 member rows of their kernel
    // shape), and whether adding i keeps a workgroup's populations within
    // the kernel's 32-bit byte offsets.
    // Automatic packing: a workgroup's time per generation grows ~6% per
    // stacked instance (its random-stream walk) while every workgroup
    // streams the whole kernel matrix from L2 each generation, so stacking
    // pays once the workgroups would fill most of the device: as many as
    // bring a launch round to one workgroup per CU, one more while that
    // leaves more than 3/4 of the CUs busy (measured at G = 4096, T = 65,
    // P = 4, 5, 8 and 32 to 1024 instances).
    const size_t n_cu = static_cast<size_t>(std::max(h->cu_count, 1));
    auto pack_limit = [&](size_t i, size_t n_active) {
        const BatchModel &b = *h->models[h->model_of[i]];
        if (options[i].ks_precision != DEAC_PRECISION_FP64 || h->sched.pack == 1 || b.Ppad != 16 || b.m.bdsf) return 1;
        int q = std::min<int>(deac::kBatchPackMax, 16 / static_cast<int>(b.m.P));
        if (h->sched.pack > 1) q = std::min(q, h->sched.pack);
        else {
            size_t a = (n_active + n_cu - 1) / n_cu;
            if ((n_active + a - 1) / a * 4 > 3 * n_cu) a++;
            q = std::min<int>(q, static_cast<int>(a));
        }
        return std::max(q, 1);
    };
    auto pack_span_ok = [&](const std::vector<size_t> &pk, size_t i) {
        const size_t rows = h->models[h->model_of[i]]->m.P * static_cast<size_t>(h->models[h->model_of[i]]->Lpad);
        size_t lo = h->pop_off[i], hi = h->pop_off[i] + rows;
        for (size_t j : pk) {
            lo = std::min(lo, h->pop_off[j]);
            hi = std::max(hi, h->pop_off[j] + rows);
        }
        return (hi - lo) * sizeof(double) <= UINT32_MAX;
    };
    for (;;) {
        // Group the active instances by kernel instantiation.
        std::map<Key, std::vector<size_t>> groups;
        size_t n_active = 0;
        for (size_t i = 0; i < n; i++)
            if (active[i]) {
                groups[key_of(i, false)].push_back(i);
                n_active++;
            }
        if (groups.empty()) break;
        // Order within a group (workgroups are dispatched in index order):
        // longest instances first -- per-generation cost ~ chunks x tiles
        // per wave -- or round robin over distinct problems.
        auto cost = [&](size_t i) {
            const BatchModel &b = *h->models[h->model_of[i]];
            return static_cast<long>(b.nchunk) * (b.NT * 4 + b.nrem);
        };
        for (auto &[key, members] : groups) {
            std::stable_sort(members.begin(), members.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
            if (h->sched.instance_order == 1) {
                std::map<int, std::vector<size_t>> per;
                std::vector<int> seen;
                for (size_t i : members) {
                    if (!per.count(h->model_of[i])) seen.push_back(h->model_of[i]);
                    per[h->model_of[i]].push_back(i);
                }
                members.clear();
                for (size_t r = 0;; r++) {
                    bool any = false;
                    for (int mdl : seen)
                        if (r < per[mdl].size()) {
                            members.push_back(per[mdl][r]);
                            any = true;
                        }
                    if (!any) break;
                }
            }
        }
        // Order of the groups' launches.
        std::vector<std::pair<Key, std::vector<size_t> *>> glist;
        for (auto &[key, members] : groups) glist.push_back({key, &members});
        if (h->sched.group_order != 2) {
            auto gcost = [&](const std::vector<size_t> *v) {
                long c = 0;
                for (size_t i : *v) c += cost(i);
                return c;
            };
            std::stable_sort(glist.begin(), glist.end(), [&](const auto &a, const auto &b) {
                return h->sched.group_order == 0 ? gcost(a.second) > gcost(b.second) : gcost(a.second) < gcost(b.second);
            });
        }
        size_t slot = 0;
        std::vector<std::pair<Key, std::pair<size_t, size_t>>> launches, evals_before, evals_after;
        std::vector<size_t> order;                       // lead instance of each slot
        std::vector<std::pair<size_t, size_t>> placed;   // (instance, slot) of this round's evolution launches
        std::vector<size_t> mixed_slots;  // evolution slots of instances not on the FP64 kernel (mixed, bdsf)
        for (auto &[key, mp] : glist) {
            const std::vector<size_t> &members = *mp;
            // Workgroups: FP64 instances of one problem with small populations
            // at the same generation share one, up to the member rows their
            // kernel shape holds anyway; the rest run one per workgroup.
            std::vector<std::vector<size_t>> packs;
            std::map<std::pair<int, long>, size_t> open;
            for (size_t i : members) {
                const int q = pack_limit(i, n_active);
                if (q > 1) {
                    auto it = open.find({h->model_of[i], next_gen[i]});
                    if (it != open.end() && packs[it->second].size() < static_cast<size_t>(q) &&
                        pack_span_ok(packs[it->second], i)) {
                        packs[it->second].push_back(i);
                        continue;
                    }
                    open[{h->model_of[i], next_gen[i]}] = packs.size();
                }
                packs.push_back({i});
            }
            const size_t begin = slot;
            for (const std::vector<size_t> &pk : packs) {
                const size_t i = pk[0];
                const BatchModel &b = *h->models[h->model_of[i]];
                const deac::Model &m = b.m;
                const int prec = options[i].ks_precision;
                deac::BatchDesc &d = descs[slot];
                d = deac::BatchDesc{};
                d.ktile = b.ktile;
                d.krem = b.krem;
                d.kmix = prec == DEAC_PRECISION_FP64 ? nullptr : b.kmix[prec];
                d.kscale = b.kscale;
                d.obs = b.obs;
                d.sig = b.sig;
                d.mterm = b.mterm;
                d.use_first = m.use_first ? 1 : 0;
                d.use_third = m.use_third ? 1 : 0;
                d.first_moment = m.first_moment;
                d.third_moment = m.third_moment;
                d.third_moment_error = m.third_moment_error;
                d.kfold = b.kfold;
                d.fold_rows = b.fold_rows;
                d.use_nfm = m.use_nfm ? 1 : 0;
                d.nfm_dt = b.nfm_dt;
                d.nfm_value = m.nfm_value;
                d.nfm_error = m.nfm_error;
                d.renorm = (m.normalize && m.renormalize) ? 1 : 0;
                d.zeroth = m.zeroth_moment;
                d.walk_rows = b.walk_rows;
                d.sub_tab = b.sub_tab;
                size_t pop_lo = h->pop_off[i];
                for (size_t j : pk) pop_lo = std::min(pop_lo, h->pop_off[j]);
                d.pop = h->pop + pop_lo;
                d.masks = h->masks + h->mask_off[i];
                d.epar = prec == DEAC_PRECISION_FP64 ? nullptr : h->epar + h->epar_off[i];
                d.gen0 = next_gen[i];
                const long len = prec == DEAC_PRECISION_FP64 ? slice : std::min(slice, refresh);
                d.gen_end = std::min<long>(next_gen[i] + len, n_abs[i] - 1);
                d.Q = static_cast<int>(pk.size());
                for (int q = 0; q < d.Q; q++) {
                    const size_t j = pk[q];
                    d.pop_row0[q] = static_cast<uint32_t>((h->pop_off[j] - pop_lo) * sizeof(double));
                    d.fitness[q] = h->fitness + j * deac::kBatchMaxP;
                    d.cr_old[q] = h->cr_old + j * deac::kBatchMaxP * 2;
                    d.f_old[q] = h->f_old + j * deac::kBatchMaxP * 2;
                    d.rng[q] = h->rng + 4 * j;
                    d.stats[q] = stats[j];
                    d.status[q] = h->status + 2 * j;
                    d.accp[q] = h->accp + j * deac::kBatchMaxP;
                    d.stats_len[q] = static_cast<long>(options[j].number_of_generations);
                    d.stats_base[q] = base[j];
                    d.stop[q] = options[j].stop_minimum_fitness;
                    // the launch ends with its shortest instance; the others go on in the next
                    d.gen_end = std::min<long>(d.gen_end, n_abs[j] - 1);
                    placed.push_back({j, slot});
                }
                d.tau_cr = m.tau_cr;
                d.tau_f = m.tau_f;
                d.Pi = static_cast<int>(m.P);
                d.P = d.Q * d.Pi;
                d.mod_magic = ~UINT64_C(0) / m.P;
                d.init = (first && !resume && std::get<0>(key) == DEAC_PRECISION_FP64) ? 1 : 0;
                // an instance alone in its workgroup whose last launch ended here
                // resumes its masks instead of drawing them again
                d.resume = (!d.init && pk.size() == 1 && masks_at[i] == next_gen[i]) ? 1 : 0;
                d.mode = 0;
                d.T = static_cast<int>(m.T);
                d.G = static_cast<int>(m.G);
                d.nsign = m.nsign;
                d.L = static_cast<int>(m.Kc);
                d.Lpad = b.Lpad;
                d.Lw = b.Lw;
                d.nchunk = b.nchunk;
                d.ntt = b.ntt;
                d.nrem = b.nrem;
                d.ntm = b.ntm;
                d.S = b.S;
                d.Lsub = b.Lsub;
                d.allow_neg = m.allow_negative ? 1 : 0;
                if (std::get<0>(key) != DEAC_PRECISION_FP64) mixed_slots.push_back(slot);
                order.push_back(i);
                slot++;
            }
            launches.push_back({key, {begin, slot - begin}});
        }
        // Evaluation launches (FP64 kernel, mode 1) of the mixed instances,
        // grouped by the FP64 kernel shape: before the first slice (initial
        // fitness and epar; skipped when a continuation finds epar current)
        // and after every slice.
        {
            std::map<Key, std::vector<size_t>> eb, ea;
            for (size_t k : mixed_slots) {
                const size_t i = order[k];
                if (h->models[h->model_of[i]]->m.bdsf && options[i].ks_precision == DEAC_PRECISION_FP64) {
                    // exact folds: only the initial population's evaluation
                    if (first && !resume) eb[key_of(i, true)].push_back(k);
                    continue;
                }
                ea[key_of(i, true)].push_back(k);
                if (first && !(resume && epar_ok[i])) eb[key_of(i, true)].push_back(k);
            }
            for (auto *grp : {&eb, &ea})
                for (auto &[key, ks] : *grp) {
                    const size_t begin = slot;
                    for (size_t k : ks) {
                        deac::BatchDesc &e = descs[slot++];
                        e = descs[k];
                        e.mode = 1;
                        e.init = 1;
                    }
                    (grp == &eb ? evals_before : evals_after).push_back({key, {begin, slot - begin}});
                }
        }
        BATCH_HIP(hipMemcpyAsync(h->descs, descs, slot * sizeof(deac::BatchDesc), hipMemcpyHostToDevice, s));
        // events: kev0 | evaluations before | kev1 | evolution | kev2 | evaluations after | kev3
        BATCH_HIP(hipEventRecord(h->kev0, s));
        for (auto &[key, range] : evals_before)
            BATCH_HIP(launch(key, h->descs + range.first, static_cast<int>(range.second), s));
        BATCH_HIP(hipEventRecord(h->kev1, s));
        const int nstreams = std::clamp(h->sched.group_streams, 1, deac_batch::kGroupStreams);
        if (launches.size() == 1 || nstreams == 1) {
            for (auto &[key, range] : launches)
                BATCH_HIP(launch(key, h->descs + range.first, static_cast<int>(range.second), s));
        } else {
            // shapes side by side, on the group streams
            BATCH_HIP(hipEventRecord(h->fork, s));
            for (int g = 0; g < nstreams; g++) BATCH_HIP(hipStreamWaitEvent(h->gstream[g], h->fork, 0));
            for (size_t k = 0; k < launches.size(); k++) {
                auto &[key, range] = launches[k];
                BATCH_HIP(launch(key, h->descs + range.first, static_cast<int>(range.second), h->gstream[k % nstreams]));
            }
            for (int g = 0; g < nstreams; g++) {
                BATCH_HIP(hipEventRecord(h->gdone[g], h->gstream[g]));
                BATCH_HIP(hipStreamWaitEvent(s, h->gdone[g], 0));
            }
        }
        BATCH_HIP(hipEventRecord(h->kev2, s));
        for (auto &[key, range] : evals_after)
            BATCH_HIP(launch(key, h->descs + range.first, static_cast<int>(range.second), s));
        BATCH_HIP(hipEventRecord(h->kev3, s));
        long *const st = h->hstatus;
        BATCH_HIP(hipMemcpyAsync(st, h->status, 2 * n * sizeof(long), hipMemcpyDeviceToHost, s));
        // the device work of the run ends here when this round is the last
        BATCH_HIP(hipEventRecord(h->ev1, s));
        BATCH_HIP(hipStreamSynchronize(s));
        float ms_before = 0.0f, ms_evolve = 0.0f, ms_after = 0.0f;
        BATCH_HIP(hipEventElapsedTime(&ms_before, h->kev0, h->kev1));
        BATCH_HIP(hipEventElapsedTime(&ms_evolve, h->kev1, h->kev2));
        BATCH_HIP(hipEventElapsedTime(&ms_after, h->kev2, h->kev3));
        prof.evolve_ms += ms_evolve;
        prof.evolve_launches += launches.size();
        for (auto &[key, range] : launches) prof.evolve_workgroups += range.second;
        prof.eval_ms += ms_before + ms_after;
        prof.eval_launches += evals_before.size() + evals_after.size();
        for (auto [i, k] : placed) {
            const long reached = st[2 * i + 1];
            // the kernel ends an instance inside its slice: anything else is
            // a device-side failure, not a result
            if (reached < descs[k].gen0 || reached > descs[k].gen_end || (st[2 * i] != 0 && st[2 * i] != 1))
                return DEAC_ERR_DEVICE;
            prof.generations += static_cast<size_t>(reached - next_gen[i]);
            // a launch that evolved a generation alone in its workgroup drew the
            // masks of the generation it ended at
            masks_at[i] = (descs[k].Q == 1 && reached > descs[k].gen0) ? reached : -1;
            next_gen[i] = reached;
            if (st[2 * i]) {
                stopped_at[i] = reached;
                active[i] = 0;
            } else if (reached >= n_abs[i] - 1) {
                active[i] = 0;
            }
        }
        first = false;
    }
    BATCH_HIP(hipEventSynchronize(h->ev1));
    float loop = 0.0f;
    BATCH_HIP(hipEventElapsedTime(&loop, h->ev0, h->ev1));
    prof.loop_ms = loop;
    h->last = prof;

    // Epilogue per instance (deac.cpp:1576-1685), as deac_run.
    std::vector<double> fit_all(n * deac::kBatchMaxP);
    BATCH_HIP(hipMemcpy(fit_all.data(), h->fitness, fit_all.size() * sizeof(double), hipMemcpyDeviceToHost));
    BATCH_HIP(hipMemcpy(rng_host.data(), h->rng, rng_host.size() * sizeof(uint64_t), hipMemcpyDeviceToHost));
    std::vector<double> row;
    for (size_t i = 0; i < n; i++) {
        const BatchModel &b = *h->models[h->model_of[i]];
        const deac::Model &m = b.m;
        const size_t P = m.P, N = options[i].number_of_generations;
        deac_result *r = &results[i];
        const double *fit = &fit_all[i * deac::kBatchMaxP];
        size_t best = 0;
        double minimum = fit[0];
        for (size_t k = 0; k < P; k++)
            if (fit[k] < minimum) {
                minimum = fit[k];
                best = k;
            }
        row.resize(m.Kc);
        BATCH_HIP(hipMemcpy(row.data(), h->pop + h->pop_off[i] + best * b.Lpad, m.Kc * sizeof(double),
                            hipMemcpyDeviceToHost));
        deac::output_transform(m, row.data(), r->dsf, r->frequency);
        size_t generation = stopped_at[i] >= 0 ? static_cast<size_t>(stopped_at[i] - base[i]) : N - 2;
        if (stats[i]) {
            std::vector<double> hs(3 * N, 0.0);
            const size_t k = generation + 1;
            BATCH_HIP(hipMemcpy(hs.data(), stats[i], 3 * N * sizeof(double), hipMemcpyDeviceToHost));
            if (r->stats_mean) std::memcpy(r->stats_mean, hs.data(), k * sizeof(double));
            if (r->stats_minimum) std::memcpy(r->stats_minimum, hs.data() + N, k * sizeof(double));
            if (r->stats_squared_mean) std::memcpy(r->stats_squared_mean, hs.data() + 2 * N, k * sizeof(double));
        }
        if (generation + 2 == N) {  // deac.cpp:1677-1685
            generation += 1;
            if (stats[i]) {
                if (r->stats_mean) r->stats_mean[generation] = deac::mean_of(fit, P);
                if (r->stats_minimum) r->stats_minimum[generation] = minimum;
                if (r->stats_squared_mean) r->stats_squared_mean[generation] = deac::squared_mean_of(fit, P);
            }
        }
        r->generation = generation;
        r->minimum_fitness = minimum;
        r->best_index = best;
        if (r->final_fitness) std::memcpy(r->final_fitness, fit, P * sizeof(double));
        std::memcpy(&rng_states[4 * i], &rng_host[4 * i], 4 * sizeof(uint64_t));
    }
    h->reached = next_gen;
    h->masks_at = masks_at;
    h->epar_valid.assign(n, 0);
    for (size_t i = 0; i < n; i++) h->epar_valid[i] = options[i].ks_precision != DEAC_PRECISION_FP64;
    return DEAC_OK;
}

}  // namespace

extern "C" {

deac_status deac_batch_run(deac_batch *h, uint64_t *rng_states, const deac_run_options *options, deac_result *results)
{
    return batch_run_impl(h, rng_states, options, results, false, nullptr);
}

deac_status deac_batch_continue(deac_batch *h, uint64_t *rng_states, const deac_run_options *options,
                                deac_result *results)
{
    return batch_run_impl(h, rng_states, options, results, true, nullptr);
}

deac_status deac_batch_run_noalloc(deac_batch *h, uint64_t *rng_states, const deac_run_options *options,
                                   deac_result *results)
{
    if (!h || !rng_states || !options || !results) return DEAC_ERR_INVALID_ARGUMENT;
    const size_t n = h->model_of.size();
    size_t need = 0;
    for (size_t i = 0; i < n; i++) {
        if (options[i].track_stats) need += 3 * options[i].number_of_generations;
        const int prec = options[i].ks_precision;
        if ((prec == DEAC_PRECISION_FP32 || prec == DEAC_PRECISION_BF16X3) && !h->models[h->model_of[i]]->kmix[prec]) {
            deac::last_message() = "instance " + std::to_string(i) + ": the mixed K.S tiles of this precision are "
                                   "built (allocated) by the first deac_batch_run at that precision";
            return DEAC_ERR_UNSUPPORTED;
        }
    }
    if (need > deac::kBatchScratchDoubles) {
        deac::last_message() = "the statistics of this call need " + std::to_string(need) +
                               " doubles; deac_batch_run_noalloc holds at most " +
                               std::to_string(deac::kBatchScratchDoubles);
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    return batch_run_impl(h, rng_states, options, results, false, h->scratch);
}

deac_status deac_batch_role_cycles(int precision, unsigned long long *cycles, int reset)
{
    if (!cycles) return DEAC_ERR_INVALID_ARGUMENT;
    const hipError_t e = precision == DEAC_PRECISION_FP64 ? deac::batch_role_cycles_fp64(cycles, reset)
                                                          : deac::batch_role_cycles_mixed(cycles, reset);
    if (e == hipErrorNotSupported) {
        (void)hipGetLastError();
        deac::last_message() = "the library was built without role timers (make ROLE_TIMERS=1)";
        return DEAC_ERR_UNSUPPORTED;
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return DEAC_ERR_DEVICE;
    }
    return DEAC_OK;
}

deac_status deac_batch_set_schedule(deac_batch *h, const deac_batch_schedule *sc)
{
    if (!h || !sc || sc->group_streams < 1 || sc->group_streams > deac_batch::kGroupStreams || sc->group_order < 0 ||
        sc->group_order > 2 || sc->instance_order < 0 || sc->instance_order > 1 || sc->pack < 0 ||
        sc->pack > deac::kBatchPackMax)
        return DEAC_ERR_INVALID_ARGUMENT;
    h->sched = *sc;
    return DEAC_OK;
}

deac_status deac_batch_get_schedule(const deac_batch *h, deac_batch_schedule *sc)
{
    if (!h || !sc) return DEAC_ERR_INVALID_ARGUMENT;
    *sc = h->sched;
    return DEAC_OK;
}

deac_status deac_batch_tune(deac_batch *h, const uint64_t *rng_states, const deac_run_options *options,
                            size_t generations, deac_batch_schedule *chosen)
{
    if (!h || !rng_states || !options || generations < 1) return DEAC_ERR_INVALID_ARGUMENT;
    const size_t n = h->model_of.size();
    // scratch outputs and a bounded run of every instance
    std::vector<deac_run_options> opts(options, options + n);
    std::vector<std::vector<double>> out(2 * n);
    std::vector<deac_result> res(n);
    for (size_t i = 0; i < n; i++) {
        opts[i].number_of_generations = generations + 1;
        opts[i].track_stats = 0;
        opts[i].stop_minimum_fitness = -1.0;  // every candidate runs every generation
        const size_t len = h->models[h->model_of[i]]->m.nsign == 2 ? 2 * h->models[h->model_of[i]]->m.G - 1
                                                                   : h->models[h->model_of[i]]->m.G;
        out[2 * i].assign(len, 0.0);
        out[2 * i + 1].assign(len, 0.0);
        res[i] = deac_result{};
        res[i].dsf = out[2 * i].data();
        res[i].frequency = out[2 * i + 1].data();
    }
    // candidates: every (streams, group order, instance order), less where
    // a knob has nothing to order
    std::map<std::tuple<int, int, int, int>, int> shapes;
    for (size_t i = 0; i < n; i++) {
        const BatchModel &b = *h->models[h->model_of[i]];
        if (b.m.bdsf && options[i].ks_precision == DEAC_PRECISION_FP64)
            shapes[{3, b.MF, b.NM, 0}] = 1;  // the fold kernel
        else
            shapes[{options[i].ks_precision, options[i].ks_precision == DEAC_PRECISION_FP64 ? b.NT : b.NQ, b.NM,
                    options[i].ks_precision == DEAC_PRECISION_FP64 ? (b.nrem > 0 ? 1 : 0) : mixed_moments(b.m)}] = 1;
    }
    const bool several_groups = shapes.size() > 1, several_models = h->models.size() > 1;
    std::vector<deac_batch_schedule> cand;
    for (int st = 1; st <= (several_groups ? deac_batch::kGroupStreams : 1); st++)
        for (int go = 0; go <= (several_groups ? 2 : 0); go++)
            for (int io = 0; io <= (several_models ? 1 : 0); io++)
                cand.push_back(deac_batch_schedule{st, go, io, 0.0, h->sched.pack});
    const deac_batch_schedule keep = h->sched;
    std::vector<uint64_t> states(rng_states, rng_states + 4 * n);
    // a warm-up run with the current schedule, then two timed continuations
    // of it per candidate (no host set-up between them)
    deac_status st = batch_run_impl(h, states.data(), opts.data(), res.data(), false, nullptr);
    if (st != DEAC_OK) return st;
    double best_ms = 1e300;
    deac_batch_schedule best = keep;
    for (const deac_batch_schedule &c : cand) {
        h->sched = c;
        double ms = 1e300;
        for (int rep = 0; rep < 2; rep++) {
            st = batch_run_impl(h, states.data(), opts.data(), res.data(), true, nullptr);
            if (st != DEAC_OK) {
                h->sched = keep;
                return st;
            }
            ms = std::min(ms, h->last.loop_ms);
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    best.ms_per_generation = best_ms / static_cast<double>(generations);
    h->sched = best;
    h->reached.clear();  // the tuning runs are not a run to continue
    {
        std::lock_guard<std::mutex> lock(g_sched_lock);
        g_sched_cache[h->signature] = best;
    }
    if (chosen) *chosen = best;
    return DEAC_OK;
}


}  // extern "C"
