// C ABI of the DEAC engine (include/deac_engine.h): device memory, uploads,
// the batch schedule of the RNG walk / mask / evolution kernels, and the
// host-side epilogue (best member, output transform, final statistics).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "deac_engine.h"
#include "host_model.hpp"
#include "kernels.hpp"
#include "ks_layout.hpp"
#include "precision_kernels.hpp"
#include "split_kernels.hpp"

namespace {


struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct deac_handle {
    deac::Model m;
    int device = 0;
    hipStream_t stream = nullptr;       // evolution
    hipStream_t draw_stream = nullptr;  // RNG walk, batches ahead of the evolution
    int MG = 0, TB = 0, stride = 0, nwg = 0, chunk = 0, batch = 0, maxwg = 0;
    bool overlap = false;               // walk beside the evolution (a CU is left free for it)
    std::vector<void *> allocations;
    double *kmat = nullptr, *nterm = nullptr, *fterm = nullptr, *tterm = nullptr, *obs = nullptr, *sig = nullptr;
    double *nfm_dt = nullptr;           // -1st-moment weights over tau (bdsf), or null
    double *pop0 = nullptr, *pop1 = nullptr, *fitness = nullptr, *cr_old = nullptr, *f_old = nullptr;
    double *emod = nullptr, *mom = nullptr;
    int *cur = nullptr;
    // Two buffer sets, used by batches of even and odd index: grid barrier
    // counter ([0]) and status ([4..6]) per set, and the batch's draws.
    unsigned *sync = nullptr;           // [2][8]
    long *iter = nullptr;               // [2][2]: split loop generation counters, per buffer set
    int *host_status = nullptr;         // [3][4], pinned: status of the sets, then of the initial fitness
    deac::MemberCtl *ctl[2] = {};
    uint64_t *raw[2] = {}, *gen_start[2] = {};
    uint64_t *state = nullptr, *skip_rows = nullptr, *lane_tab = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // per set: walk start, walk end, mask start, mask end, evolve start,
    // evolve end (draws free), status copied
    hipEvent_t kev[2][7] = {};
    deac_run_profile last{};
    // Split loop: contraction operands of the precision it last ran at.
    struct Split {
        int precision = -1;
        deac::KsLayout y;
        void *A = nullptr, *Al = nullptr;  // kernel matrix (BF16X3: hi, lo)
        void *B = nullptr, *Bl = nullptr;  // genome operand
        float *part = nullptr;             // FP32 genome slices
        int *ex = nullptr;                 // per-row exponents
    } split;

    ~deac_handle()
    {
        if (stream) (void)hipStreamSynchronize(stream);
        if (draw_stream) (void)hipStreamSynchronize(draw_stream);
        for (void *p : allocations) (void)hipFree(p);
        if (host_status) (void)hipHostFree(host_status);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (auto &set : kev)
            for (hipEvent_t e : set)
                if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
        if (draw_stream) (void)hipStreamDestroy(draw_stream);
    }

    template <class T>
    deac_status alloc(T *&out, size_t count)
    {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return DEAC_ERR_OUT_OF_MEMORY;
        allocations.push_back(p);
        out = static_cast<T *>(p);
        return DEAC_OK;
    }

    void release(void *&p)
    {
        if (!p) return;
        allocations.erase(std::find(allocations.begin(), allocations.end(), p));
        (void)hipFree(p);
        p = nullptr;
    }
};

namespace {

#define DEAC_TRY(expr)                         \
    do {                                       \
        const deac_status st_ = (expr);        \
        if (st_ != DEAC_OK) return st_;        \
    } while (0)
#define DEAC_HIP(expr)                         \
    do {                                       \
        if ((expr) != hipSuccess) {            \
            (void)hipGetLastError();           \
            return DEAC_ERR_DEVICE;            \
        }                                      \
    } while (0)

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Device population layout: member-major rows of `stride` doubles, rows for
// MG*4 members, zero padded.
void pad_rows(const std::vector<double> &pop, size_t P, size_t Kc, int MG, int stride, std::vector<double> &out)
{
    out.assign(static_cast<size_t>(MG) * deac::kChainRows * stride, 0.0);
    for (size_t i = 0; i < P; i++) std::memcpy(&out[i * stride], &pop[i * Kc], Kc * sizeof(double));
}

// Contraction operands of the split loop at `precision` (kept until another
// precision is asked for): the kernel matrix in the contraction's lane order,
// the genome operand and slice buffers for the population.
deac_status prepare_split(deac_handle *h, int precision)
{
    deac_handle::Split &sp = h->split;
    if (sp.precision == precision) return DEAC_OK;
    for (void **p : {&sp.A, &sp.Al, &sp.B, &sp.Bl, reinterpret_cast<void **>(&sp.part),
                     reinterpret_cast<void **>(&sp.ex)})
        h->release(*p);
    sp.precision = -1;
    const deac::Model &m = h->m;
    sp.y = deac::ks_layout(m, precision, m.P);
    const deac::KsLayout &y = sp.y;
    auto up = [](void *dst, const void *src, size_t bytes) {
        return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess ? DEAC_OK : DEAC_ERR_DEVICE;
    };
    if (precision == DEAC_PRECISION_FP64) {
        std::vector<double> A;
        deac::ks_kernel_fp64(m, y, A);
        double *a = nullptr, *b = nullptr;
        DEAC_TRY(h->alloc(a, A.size()));
        sp.A = a;
        DEAC_TRY(h->alloc(b, y.b_count()));
        sp.B = b;
        DEAC_TRY(up(a, A.data(), A.size() * sizeof(double)));
    } else if (precision == DEAC_PRECISION_FP32) {
        std::vector<float> A;
        deac::ks_kernel_fp32(m, y, A);
        float *a = nullptr, *b = nullptr;
        DEAC_TRY(h->alloc(a, A.size()));
        sp.A = a;
        DEAC_TRY(h->alloc(b, y.b_count()));
        sp.B = b;
        DEAC_TRY(up(a, A.data(), A.size() * sizeof(float)));
    } else {
        std::vector<uint16_t> Ah, Al;
        deac::ks_kernel_bf16x3(m, y, Ah, Al);
        uint16_t *ah = nullptr, *al = nullptr, *bh = nullptr, *bl = nullptr;
        DEAC_TRY(h->alloc(ah, Ah.size()));
        sp.A = ah;
        DEAC_TRY(h->alloc(al, Al.size()));
        sp.Al = al;
        DEAC_TRY(h->alloc(bh, y.b_count()));
        sp.B = bh;
        DEAC_TRY(h->alloc(bl, y.b_count()));
        sp.Bl = bl;
        DEAC_TRY(up(ah, Ah.data(), Ah.size() * sizeof(uint16_t)));
        DEAC_TRY(up(al, Al.data(), Al.size() * sizeof(uint16_t)));
    }
    if (precision != DEAC_PRECISION_FP64) {
        DEAC_TRY(h->alloc(sp.part, static_cast<size_t>(y.slices) * y.mpad() * y.npad()));
        DEAC_TRY(h->alloc(sp.ex, y.T));
        DEAC_TRY(up(sp.ex, y.ex.data(), y.T * sizeof(int)));
    }
    sp.precision = precision;
    return DEAC_OK;
}

}  // namespace

extern "C" {

const char *deac_status_string(deac_status s)
{
    switch (s) {
    case DEAC_OK: return "success";
    case DEAC_ERR_INVALID_ARGUMENT: return "invalid argument";
    case DEAC_ERR_NO_DEVICE: return "no usable gfx950 device";
    case DEAC_ERR_OUT_OF_MEMORY: return "out of memory";
    case DEAC_ERR_DEVICE: return "device error";
    case DEAC_ERR_UNSUPPORTED: return "unsupported";
    }
    return "unknown status";
}

const char *deac_last_error_message(void) { return deac::last_message().c_str(); }

size_t deac_output_length(const deac_problem *p)
{
    if (!p) return 0;
    return (p->zero_temperature || p->spectra_type == DEAC_BDSF) ? p->genome_size : 2 * p->genome_size - 1;
}

void deac_seed_state(uint64_t seed, uint64_t state[4]) { deac::seed_state(seed, state); }

void deac_rng_advance(uint64_t state[4], uint64_t n)
{
    const deac::Gf2Matrix m = deac::transition_power(n);
    uint64_t out[4];
    m.apply(state, out);
    std::memcpy(state, out, sizeof out);
}

deac_status deac_frequency_grid(size_t G, double omega_max, double *f)
{
    if (G < 2 || !f) return DEAC_ERR_INVALID_ARGUMENT;
    const double d = omega_max / static_cast<double>(G - 1);  // deac.cpp:2121-2125
    for (size_t i = 0; i < G; i++) f[i] = static_cast<double>(i) * d;
    return DEAC_OK;
}

deac_status deac_create(const deac_problem *p, deac_handle **out, const char **message)
{
    if (message) *message = nullptr;
    if (!p || !out) return DEAC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    const deac_status v = deac::validate(*p, deac::last_message());
    if (v != DEAC_OK) {
        if (message) *message = deac::last_message().c_str();
        return v;
    }

    int device = 0, count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return DEAC_ERR_NO_DEVICE;
    }
    DEAC_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    DEAC_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return DEAC_ERR_NO_DEVICE;

    deac_handle *h = new deac_handle();
    h->m = deac::build_model(*p);
    h->device = device;
    const deac::Model &m = h->m;
    auto fail = [&](deac_status s) {
        delete h;
        return s;
    };
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&h->draw_stream, hipStreamNonBlocking) != hipSuccess)
        return fail(DEAC_ERR_DEVICE);
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) return fail(DEAC_ERR_DEVICE);
    for (auto &set : h->kev)
        for (hipEvent_t &e : set)
            if (hipEventCreate(&e) != hipSuccess) return fail(DEAC_ERR_DEVICE);
    if (hipHostMalloc(reinterpret_cast<void **>(&h->host_status), 12 * sizeof(int)) != hipSuccess) {
        h->host_status = nullptr;
        return fail(DEAC_ERR_OUT_OF_MEMORY);
    }

    const int T = static_cast<int>(m.T), P = static_cast<int>(m.P), Kc = static_cast<int>(m.Kc);
    h->MG = (P + deac::kChainRows - 1) / deac::kChainRows;
    h->TB = (T + deac::kChainTau - 1) / deac::kChainTau;
    h->stride = round_up(Kc, deac::kChunk);
    h->chunk = (Kc + 63) / 64;
    // RNG batch: generations whose mask draws are buffered at once (<= 256 MiB
    // per buffer set).
    const size_t per_gen = static_cast<size_t>(P) * Kc * sizeof(uint64_t);
    h->batch = static_cast<int>(std::clamp<size_t>((size_t(256) << 20) / per_gen, 1, 1024));

    const int maxwg = deac::evolve_max_workgroups(device);
    if (maxwg <= 0) return fail(DEAC_ERR_DEVICE);
    h->maxwg = maxwg;
    // At least one workgroup per chain task; beyond that, enough for the
    // trial and selection phases (one per 512 trial items, 32 to 128: more
    // cost more at every grid barrier than they save).
    const size_t trial_items = static_cast<size_t>(P) * h->stride / 2;
    int nwg = std::max(h->MG * h->TB, static_cast<int>(std::clamp<size_t>(trial_items / 512, 32, 128)));
    const char *user_nwg = std::getenv("DEAC_EVOLVE_WORKGROUPS");
    if (user_nwg) nwg = std::atoi(user_nwg);
    h->nwg = std::clamp(nwg, 1, maxwg);
    // Leave a CU to the walk when that costs no extra round of chain tasks:
    // the walk then runs beside the evolution instead of in line.
    const int tasks = h->MG * h->TB;
    if (!user_nwg && h->nwg == maxwg && maxwg > 1 &&
        (tasks + maxwg - 2) / (maxwg - 1) == (tasks + maxwg - 1) / maxwg)
        h->nwg = maxwg - 1;
    h->overlap = h->nwg < maxwg;

    // Device buffers.
    const size_t kmat_n = static_cast<size_t>(h->TB) * deac::kChainTau * h->stride;
    const size_t pop_n = static_cast<size_t>(h->MG) * deac::kChainRows * h->stride;
    deac_status st = DEAC_OK;
    auto A = [&](auto *&ptr, size_t n) {
        if (st == DEAC_OK) st = h->alloc(ptr, n);
    };
    A(h->kmat, kmat_n);
    A(h->obs, T);
    A(h->sig, T);
    A(h->pop0, pop_n);
    A(h->pop1, pop_n);
    A(h->cur, P);
    A(h->fitness, P);
    A(h->cr_old, static_cast<size_t>(P) * m.nsign);
    A(h->f_old, static_cast<size_t>(P) * m.nsign);
    A(h->emod, static_cast<size_t>(P) * T);
    A(h->mom, static_cast<size_t>(P) * 4);
    A(h->sync, 16);
    A(h->iter, 4);
    for (int k = 0; k < 2; k++) {
        A(h->ctl[k], static_cast<size_t>(h->batch) * P);
        A(h->raw[k], static_cast<size_t>(h->batch) * P * Kc);
        A(h->gen_start[k], static_cast<size_t>(h->batch) * 4);
    }
    A(h->state, 4);
    A(h->skip_rows, 256 * 4);
    A(h->lane_tab, static_cast<size_t>(64) * 64 * 16 * 4);
    if (!m.nterm.empty()) A(h->nterm, Kc);
    if (!m.fterm.empty()) A(h->fterm, Kc);
    if (!m.tterm.empty()) A(h->tterm, Kc);
    if (m.use_nfm) A(h->nfm_dt, T);
    if (st != DEAC_OK) return fail(st);

    // K rows [tau][stride], zero padded to whole tau blocks and chunks.
    std::vector<double> kl(kmat_n, 0.0);
    for (size_t t = 0; t < m.T; t++) std::memcpy(&kl[t * h->stride], &m.kernel[t * m.Kc], m.Kc * sizeof(double));
    // Jump tables: T^(nsign*G) by rows for the walk, T^(l*chunk) by nibbles
    // for the mask generator.
    std::vector<uint64_t> rows(256 * 4), tabs(static_cast<size_t>(64) * 64 * 16 * 4);
    deac::matrix_rows(deac::transition_power(static_cast<uint64_t>(Kc)), rows.data());
    const deac::Gf2Matrix step = deac::transition_power(static_cast<uint64_t>(h->chunk));
    deac::Gf2Matrix acc = deac::Gf2Matrix::identity();
    for (int l = 0; l < 64; l++) {
        deac::nibble_table(acc, tabs.data() + static_cast<size_t>(l) * 64 * 16 * 4);
        acc = step * acc;
    }
    auto up = [&](void *dst, const void *src, size_t bytes) {
        if (st == DEAC_OK && hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess) st = DEAC_ERR_DEVICE;
    };
    up(h->kmat, kl.data(), kl.size() * sizeof(double));
    up(h->obs, m.isf.data(), m.T * sizeof(double));
    up(h->sig, m.err.data(), m.T * sizeof(double));
    up(h->skip_rows, rows.data(), rows.size() * sizeof(uint64_t));
    up(h->lane_tab, tabs.data(), tabs.size() * sizeof(uint64_t));
    if (h->nterm) up(h->nterm, m.nterm.data(), m.Kc * sizeof(double));
    if (h->fterm) up(h->fterm, m.fterm.data(), m.Kc * sizeof(double));
    if (h->tterm) up(h->tterm, m.tterm.data(), m.Kc * sizeof(double));
    if (h->nfm_dt) up(h->nfm_dt, m.nfm_dt.data(), m.T * sizeof(double));
    if (st != DEAC_OK) return fail(st);
    *out = h;
    return DEAC_OK;
}

void deac_destroy(deac_handle *h) { delete h; }

deac_status deac_last_run_profile(const deac_handle *h, deac_run_profile *profile)
{
    if (!h || !profile) return DEAC_ERR_INVALID_ARGUMENT;
    *profile = h->last;
    return DEAC_OK;
}

deac_status deac_run(deac_handle *h, uint64_t rng_state[4], const deac_run_options *o, deac_result *r)
{
    if (!h || !rng_state || !o || !r || !r->dsf || !r->frequency) return DEAC_ERR_INVALID_ARGUMENT;
    if (o->number_of_generations < 2) {
        deac::last_message() = "number_of_generations must be at least 2";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (o->ks_precision != DEAC_PRECISION_FP64 && o->ks_precision != DEAC_PRECISION_FP32 &&
        o->ks_precision != DEAC_PRECISION_BF16X3) {
        deac::last_message() = "ks_precision must be DEAC_PRECISION_FP64, DEAC_PRECISION_FP32 or DEAC_PRECISION_BF16X3";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if ((rng_state[0] | rng_state[1] | rng_state[2] | rng_state[3]) == 0) {
        // the all-zero state is a fixed point of xoshiro256+: every draw is 0
        // and the donor draws (deac.cpp:200-216) would never end
        deac::last_message() = "rng_state must not be all zero (a fixed point of xoshiro256+)";
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    const deac::Model &m = h->m;
    // bdsf always runs the split loop: its trial contraction is the
    // reference build's rounded-product fold (split_fold_kernel), not the
    // fused kernel's exact FMA chains
    const bool split = o->ks_precision != DEAC_PRECISION_FP64 || o->split_loop != 0 || m.bdsf;
    const size_t N = o->number_of_generations, P = m.P, Kc = m.Kc;
    const int nsign = m.nsign;
    hipStream_t s = h->stream;
    DEAC_HIP(hipSetDevice(h->device));
    if (split) DEAC_TRY(prepare_split(h, o->ks_precision));

    // Initial population and normalization on the host (deac.cpp:430-603).
    deac::Xoshiro rng{{rng_state[0], rng_state[1], rng_state[2], rng_state[3]}};
    std::vector<double> pop, lay;
    deac::initial_population(m, rng, pop);
    pad_rows(pop, P, Kc, h->MG, h->stride, lay);
    std::vector<double> zeros(lay.size(), 0.0);
    std::vector<int> cur(P, 0);
    std::vector<double> cr(P * nsign, m.cr0), fw(P * nsign, m.f0);  // deac.cpp:990-1032

    double *stats = nullptr;
    const bool track = o->track_stats != 0;
    if (track && h->alloc(stats, 3 * N) != DEAC_OK) return DEAC_ERR_OUT_OF_MEMORY;
    void *stats_alloc = stats;

    DEAC_HIP(hipMemcpyAsync(h->pop0, lay.data(), lay.size() * sizeof(double), hipMemcpyHostToDevice, s));
    DEAC_HIP(hipMemcpyAsync(h->pop1, zeros.data(), zeros.size() * sizeof(double), hipMemcpyHostToDevice, s));
    DEAC_HIP(hipMemcpyAsync(h->cur, cur.data(), P * sizeof(int), hipMemcpyHostToDevice, s));
    DEAC_HIP(hipMemcpyAsync(h->cr_old, cr.data(), cr.size() * sizeof(double), hipMemcpyHostToDevice, s));
    DEAC_HIP(hipMemcpyAsync(h->f_old, fw.data(), fw.size() * sizeof(double), hipMemcpyHostToDevice, s));
    DEAC_HIP(hipMemcpyAsync(h->state, rng.s, sizeof rng.s, hipMemcpyHostToDevice, s));

    deac::EvoArgs ea{};
    ea.P = static_cast<int>(P);
    ea.T = static_cast<int>(m.T);
    ea.G = static_cast<int>(m.G);
    ea.nsign = nsign;
    ea.Kc = static_cast<int>(Kc);
    ea.stride = h->stride;
    ea.MG = h->MG;
    ea.TB = h->TB;
    ea.use_first = m.use_first;
    ea.use_third = m.use_third;
    ea.renorm = m.normalize && m.renormalize;
    ea.allow_neg = m.allow_negative;
    ea.track_stats = track;
    ea.stop = o->stop_minimum_fitness;
    ea.zeroth = m.zeroth_moment;
    ea.first_moment = m.first_moment;
    ea.third_moment = m.third_moment;
    ea.third_moment_error = m.third_moment_error;
    ea.use_nfm = m.use_nfm;
    ea.nfm_value = m.nfm_value;
    ea.nfm_error = m.nfm_error;
    ea.nfm_dt = h->nfm_dt;
    ea.fold_ks = m.bdsf;
    ea.kmat = h->kmat;
    ea.nterm = h->nterm;
    ea.fterm = h->fterm;
    ea.tterm = h->tterm;
    ea.obs = h->obs;
    ea.sig = h->sig;
    ea.pop0 = h->pop0;
    ea.pop1 = h->pop1;
    ea.cur = h->cur;
    ea.fitness = h->fitness;
    ea.cr_old = h->cr_old;
    ea.f_old = h->f_old;
    ea.emod = h->emod;
    ea.mom = h->mom;
    ea.stats = stats;
    ea.stats_len = N;

    deac::WalkArgs wa{};
    wa.state = h->state;
    wa.skip_rows = h->skip_rows;
    wa.P = static_cast<int>(P);
    wa.nsign = nsign;
    wa.mod_magic = ~UINT64_C(0) / P;
    wa.tau_cr = m.tau_cr;
    wa.tau_f = m.tau_f;
    deac::MaskArgs ma{};
    ma.lane_tab = h->lane_tab;
    ma.P = static_cast<int>(P);
    ma.Kc = static_cast<int>(Kc);
    ma.chunk = h->chunk;

    // Generation loop: ii = 0 .. N-2 (deac.cpp:1138), in batches.  Batch b
    // uses buffer set b%2.  Its walk runs on the draw stream while batch b-1
    // evolves (on a CU of its own: the walk kernel holds enough LDS that it
    // cannot share one with an evolution workgroup), then its mask draws and
    // its evolution run on the evolution stream; the mask kernel spreads over
    // every CU and costs less alone than beside the evolution.  The evolution
    // of b is queued before the host has read batch b-1's status.  Behind a
    // stop it changes nothing: the stop test sits at the top of a generation,
    // so the fitnesses it finds are those that stopped batch b-1, and it
    // stops at its first generation (its status and its statistics entry,
    // past the stop, are never read).  Batches start at 8
    // generations and double up to the buffer size, so that the first walk,
    // which nothing hides, is short.  When the evolution fills every CU the
    // walk runs in line on the evolution stream.
    const size_t total = N - 1;
    std::vector<std::pair<size_t, int>> plan;  // first generation, generations
    for (size_t g = 0, len = std::min(8, h->batch); g < total;) {
        const int nb = static_cast<int>(std::min(len, total - g));
        plan.emplace_back(g, nb);
        g += static_cast<size_t>(nb);
        len = std::min(2 * len, static_cast<size_t>(h->batch));
    }
    // the split loop's kernels leave room on the CUs for the walk
    hipStream_t d = split || h->overlap ? h->draw_stream : s;
    auto draws = [&](size_t b) -> deac_status {
        const int k = static_cast<int>(b & 1);
        hipEvent_t *e = h->kev[k];
        if (b >= 2) DEAC_HIP(hipStreamWaitEvent(d, e[5], 0));  // batch b-2 done with the set
        wa.gen_start = h->gen_start[k];
        wa.ctl = h->ctl[k];
        wa.ngen = plan[b].second;
        DEAC_HIP(hipEventRecord(e[0], d));
        DEAC_HIP(deac::launch_walk(wa, d));
        DEAC_HIP(hipEventRecord(e[1], d));
        return DEAC_OK;
    };
    auto evolve = [&](size_t b) -> deac_status {
        const int k = static_cast<int>(b & 1);
        hipEvent_t *e = h->kev[k];
        unsigned *sy = h->sync + 8 * k;
        ea.ctl = h->ctl[k];
        ea.raw = h->raw[k];
        ea.bar = sy;
        ea.status = reinterpret_cast<int *>(sy + 4);
        ea.gen0 = static_cast<long>(plan[b].first);
        ea.ngen = plan[b].second;
        ma.ctl = h->ctl[k];
        ma.raw = h->raw[k];
        ma.ngen = plan[b].second;
        DEAC_HIP(hipStreamWaitEvent(s, e[1], 0));  // walk done
        DEAC_HIP(hipEventRecord(e[2], s));
        DEAC_HIP(deac::launch_mask(ma, s));
        DEAC_HIP(hipEventRecord(e[3], s));
        DEAC_HIP(deac::launch_zero_u32(sy, 8, s));
        DEAC_HIP(hipEventRecord(e[4], s));
        DEAC_HIP(deac::launch_evolve(ea, h->nwg, s));
        DEAC_HIP(hipEventRecord(e[5], s));
        DEAC_HIP(hipMemcpyAsync(h->host_status + 4 * k, sy + 4, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
        DEAC_HIP(hipEventRecord(e[6], s));
        return DEAC_OK;
    };
    deac_run_profile prof{};
    size_t done = 0;          // generations evolved
    long stopped_at = -1;     // generation index of the stop, if any
    auto pipeline = [&]() -> deac_status {
        // Initial fitness (deac.cpp:857-976), on buffer set 1; its status
        // goes to host slot 2.
        ea.init_mode = 1;
        ea.gen0 = 0;
        ea.ngen = 0;
        ea.bar = h->sync + 8;
        ea.status = reinterpret_cast<int *>(h->sync + 12);
        DEAC_HIP(deac::launch_zero_u32(h->sync + 8, 8, s));
        DEAC_HIP(deac::launch_evolve(ea, h->nwg, s));
        DEAC_HIP(hipMemcpyAsync(h->host_status + 8, h->sync + 12, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
        ea.init_mode = 0;
        DEAC_HIP(hipEventRecord(h->ev0, s));
        DEAC_HIP(hipStreamWaitEvent(d, h->ev0, 0));  // RNG state uploaded
        DEAC_TRY(draws(0));
        if (plan.size() > 1) DEAC_TRY(draws(1));
        DEAC_TRY(evolve(0));
        for (size_t b = 0; b < plan.size(); b++) {
            if (b + 1 < plan.size()) DEAC_TRY(evolve(b + 1));
            const int k = static_cast<int>(b & 1);
            hipEvent_t *e = h->kev[k];
            DEAC_HIP(hipEventSynchronize(e[6]));
            if (b == 0 && h->host_status[8 + 2]) return DEAC_ERR_DEVICE;  // initial fitness: barrier timeout
            float t_walk = 0, t_mask = 0, t_evo = 0;
            DEAC_HIP(hipEventElapsedTime(&t_walk, e[0], e[1]));
            DEAC_HIP(hipEventElapsedTime(&t_mask, e[2], e[3]));
            DEAC_HIP(hipEventElapsedTime(&t_evo, e[4], e[5]));
            prof.walk_ms += t_walk;
            prof.mask_ms += t_mask;
            prof.evolve_ms += t_evo;
            prof.evolve_launches += 1;
            const int *status = h->host_status + 4 * k;
            if (status[2]) return DEAC_ERR_DEVICE;  // a grid barrier timed out
            if (status[0]) {
                // RNG state at the top of the stopping generation; the set is
                // not rewritten before this copy (draws of b+2 not queued)
                stopped_at = static_cast<long>(plan[b].first) + status[1];
                uint64_t st4[4];
                DEAC_HIP(hipMemcpy(st4, h->gen_start[k] + static_cast<size_t>(status[1]) * 4, sizeof st4,
                                   hipMemcpyDeviceToHost));
                std::memcpy(rng.s, st4, sizeof st4);
                done += static_cast<size_t>(status[1]);
                return DEAC_OK;
            }
            done += static_cast<size_t>(plan[b].second);
            if (b + 2 < plan.size()) DEAC_TRY(draws(b + 2));
        }
        return DEAC_OK;
    };
    // The split loop: the same generations as separate kernels, batch by
    // batch on the plan above.  Batch b's draws (walk, then mask) run on the
    // draw stream while batch b-1 runs, into buffer set b%2; then, per
    // generation, top, trial, renormalization, the contraction at the chosen
    // precision, moments and selection run on the evolution stream.  One
    // status word serves the whole run: once a stop sets it, every kernel
    // queued behind (the rest of the batch and the next one) returns at once.
    // graphs of the split loop (per run: the kernel arguments hold this
    // run's buffers and options); DEAC_SPLIT_GRAPH sets the generations per
    // graph, 0 runs plain launches
    int graph_len = 8;
    if (const char *e = std::getenv("DEAC_SPLIT_GRAPH")) graph_len = std::max(0, std::atoi(e));
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    struct GraphFree {
        hipGraphExec_t *g;
        ~GraphFree()
        {
            for (int k = 0; k < 2; k++)
                if (g[k]) (void)hipGraphExecDestroy(g[k]);
        }
    } graph_free{gexec};
    auto split_loop = [&]() -> deac_status {
        const deac_handle::Split &sp = h->split;
        const deac::KsLayout &y = sp.y;
        const int Kb = static_cast<int>(y.Kb);
        ea.bar = h->sync;
        ea.status = reinterpret_cast<int *>(h->sync + 4);
        const int *halt = ea.status;
        // model ISF (and moment rows) of the contraction's genomes: the
        // current ones for the initial fitness, the trials otherwise
        auto evaluate = [&](const deac::EvoArgs &a, int init) -> deac_status {
            if (y.precision == DEAC_PRECISION_FP64 && a.fold_ks && !init) {
                DEAC_HIP(deac::launch_split_fold(a, s));
            } else if (y.precision == DEAC_PRECISION_FP64) {
                double *B = static_cast<double *>(sp.B);
                DEAC_HIP(deac::launch_split_pack_fp64(a, init, B, y.Nt, Kb, s));
                DEAC_HIP(deac::launch_ks_fp64(static_cast<const double *>(sp.A), B, a.emod, a.T, a.P, y.Mt, y.Nt,
                                              Kb, s, nullptr, nullptr, halt));
            } else {
                if (y.precision == DEAC_PRECISION_FP32) {
                    float *B = static_cast<float *>(sp.B);
                    DEAC_HIP(deac::launch_split_pack_fp32(a, init, B, y.Nt, Kb, s));
                    DEAC_HIP(deac::launch_ks_fp32(static_cast<const float *>(sp.A), B, sp.part, y.Mt, y.Nt, Kb,
                                                  y.kb_per_slice, y.slices, s, nullptr, nullptr, halt));
                } else {
                    DEAC_HIP(deac::launch_split_pack_bf16x3(a, init, sp.B, sp.Bl, y.Nt, Kb, s));
                    DEAC_HIP(deac::launch_ks_bf16x3(sp.A, sp.Al, sp.B, sp.Bl, sp.part, y.Mt, y.Nt, Kb, y.kb_per_slice,
                                                    y.slices, s, nullptr, nullptr, halt));
                }
                DEAC_HIP(deac::launch_split_model(a, sp.part, y.slices, static_cast<int>(y.npad()),
                                                  y.mpad() * y.npad(), sp.ex, s));
            }
            if (a.use_first || a.use_third) DEAC_HIP(deac::launch_split_moments(a, init, s));
            return DEAC_OK;
        };
        // One generation g of a batch: top and trials, renormalization,
        // contraction, moments, selection; with a.iter, the generation comes
        // from the device counter, which the last kernel advances.
        auto one_generation = [&](const deac::EvoArgs &a, int g) -> deac_status {
            DEAC_HIP(deac::launch_split_trial(a, g, s));  // with the top of generation g
            if (a.renorm) DEAC_HIP(deac::launch_split_renorm(a, s));
            DEAC_TRY(evaluate(a, 0));
            DEAC_HIP(deac::launch_split_select(a, g, 0, s));
            if (a.iter) DEAC_HIP(deac::launch_split_iter(a.iter, 0, 0, s));
            return DEAC_OK;
        };
        // Graphs: the kernels of graph_len generations captured once per
        // buffer set and run, replayed for every full group of a batch's
        // generations (a launch per group instead of ~5 per generation);
        // the remainder of a batch runs as plain launches.  A capture the
        // runtime refuses leaves the run on plain launches.
        auto capture = [&](int k, const deac::EvoArgs &a) -> bool {
            hipGraph_t graph = nullptr;
            if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                (void)hipGetLastError();
                return false;
            }
            deac_status st = DEAC_OK;
            for (int r = 0; r < graph_len && st == DEAC_OK; r++) st = one_generation(a, 0);
            const hipError_t e = hipStreamEndCapture(s, &graph);
            bool ok = st == DEAC_OK && e == hipSuccess && graph;
            if (ok) ok = hipGraphInstantiate(&gexec[k], graph, nullptr, nullptr, 0) == hipSuccess;
            if (graph) (void)hipGraphDestroy(graph);
            if (!ok) {
                (void)hipGetLastError();
                gexec[k] = nullptr;
            }
            return ok;
        };
        // Initial fitness (deac.cpp:857-976).
        DEAC_HIP(deac::launch_zero_u32(h->sync, 8, s));
        DEAC_TRY(evaluate(ea, 1));
        DEAC_HIP(deac::launch_split_select(ea, 0, 1, s));
        DEAC_HIP(hipEventRecord(h->ev0, s));
        DEAC_HIP(hipStreamWaitEvent(d, h->ev0, 0));  // RNG state uploaded
        auto walk = [&](size_t b) -> deac_status {
            const int k = static_cast<int>(b & 1);
            hipEvent_t *e = h->kev[k];
            if (b >= 2) DEAC_HIP(hipStreamWaitEvent(d, e[5], 0));  // batch b-2 done with the set
            wa.gen_start = h->gen_start[k];
            wa.ctl = h->ctl[k];
            wa.ngen = plan[b].second;
            ma.ctl = h->ctl[k];
            ma.raw = h->raw[k];
            ma.ngen = plan[b].second;
            DEAC_HIP(hipEventRecord(e[0], d));
            DEAC_HIP(deac::launch_walk(wa, d));
            DEAC_HIP(hipEventRecord(e[1], d));
            DEAC_HIP(deac::launch_mask(ma, d));
            DEAC_HIP(hipEventRecord(e[2], d));
            return DEAC_OK;
        };
        auto generations = [&](size_t b) -> deac_status {
            const int k = static_cast<int>(b & 1);
            hipEvent_t *e = h->kev[k];
            const int nb = plan[b].second;
            DEAC_HIP(hipStreamWaitEvent(s, e[2], 0));  // draws ready
            DEAC_HIP(hipEventRecord(e[3], s));
            ea.ctl = h->ctl[k];
            ea.raw = h->raw[k];
            ea.gen0 = static_cast<long>(plan[b].first);
            ea.ngen = nb;
            int g = 0;
            if (graph_len > 0 && nb >= graph_len) {
                deac::EvoArgs eg = ea;
                eg.iter = h->iter + 2 * k;
                DEAC_HIP(deac::launch_split_iter(eg.iter, eg.gen0, 1, s));
                if (!gexec[k] && !capture(k, eg)) graph_len = 0;
                if (gexec[k]) {
                    for (; g + graph_len <= nb; g += graph_len) {
                        DEAC_HIP(hipGraphLaunch(gexec[k], s));
                        prof.graph_launches += 1;
                    }
                    for (; g < nb; g++) DEAC_TRY(one_generation(eg, 0));
                }
            }
            for (; g < nb; g++) DEAC_TRY(one_generation(ea, g));
            DEAC_HIP(hipEventRecord(e[5], s));
            DEAC_HIP(hipMemcpyAsync(h->host_status + 4 * k, h->sync + 4, 4 * sizeof(int), hipMemcpyDeviceToHost, s));
            DEAC_HIP(hipEventRecord(e[6], s));
            return DEAC_OK;
        };
        DEAC_TRY(walk(0));
        if (plan.size() > 1) DEAC_TRY(walk(1));
        DEAC_TRY(generations(0));
        for (size_t b = 0; b < plan.size(); b++) {
            if (b + 1 < plan.size()) DEAC_TRY(generations(b + 1));
            const int k = static_cast<int>(b & 1);
            hipEvent_t *e = h->kev[k];
            DEAC_HIP(hipEventSynchronize(e[6]));
            float t_walk = 0, t_mask = 0, t_evo = 0;
            DEAC_HIP(hipEventElapsedTime(&t_walk, e[0], e[1]));
            DEAC_HIP(hipEventElapsedTime(&t_mask, e[1], e[2]));
            DEAC_HIP(hipEventElapsedTime(&t_evo, e[3], e[5]));
            prof.walk_ms += t_walk;
            prof.mask_ms += t_mask;
            prof.evolve_ms += t_evo;
            prof.evolve_launches += 1;
            // the copy after batch b: a stop there, at generation status[1] of
            // the batch (later batches have not run, or returned at once)
            const int *status = h->host_status + 4 * k;
            if (status[0]) {
                stopped_at = static_cast<long>(plan[b].first) + status[1];
                uint64_t st4[4];
                DEAC_HIP(hipMemcpy(st4, h->gen_start[k] + static_cast<size_t>(status[1]) * 4, sizeof st4,
                                   hipMemcpyDeviceToHost));
                std::memcpy(rng.s, st4, sizeof st4);
                done += static_cast<size_t>(status[1]);
                return DEAC_OK;
            }
            done += static_cast<size_t>(plan[b].second);
            if (b + 2 < plan.size()) DEAC_TRY(walk(b + 2));
        }
        return DEAC_OK;
    };
    const deac_status ps = split ? split_loop() : pipeline();
    // nothing stays in flight past this call, whatever the outcome
    const bool drained = hipEventRecord(h->ev1, s) == hipSuccess && hipStreamSynchronize(s) == hipSuccess &&
                         hipStreamSynchronize(d) == hipSuccess;
    if (ps != DEAC_OK) return ps;
    if (!drained) {
        (void)hipGetLastError();
        return DEAC_ERR_DEVICE;
    }
    float ms = 0.0f;
    DEAC_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    prof.loop_ms = ms;
    prof.generations = done;
    h->last = prof;
    if (stopped_at < 0) DEAC_HIP(hipMemcpy(rng.s, h->state, sizeof rng.s, hipMemcpyDeviceToHost));
    std::memcpy(rng_state, rng.s, sizeof rng.s);

    // Epilogue on the host (deac.cpp:1576-1685).
    std::vector<double> fit(P), l0(lay.size()), l1(lay.size());
    DEAC_HIP(hipMemcpy(fit.data(), h->fitness, P * sizeof(double), hipMemcpyDeviceToHost));
    DEAC_HIP(hipMemcpy(cur.data(), h->cur, P * sizeof(int), hipMemcpyDeviceToHost));
    DEAC_HIP(hipMemcpy(l0.data(), h->pop0, l0.size() * sizeof(double), hipMemcpyDeviceToHost));
    DEAC_HIP(hipMemcpy(l1.data(), h->pop1, l1.size() * sizeof(double), hipMemcpyDeviceToHost));
    size_t generation = stopped_at >= 0 ? static_cast<size_t>(stopped_at) : N - 2;
    size_t best = 0;
    double minimum = fit[0];
    for (size_t i = 0; i < P; i++)
        if (fit[i] < minimum) {
            minimum = fit[i];
            best = i;
        }
    const std::vector<double> &lb = cur[best] ? l1 : l0;
    deac::output_transform(m, &lb[best * h->stride], r->dsf, r->frequency);

    if (track) {
        std::vector<double> host_stats(3 * N, 0.0);
        const size_t n = generation + 1;
        DEAC_HIP(hipMemcpy(host_stats.data(), stats, 3 * N * sizeof(double), hipMemcpyDeviceToHost));
        if (r->stats_mean) std::memcpy(r->stats_mean, host_stats.data(), n * sizeof(double));
        if (r->stats_minimum) std::memcpy(r->stats_minimum, host_stats.data() + N, n * sizeof(double));
        if (r->stats_squared_mean) std::memcpy(r->stats_squared_mean, host_stats.data() + 2 * N, n * sizeof(double));
    }
    if (generation + 2 == N) {  // deac.cpp:1677-1685
        generation += 1;
        if (track) {
            if (r->stats_mean) r->stats_mean[generation] = deac::mean_of(fit.data(), P);
            if (r->stats_minimum) r->stats_minimum[generation] = minimum;
            if (r->stats_squared_mean) r->stats_squared_mean[generation] = deac::squared_mean_of(fit.data(), P);
        }
    }
    r->generation = generation;
    r->minimum_fitness = minimum;
    r->best_index = best;
    if (r->final_fitness) std::memcpy(r->final_fitness, fit.data(), P * sizeof(double));
    if (stats_alloc) {
        h->allocations.erase(std::find(h->allocations.begin(), h->allocations.end(), stats_alloc));
        (void)hipFree(stats_alloc);
    }
    return DEAC_OK;
}

}  // extern "C"
