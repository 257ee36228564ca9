// deac_model_isf: the model ISF E = K.S (deac.cpp:1385-1395) and the reduced
// chi^2 (deac.cpp:120-126, 1540-1542) of a stack of genomes at a chosen K.S
// precision -- the precision sweep of the K.S contraction.  FP64 is the
// engine's exact evaluation; FP32 and BF16X3 are the mixed-precision forms
// (K scaled per tau row by a power of two, FP32 accumulation), whose error
// against FP64 this entry point exists to measure.  chi^2 is FP64 in the
// reference's order at every precision.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "deac_engine.h"
#include "host_model.hpp"
#include "ks_layout.hpp"
#include "precision_kernels.hpp"

namespace {


// Device buffers of one call, freed on every exit path.
struct Buffers {
    std::vector<void *> ptrs;
    ~Buffers()
    {
        for (void *p : ptrs) (void)hipFree(p);
    }
    template <class T>
    bool alloc(T *&out, size_t count)
    {
        void *p = nullptr;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        ptrs.push_back(p);
        out = static_cast<T *>(p);
        return true;
    }
};

// chi^2 / T in the reference's order: rounded squares added in tau order, the
// last one fused when T is odd (deac.cpp:120-126, 1540-1542).
double reduced_chi2(const deac::Model &m, const double *e)
{
    const size_t T = m.T;
    double acc = 0.0;
    for (size_t t = 0; t < T; t++) {
        const double d = (m.isf[t] - e[t]) / m.err[t];
        acc = ((T & 1) && t == T - 1) ? std::fma(d, d, acc) : acc + d * d;
    }
    return acc / static_cast<double>(T);
}

}  // namespace

#define PREC_HIP(call)                                  \
    do {                                                \
        if ((call) != hipSuccess) {                     \
            (void)hipGetLastError();                    \
            return DEAC_ERR_DEVICE;                     \
        }                                               \
    } while (0)

extern "C" size_t deac_genome_length(const deac_problem *p)
{
    if (!p) return 0;
    return (p->zero_temperature || p->spectra_type == DEAC_BDSF) ? p->genome_size : 2 * p->genome_size;
}

extern "C" deac_status deac_model_isf(const deac_problem *p, const double *genomes, size_t count, int precision,
                                      double *model_isf, double *chi2, double *kernel_ms, const char **message)
{
    if (message) *message = nullptr;
    if (!p || !genomes || count == 0) return DEAC_ERR_INVALID_ARGUMENT;
    if (precision != DEAC_PRECISION_FP64 && precision != DEAC_PRECISION_FP32 && precision != DEAC_PRECISION_BF16X3) {
        deac::last_message() = "precision must be DEAC_PRECISION_FP64, DEAC_PRECISION_FP32 or DEAC_PRECISION_BF16X3";
        if (message) *message = deac::last_message().c_str();
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (count > (1u << 24)) {
        deac::last_message() = "at most 2^24 genomes per call";
        if (message) *message = deac::last_message().c_str();
        return DEAC_ERR_UNSUPPORTED;
    }
    const deac_status v = deac::validate(*p, deac::last_message());
    if (v != DEAC_OK) {
        if (message) *message = deac::last_message().c_str();
        return v;
    }
    int devices = 0, device = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) {
        (void)hipGetLastError();
        return DEAC_ERR_NO_DEVICE;
    }
    PREC_HIP(hipGetDevice(&device));
    hipDeviceProp_t prop;
    PREC_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return DEAC_ERR_NO_DEVICE;

    const deac::Model m = deac::build_model(*p);
    const deac::KsLayout y = deac::ks_layout(m, precision, count);
    const int T = static_cast<int>(m.T), N = static_cast<int>(count);
    const int Kb = static_cast<int>(y.Kb);
    std::vector<double> E(static_cast<size_t>(N) * T);

    hipStream_t s = nullptr;
    PREC_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    PREC_HIP(hipEventCreate(&e0));
    PREC_HIP(hipEventCreate(&e1));
    struct EventGuard {
        hipEvent_t a, b;
        ~EventGuard()
        {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } eg{e0, e1};
    Buffers buf;
    const size_t asz = y.a_count(), bsz = y.b_count();

    if (precision == DEAC_PRECISION_FP64) {
        std::vector<double> A, B;
        deac::ks_kernel_fp64(m, y, A);
        deac::ks_genomes_fp64(y, genomes, B);
        double *dA, *dB, *dE;
        if (!buf.alloc(dA, asz) || !buf.alloc(dB, bsz) || !buf.alloc(dE, E.size())) return DEAC_ERR_OUT_OF_MEMORY;
        PREC_HIP(hipMemcpyAsync(dA, A.data(), asz * sizeof(double), hipMemcpyHostToDevice, s));
        PREC_HIP(hipMemcpyAsync(dB, B.data(), bsz * sizeof(double), hipMemcpyHostToDevice, s));
        PREC_HIP(deac::launch_ks_fp64(dA, dB, dE, T, N, y.Mt, y.Nt, Kb, s, e0, e1, nullptr));
        PREC_HIP(hipMemcpyAsync(E.data(), dE, E.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        PREC_HIP(hipStreamSynchronize(s));
    } else {
        const size_t mn = y.mpad() * y.npad();
        float *part, *out;
        if (!buf.alloc(part, static_cast<size_t>(y.slices) * mn) || !buf.alloc(out, mn)) return DEAC_ERR_OUT_OF_MEMORY;
        // Each branch queues the slice sum before its host operands are
        // freed (a few ms for large stacks), so the end event follows the
        // contraction directly.
        if (precision == DEAC_PRECISION_FP32) {
            std::vector<float> A, B;
            deac::ks_kernel_fp32(m, y, A);
            deac::ks_genomes_fp32(y, genomes, B);
            float *dA, *dB;
            if (!buf.alloc(dA, asz) || !buf.alloc(dB, bsz)) return DEAC_ERR_OUT_OF_MEMORY;
            PREC_HIP(hipMemcpyAsync(dA, A.data(), asz * sizeof(float), hipMemcpyHostToDevice, s));
            PREC_HIP(hipMemcpyAsync(dB, B.data(), bsz * sizeof(float), hipMemcpyHostToDevice, s));
            PREC_HIP(deac::launch_ks_fp32(dA, dB, part, y.Mt, y.Nt, Kb, y.kb_per_slice, y.slices, s, e0, nullptr,
                                          nullptr));
            PREC_HIP(deac::launch_reduce_slices(part, out, y.slices, mn, s, nullptr, e1));
        } else {
            std::vector<uint16_t> Ah, Al, Bh, Bl;
            deac::ks_kernel_bf16x3(m, y, Ah, Al);
            deac::ks_genomes_bf16x3(y, genomes, Bh, Bl);
            uint16_t *dAh, *dAl, *dBh, *dBl;
            if (!buf.alloc(dAh, asz) || !buf.alloc(dAl, asz) || !buf.alloc(dBh, bsz) || !buf.alloc(dBl, bsz))
                return DEAC_ERR_OUT_OF_MEMORY;
            PREC_HIP(hipMemcpyAsync(dAh, Ah.data(), asz * 2, hipMemcpyHostToDevice, s));
            PREC_HIP(hipMemcpyAsync(dAl, Al.data(), asz * 2, hipMemcpyHostToDevice, s));
            PREC_HIP(hipMemcpyAsync(dBh, Bh.data(), bsz * 2, hipMemcpyHostToDevice, s));
            PREC_HIP(hipMemcpyAsync(dBl, Bl.data(), bsz * 2, hipMemcpyHostToDevice, s));
            PREC_HIP(deac::launch_ks_bf16x3(dAh, dAl, dBh, dBl, part, y.Mt, y.Nt, Kb, y.kb_per_slice, y.slices, s,
                                            e0, nullptr, nullptr));
            PREC_HIP(deac::launch_reduce_slices(part, out, y.slices, mn, s, nullptr, e1));
        }
        std::vector<float> E32(mn);
        PREC_HIP(hipMemcpyAsync(E32.data(), out, E32.size() * sizeof(float), hipMemcpyDeviceToHost, s));
        PREC_HIP(hipStreamSynchronize(s));
        const size_t Npad = y.npad();
        for (int n = 0; n < N; n++)
            for (int t = 0; t < T; t++)
                E[static_cast<size_t>(n) * T + t] =
                    std::ldexp(static_cast<double>(E32[static_cast<size_t>(t) * Npad + n]), y.ex[t]);
    }
    if (kernel_ms) {
        float ms = 0.0f;
        PREC_HIP(hipEventElapsedTime(&ms, e0, e1));
        *kernel_ms = ms;
    }
    if (model_isf) std::memcpy(model_isf, E.data(), E.size() * sizeof(double));
    if (chi2)
        for (int n = 0; n < N; n++) chi2[n] = reduced_chi2(m, &E[static_cast<size_t>(n) * T]);
    return DEAC_OK;
}
