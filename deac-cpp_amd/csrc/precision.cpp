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
#include "precision_kernels.hpp"

namespace {

thread_local std::string g_precision_message;

// Round-to-nearest-even FP32 -> bfloat16 bits (finite inputs).
uint16_t bf16_bits(float x)
{
    uint32_t u;
    std::memcpy(&u, &x, sizeof u);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    return static_cast<uint16_t>(u >> 16);
}

float bf16_value(uint16_t b)
{
    const uint32_t u = static_cast<uint32_t>(b) << 16;
    float x;
    std::memcpy(&x, &u, sizeof x);
    return x;
}

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
    return p->zero_temperature ? p->genome_size : 2 * p->genome_size;
}

extern "C" deac_status deac_model_isf(const deac_problem *p, const double *genomes, size_t count, int precision,
                                      double *model_isf, double *chi2, double *kernel_ms, const char **message)
{
    if (message) *message = nullptr;
    if (!p || !genomes || count == 0) return DEAC_ERR_INVALID_ARGUMENT;
    if (precision != DEAC_PRECISION_FP64 && precision != DEAC_PRECISION_FP32 && precision != DEAC_PRECISION_BF16X3) {
        g_precision_message = "precision must be DEAC_PRECISION_FP64, DEAC_PRECISION_FP32 or DEAC_PRECISION_BF16X3";
        if (message) *message = g_precision_message.c_str();
        return DEAC_ERR_INVALID_ARGUMENT;
    }
    if (count > (1u << 24)) {
        g_precision_message = "at most 2^24 genomes per call";
        if (message) *message = g_precision_message.c_str();
        return DEAC_ERR_UNSUPPORTED;
    }
    const deac_status v = deac::validate(*p, g_precision_message);
    if (v != DEAC_OK) {
        if (message) *message = g_precision_message.c_str();
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
    const int T = static_cast<int>(m.T), N = static_cast<int>(count);
    const size_t L = m.Kc;
    const int Mt = (T + 15) / 16, Nt = (N + 15) / 16;
    const size_t Mpad = static_cast<size_t>(Mt) * 16, Npad = static_cast<size_t>(Nt) * 16;
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

    if (precision == DEAC_PRECISION_FP64) {
        // Padding: K with -0.0 and S with +0.0, so padded products are -0.0 and
        // leave every chain exactly as it is.
        const size_t Kb = (L + 4 * deac::kSweepDepth - 1) / (4 * deac::kSweepDepth) * deac::kSweepDepth;
        std::vector<double> A(static_cast<size_t>(Mt) * Kb * 64, -0.0), B(static_cast<size_t>(Nt) * Kb * 64, 0.0);
        for (int mt = 0; mt < Mt; mt++)
            for (size_t kb = 0; kb < Kb; kb++)
                for (int l = 0; l < 64; l++) {
                    const size_t t = static_cast<size_t>(mt) * 16 + (l & 15), k = 4 * kb + (l >> 4);
                    if (t < m.T && k < L) A[(static_cast<size_t>(mt) * Kb + kb) * 64 + l] = m.kernel[t * L + k];
                }
        for (int nt = 0; nt < Nt; nt++)
            for (size_t kb = 0; kb < Kb; kb++)
                for (int l = 0; l < 64; l++) {
                    const size_t n = static_cast<size_t>(nt) * 16 + (l & 15), k = 4 * kb + (l >> 4);
                    if (n < count && k < L) B[(static_cast<size_t>(nt) * Kb + kb) * 64 + l] = genomes[n * L + k];
                }
        double *dA, *dB, *dE;
        if (!buf.alloc(dA, A.size()) || !buf.alloc(dB, B.size()) || !buf.alloc(dE, E.size())) return DEAC_ERR_OUT_OF_MEMORY;
        PREC_HIP(hipMemcpyAsync(dA, A.data(), A.size() * sizeof(double), hipMemcpyHostToDevice, s));
        PREC_HIP(hipMemcpyAsync(dB, B.data(), B.size() * sizeof(double), hipMemcpyHostToDevice, s));
        PREC_HIP(deac::launch_ks_fp64(dA, dB, dE, T, N, Mt, Nt, static_cast<int>(Kb), s, e0, e1));
        PREC_HIP(hipMemcpyAsync(E.data(), dE, E.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        PREC_HIP(hipStreamSynchronize(s));
    } else {
        // K row t scaled by 2^-e_t (exact), e_t the exponent of its largest
        // magnitude, so every row's largest entry lies in [1, 2).
        std::vector<int> ex(m.T, 0);
        for (size_t t = 0; t < m.T; t++) {
            double mx = 0.0;
            for (size_t k = 0; k < L; k++) mx = std::fmax(mx, std::fabs(m.kernel[t * L + k]));
            ex[t] = mx > 0.0 ? std::ilogb(mx) : 0;
        }
        auto kval = [&](size_t t, size_t k) { return std::ldexp(m.kernel[t * L + k], -ex[t]); };
        const bool bf = precision == DEAC_PRECISION_BF16X3;
        const size_t kblock = bf ? 32 : 4;
        const size_t depth = bf ? deac::kSweepDepth16 : deac::kSweepDepth;
        const size_t Kb = (L + kblock * depth - 1) / (kblock * depth) * depth;  // zero padded
        const size_t per = bf ? 8 : 1;  // values per lane and k-block
        // genome slices: enough workgroups (about 2048) to fill the device,
        // at least 8 k-blocks per slice
        const size_t tiles = static_cast<size_t>(Mt) * ((Nt + 3) / 4);
        const size_t want = std::max<size_t>(1, std::min<size_t>((2048 + tiles - 1) / tiles, std::max<size_t>(Kb / 8, 1)));
        const int kb_per_slice = static_cast<int>(((Kb + want - 1) / want + depth - 1) / depth * depth);
        const int slices = static_cast<int>((Kb + kb_per_slice - 1) / kb_per_slice);
        float *part, *out;
        if (!buf.alloc(part, static_cast<size_t>(slices) * Mpad * Npad) || !buf.alloc(out, Mpad * Npad))
            return DEAC_ERR_OUT_OF_MEMORY;
        const size_t asz = static_cast<size_t>(Mt) * Kb * 64 * per, bsz = static_cast<size_t>(Nt) * Kb * 64 * per;
        // element (row r of a tile, position k) of operand slot (tile, kb, lane, e)
        auto a_at = [&](int mt, size_t kb, int l, size_t e, size_t &t, size_t &k) {
            t = static_cast<size_t>(mt) * 16 + (l & 15);
            k = kblock * kb + (bf ? 8 * static_cast<size_t>(l >> 4) + e : static_cast<size_t>(l >> 4));
        };
        if (!bf) {
            std::vector<float> A(asz, 0.0f), B(bsz, 0.0f);
            for (int mt = 0; mt < Mt; mt++)
                for (size_t kb = 0; kb < Kb; kb++)
                    for (int l = 0; l < 64; l++) {
                        size_t t, k;
                        a_at(mt, kb, l, 0, t, k);
                        if (t < m.T && k < L) A[(static_cast<size_t>(mt) * Kb + kb) * 64 + l] = static_cast<float>(kval(t, k));
                    }
            for (int nt = 0; nt < Nt; nt++)
                for (size_t kb = 0; kb < Kb; kb++)
                    for (int l = 0; l < 64; l++) {
                        size_t n, k;
                        a_at(nt, kb, l, 0, n, k);
                        if (n < count && k < L) B[(static_cast<size_t>(nt) * Kb + kb) * 64 + l] = static_cast<float>(genomes[n * L + k]);
                    }
            float *dA, *dB;
            if (!buf.alloc(dA, asz) || !buf.alloc(dB, bsz)) return DEAC_ERR_OUT_OF_MEMORY;
            PREC_HIP(hipMemcpyAsync(dA, A.data(), asz * sizeof(float), hipMemcpyHostToDevice, s));
            PREC_HIP(hipMemcpyAsync(dB, B.data(), bsz * sizeof(float), hipMemcpyHostToDevice, s));
            PREC_HIP(deac::launch_ks_fp32(dA, dB, part, Mt, Nt, static_cast<int>(Kb), kb_per_slice, slices, s, e0,
                                          nullptr));
            PREC_HIP(deac::launch_reduce_slices(part, out, slices, Mpad * Npad, s, nullptr, e1));
        } else {
            // x = hi + lo: hi = bf16(x), lo = bf16(x - hi) (the remainder in FP32)
            std::vector<uint16_t> Ah(asz, 0), Al(asz, 0), Bh(bsz, 0), Bl(bsz, 0);
            auto split = [](double x, uint16_t &h, uint16_t &lo) {
                const float f = static_cast<float>(x);
                h = bf16_bits(f);
                lo = bf16_bits(f - bf16_value(h));
            };
            for (int mt = 0; mt < Mt; mt++)
                for (size_t kb = 0; kb < Kb; kb++)
                    for (int l = 0; l < 64; l++)
                        for (size_t e = 0; e < 8; e++) {
                            size_t t, k;
                            a_at(mt, kb, l, e, t, k);
                            if (t < m.T && k < L) {
                                const size_t i = ((static_cast<size_t>(mt) * Kb + kb) * 64 + l) * 8 + e;
                                split(kval(t, k), Ah[i], Al[i]);
                            }
                        }
            for (int nt = 0; nt < Nt; nt++)
                for (size_t kb = 0; kb < Kb; kb++)
                    for (int l = 0; l < 64; l++)
                        for (size_t e = 0; e < 8; e++) {
                            size_t n, k;
                            a_at(nt, kb, l, e, n, k);
                            if (n < count && k < L) {
                                const size_t i = ((static_cast<size_t>(nt) * Kb + kb) * 64 + l) * 8 + e;
                                split(genomes[n * L + k], Bh[i], Bl[i]);
                            }
                        }
            uint16_t *dAh, *dAl, *dBh, *dBl;
            if (!buf.alloc(dAh, asz) || !buf.alloc(dAl, asz) || !buf.alloc(dBh, bsz) || !buf.alloc(dBl, bsz))
                return DEAC_ERR_OUT_OF_MEMORY;
            PREC_HIP(hipMemcpyAsync(dAh, Ah.data(), asz * 2, hipMemcpyHostToDevice, s));
            PREC_HIP(hipMemcpyAsync(dAl, Al.data(), asz * 2, hipMemcpyHostToDevice, s));
            PREC_HIP(hipMemcpyAsync(dBh, Bh.data(), bsz * 2, hipMemcpyHostToDevice, s));
            PREC_HIP(hipMemcpyAsync(dBl, Bl.data(), bsz * 2, hipMemcpyHostToDevice, s));
            PREC_HIP(deac::launch_ks_bf16x3(dAh, dAl, dBh, dBl, part, Mt, Nt, static_cast<int>(Kb), kb_per_slice,
                                            slices, s, e0, nullptr));
            PREC_HIP(deac::launch_reduce_slices(part, out, slices, Mpad * Npad, s, nullptr, e1));
        }
        std::vector<float> E32(Mpad * Npad);
        PREC_HIP(hipMemcpyAsync(E32.data(), out, E32.size() * sizeof(float), hipMemcpyDeviceToHost, s));
        PREC_HIP(hipStreamSynchronize(s));
        for (int n = 0; n < N; n++)
            for (int t = 0; t < T; t++)
                E[static_cast<size_t>(n) * T + t] = std::ldexp(static_cast<double>(E32[static_cast<size_t>(t) * Npad + n]), ex[t]);
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
