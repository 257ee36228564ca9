// Mixed-precision K.S instantiations of the batched DEAC kernel
// (batch_device.hpp): FP32 and BF16X3 delta-form contractions.
#include "batch_device.hpp"

namespace deac {

namespace {

// MOM = 1: the instantiation with moment constraints (the kernel's REM
// parameter, which the mixed contraction does not otherwise use)
template <int PREC, int NQ, int NM, int MOM>
hipError_t launch_mixed_one(const BatchDesc *descs, int count, hipStream_t s)
{
    constexpr size_t lds_bytes = static_cast<size_t>(kBatchRegionDoubles) * sizeof(double);
    static bool configured = false;
    if (!configured) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<PREC, NQ, NM, MOM>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_bytes));
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((batch_kernel<PREC, NQ, NM, MOM>), dim3(count), dim3(kBatchThreads), lds_bytes, s, descs);
    return hipGetLastError();
}

template <int PREC, int NM, int MOM>
hipError_t launch_mixed_nq(const BatchDesc *descs, int count, int nq, hipStream_t s)
{
    switch (nq) {
    case 1: return launch_mixed_one<PREC, 1, NM, MOM>(descs, count, s);
    case 2: return launch_mixed_one<PREC, 2, NM, MOM>(descs, count, s);
    case 3: return launch_mixed_one<PREC, 3, NM, MOM>(descs, count, s);
    case 5: return launch_mixed_one<PREC, 5, NM, MOM>(descs, count, s);
    case 9: if constexpr (NM >= 2) return launch_mixed_one<PREC, 9, NM, MOM>(descs, count, s);
    }
    return hipErrorInvalidValue;
}

template <int PREC, int MOM>
hipError_t launch_mixed_nm(const BatchDesc *descs, int count, int nq, int nm, hipStream_t s)
{
    if (nm == 1) return launch_mixed_nq<PREC, 1, MOM>(descs, count, nq, s);
    if (nm == 2) return launch_mixed_nq<PREC, 2, MOM>(descs, count, nq, s);
    return launch_mixed_nq<PREC, 4, MOM>(descs, count, nq, s);
}

}  // namespace

int batch_mixed_nq(int ntm, int nm)
{
    // pairs per chain wave: ceil(ntm / 2) with two member tiles (a wave keeps
    // one member tile), ceil(ntm / 4) with one, ntm with four (a wave keeps
    // one member tile and takes every kernel tile)
    const int need = nm == 4 ? ntm : nm == 2 ? (ntm + 1) / 2 : (ntm + 3) / 4;
    for (int nq : kBatchMixedNq)
        if (nq >= need && (nm >= 2 || nq <= 5)) return nq;
    return 0;
}

hipError_t launch_batch_mixed(const BatchDesc *descs, int count, int precision, int nq, int nm, int moments,
                              hipStream_t s)
{
    if (nm != 1 && nm != 2 && nm != 4) return hipErrorInvalidValue;
    if (precision == kPrecFp32)
        return moments ? launch_mixed_nm<kPrecFp32, 1>(descs, count, nq, nm, s)
                       : launch_mixed_nm<kPrecFp32, 0>(descs, count, nq, nm, s);
    if (precision == kPrecBf16)
        return moments ? launch_mixed_nm<kPrecBf16, 1>(descs, count, nq, nm, s)
                       : launch_mixed_nm<kPrecBf16, 0>(descs, count, nq, nm, s);
    return hipErrorInvalidValue;
}


hipError_t batch_role_cycles_mixed(unsigned long long *out, int reset)
{
#ifdef DEAC_ROLE_TIMERS
    hipError_t e = hipMemcpyFromSymbol(out, HIP_SYMBOL(g_role), sizeof(unsigned long long) * 3 * kBatchWaves);
    if (e != hipSuccess || !reset) return e;
    const unsigned long long z[3 * kBatchWaves] = {};
    return hipMemcpyToSymbol(HIP_SYMBOL(g_role), z, sizeof z);
#else
    (void)out;
    (void)reset;
    return hipErrorNotSupported;
#endif
}

}  // namespace deac
