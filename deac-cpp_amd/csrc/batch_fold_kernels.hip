// Detailed-balance (bdsf) instantiations of the batched DEAC kernel
// (batch_device.hpp): FP64 throughout, the trials' model ISF a fold of
// rounded products on the VALU (fold_chains).
#include "batch_device.hpp"

namespace deac {

namespace {

template <int MF, int NM>
hipError_t launch_fold_one(const BatchDesc *descs, int count, hipStream_t s)
{
    constexpr size_t lds_bytes = static_cast<size_t>(kBatchRegionDoubles) * sizeof(double);
    static bool configured = false;
    if (!configured) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<kPrecFold, MF, NM, 0>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_bytes));
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((batch_kernel<kPrecFold, MF, NM, 0>), dim3(count), dim3(kBatchThreads), lds_bytes, s, descs);
    return hipGetLastError();
}

template <int NM>
hipError_t launch_fold_mf(const BatchDesc *descs, int count, int mf, hipStream_t s)
{
    if (mf == 8) return launch_fold_one<8, NM>(descs, count, s);
    if (mf == 16) return launch_fold_one<16, NM>(descs, count, s);
    if (mf == 17) return launch_fold_one<17, NM>(descs, count, s);
    if (mf == 24) return launch_fold_one<24, NM>(descs, count, s);
    if (mf == 32) return launch_fold_one<32, NM>(descs, count, s);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_batch_fold(const BatchDesc *descs, int count, int mf, int nm, hipStream_t s)
{
    if (nm == 1) return launch_fold_mf<1>(descs, count, mf, s);
    if (nm == 2) return launch_fold_mf<2>(descs, count, mf, s);
    if (nm == 4) return launch_fold_mf<4>(descs, count, mf, s);
    return hipErrorInvalidValue;
}

}  // namespace deac
