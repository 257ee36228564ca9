// Exact FP64 instantiations of the batched DEAC kernel (batch_device.hpp).
#include "batch_device.hpp"

namespace deac {

namespace {

template <int NT, int NM, int REM>
hipError_t launch_one(const BatchDesc *descs, int count, hipStream_t s)
{
    constexpr size_t lds_bytes = static_cast<size_t>(kBatchRegionDoubles) * sizeof(double);
    static bool configured = false;
    if (!configured) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<kPrecFp64, NT, NM, REM>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_bytes));
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((batch_kernel<kPrecFp64, NT, NM, REM>), dim3(count), dim3(kBatchThreads), lds_bytes, s, descs);
    return hipGetLastError();
}

template <int NM, int REM>
hipError_t launch_nt(const BatchDesc *descs, int count, int nt, hipStream_t s)
{
    if (nt == 1) return launch_one<1, NM, REM>(descs, count, s);
    if (nt == 2) return launch_one<2, NM, REM>(descs, count, s);
    if (nt == 4) return launch_one<4, NM, REM>(descs, count, s);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_batch(const BatchDesc *descs, int count, int nt, int nm, int rem, hipStream_t s)
{
    if (nm == 1) return rem ? launch_nt<1, 1>(descs, count, nt, s) : launch_nt<1, 0>(descs, count, nt, s);
    if (nm == 2) return rem ? launch_nt<2, 1>(descs, count, nt, s) : launch_nt<2, 0>(descs, count, nt, s);
    return hipErrorInvalidValue;
}

}  // namespace deac
