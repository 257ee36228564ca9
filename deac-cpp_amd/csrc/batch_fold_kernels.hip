// Detailed-balance (bdsf) instantiations of the batched DEAC kernel
// (batch_device.hpp): FP64 throughout, the trials' model ISF a fold of
// rounded products on the VALU (fold_chains, register tiles of RL
// timeslices x 4 members per lane).
#include "batch_device.hpp"

namespace deac {

namespace {

template <int FS, int NM>
hipError_t launch_fold_one(const BatchDesc *descs, int count, hipStream_t s)
{
    constexpr size_t lds_bytes = static_cast<size_t>(kBatchRegionDoubles) * sizeof(double);
    static bool configured = false;
    if (!configured) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<kPrecFold, FS, NM, 0>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_bytes));
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((batch_kernel<kPrecFold, FS, NM, 0>), dim3(count), dim3(kBatchThreads), lds_bytes, s, descs);
    return hipGetLastError();
}

// shape = 2 RL + XR, the shapes of T <= 128: R = 64 RL / NM <= 128, and
// extra timeslices only past R < 128
template <int NM>
hipError_t launch_fold_shape(const BatchDesc *descs, int count, int shape, hipStream_t s)
{
    switch (shape) {
    case 2: return launch_fold_one<2, NM>(descs, count, s);
    case 3: return launch_fold_one<3, NM>(descs, count, s);
    case 4: return launch_fold_one<4, NM>(descs, count, s);
    }
    if constexpr (NM >= 2) {
        if (shape == 5) return launch_fold_one<5, NM>(descs, count, s);
        if (shape == 8) return launch_fold_one<8, NM>(descs, count, s);
    }
    if constexpr (NM >= 4) {
        if (shape == 9) return launch_fold_one<9, NM>(descs, count, s);
        if (shape == 16) return launch_fold_one<16, NM>(descs, count, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_batch_fold(const BatchDesc *descs, int count, int shape, int nm, hipStream_t s)
{
    if (nm == 1) return launch_fold_shape<1>(descs, count, shape, s);
    if (nm == 2) return launch_fold_shape<2>(descs, count, shape, s);
    if (nm == 4) return launch_fold_shape<4>(descs, count, shape, s);
    return hipErrorInvalidValue;
}

}  // namespace deac
