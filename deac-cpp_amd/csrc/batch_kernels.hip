// Exact FP64 instantiations of the batched DEAC kernel (batch_device.hpp),
// and the device scratch of the allocation-free batch entry point.
#include <mutex>

#include "batch_device.hpp"

namespace deac {

namespace {

// Statistics staging of deac_batch_run_noalloc: one array per GPU, held
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory: only the address from hipGetSymbolAddress (kept per
// device below) may reach a kernel or a copy.
static __device__ __align__(256) double scratch[kBatchScratchDoubles];
std::mutex g_scratch_lock;
double *g_scratch_ptr[kBatchMaxDevices] = {};

}  // namespace

hipError_t batch_scratch(int device, double **ptr)
{
    if (device < 0 || device >= kBatchMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_lock);
    if (!g_scratch_ptr[device]) {
        int current = -1;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess) return e;
        if (current != device) return hipErrorInvalidDevice;  // the caller makes the device current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = static_cast<double *>(p);
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

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
    // four member tiles: at most two kernel tiles per chain wave (accumulator registers)
    if constexpr (NM < 4)
        if (nt == 4) return launch_one<4, NM, REM>(descs, count, s);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_batch(const BatchDesc *descs, int count, int nt, int nm, int rem, hipStream_t s)
{
    if (nm == 1) return rem ? launch_nt<1, 1>(descs, count, nt, s) : launch_nt<1, 0>(descs, count, nt, s);
    if (nm == 2) return rem ? launch_nt<2, 1>(descs, count, nt, s) : launch_nt<2, 0>(descs, count, nt, s);
    if (nm == 4) return rem ? launch_nt<4, 1>(descs, count, nt, s) : launch_nt<4, 0>(descs, count, nt, s);
    return hipErrorInvalidValue;
}


hipError_t batch_role_cycles_fp64(unsigned long long *out, int reset)
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
