/* Instantiations of the simulation kernel for the mainen mechanism set. */
#include "sim_kernel.h"

template <int NS>
static hipError_t launch_ns(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    if (multi) {
        auto k = ngpu_sim_kernel<SetMainen, NS, true>;
        if (lds > 65536) { hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); if (e != hipSuccess) return e; }
        hipLaunchKernelGGL(k, grid, block, lds, s, a);
    } else {
        auto k = ngpu_sim_kernel<SetMainen, NS, false>;
        if (lds > 65536) { hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); if (e != hipSuccess) return e; }
        hipLaunchKernelGGL(k, grid, block, lds, s, a);
    }
    return hipGetLastError();
}

hipError_t ngpu_launch_mainen(int nslot, bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    switch (nslot) {
    case 1: return launch_ns<1>(multi, grid, block, lds, s, a);
    case 2: return launch_ns<2>(multi, grid, block, lds, s, a);
    case 4: return launch_ns<4>(multi, grid, block, lds, s, a);
    case 6: return launch_ns<6>(multi, grid, block, lds, s, a);
    case 8: return launch_ns<8>(multi, grid, block, lds, s, a);
    default: return hipErrorInvalidValue;
    }
}

hipError_t ngpu_launch_batch_mainen(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b)
{
    auto k = ngpu_batch_kernel<SetMainen, NG_BATCH_NSLOT, 4>;
    if (lds > 65536) { hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds); if (e != hipSuccess) return e; }
    hipLaunchKernelGGL(k, grid, dim3(NG_BATCH_THREADS), lds, s, b);
    return hipGetLastError();
}
