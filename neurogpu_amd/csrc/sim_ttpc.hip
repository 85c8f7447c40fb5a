/* Instantiations of the simulation kernels for the ttpc mechanism set. */
#include "sim_kernel.h"

hipError_t ngpu_launch_ttpc(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    return ng_launch_set<SetTTPC>(multi, grid, block, lds, s, a);
}

hipError_t ngpu_launch_batch_ttpc(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b)
{
    return ng_launch_batch_set<SetTTPC>(grid, lds, s, b);
}

#if NG_PHASE_CYCLES
/* The profiling build's per-phase cycle sums of this set's kernels (see
 * NG_PHASE_CYCLES in sim_kernel.h); reset clears them after reading. */
extern "C" int ngpu_phase_cycles_ttpc(unsigned long long *out, int reset)
{
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(ng_phase_cycles), sizeof(unsigned long long) * 8) != hipSuccess) return -1;
    if (reset) {
        const unsigned long long z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (hipMemcpyToSymbol(HIP_SYMBOL(ng_phase_cycles), z, sizeof z) != hipSuccess) return -1;
    }
    return 0;
}
#endif
