/* Instantiations of the simulation kernels for the passive mechanism set. */
#include "sim_kernel.h"

hipError_t ngpu_launch_passive(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    return ng_launch_set<SetPassive>(multi, grid, block, lds, s, a);
}

hipError_t ngpu_launch_batch_passive(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b)
{
    return ng_launch_batch_set<SetPassive>(grid, lds, s, b);
}
