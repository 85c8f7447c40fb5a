/* Instantiations of the simulation kernels for the mainen mechanism set. */
#include "sim_kernel.h"

hipError_t ngpu_launch_mainen(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    return ng_launch_set<SetMainen>(multi, grid, block, lds, s, a);
}

hipError_t ngpu_launch_batch_mainen(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b)
{
    return ng_launch_batch_set<SetMainen>(grid, lds, s, b);
}
