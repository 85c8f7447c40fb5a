/* Instantiations of the simulation kernels for the TTPC mechanism set
 * generated from the model's NMODL files (gen_ttpc.h, neurogpu_amd.nmodl). */
#include "sim_kernel.h"
#include "gen_ttpc.h"

hipError_t ngpu_launch_ttpcgen(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    return ng_launch_set<SetGen_ttpc>(multi, grid, block, lds, s, a);
}

hipError_t ngpu_launch_batch_ttpcgen(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b)
{
    return ng_launch_batch_set<SetGen_ttpc>(grid, lds, s, b);
}
