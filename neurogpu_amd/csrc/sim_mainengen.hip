/* Instantiations of the simulation kernels for the Mainen mechanism set
 * generated from the model's NMODL files (gen_mainen.h, neurogpu_amd.nmodl). */
#include "sim_kernel.h"
#include "gen_mainen.h"

hipError_t ngpu_launch_mainengen(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    return ng_launch_set<SetGen_mainen>(multi, grid, block, lds, s, a);
}

hipError_t ngpu_launch_batch_mainengen(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b)
{
    return ng_launch_batch_set<SetGen_mainen>(grid, lds, s, b);
}
