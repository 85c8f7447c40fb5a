/*
 * params.hip -- per-set parameter tables assembled on the device.
 *
 * The GA evaluators build the whole parameter matrix on the host every
 * generation: each individual's free parameters are written into a template
 * row at the positions that hold the sentinel 12300 + i
 * (Figures/Figure4_DEAP/extractModel_mappings.py:158-171, :258-270), the
 * rows are written to AllParams.csv and read back by the simulator.  Here
 * the template is compiled once into the position tables (ngpu_set_param_template,
 * every sentinel column becomes a run-varying column), and each generation
 * uploads only its genome matrix [nsets][nfree]; this kernel scatters it into
 * the per-set table of varying columns pv [nsets][nvar] that the simulation
 * kernels read: pv[s][k] = genome[s][src[k]].
 */
#include <hip/hip_runtime.h>

__global__ void ngpu_scatter_genomes_kernel(const double *__restrict__ genome, int nsets, int nfree,
                                            const int *__restrict__ src, int nvar, double *__restrict__ pv)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)nsets * nvar) return;
    const long long s = i / nvar;
    const int k = (int)(i - s * nvar);
    pv[i] = genome[s * nfree + src[k]];
}

hipError_t ngpu_launch_scatter_genomes(const double *genome, int nsets, int nfree, const int *src, int nvar, double *pv,
                                       hipStream_t stream)
{
    const long long n = (long long)nsets * nvar;
    const int bs = 256;
    hipLaunchKernelGGL(ngpu_scatter_genomes_kernel, dim3((unsigned)((n + bs - 1) / bs)), dim3(bs), 0, stream, genome, nsets,
                       nfree, src, nvar, pv);
    return hipGetLastError();
}
