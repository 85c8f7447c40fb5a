/*
 * features.hip -- spike features of every recorded trace, computed on the
 * device right after the simulation (no trace round trip to the host).
 *
 * Stands in for the per-individual eFEL pass of the GA evaluators
 * (PYS/efel_ext.py:23-97, threshold -30 mV at :15).  The features are this
 * build's own definitions (eFEL is not available to pin them), per trace of
 * record site 0 with t_i the start time of step i:
 *   0 spike count      upward crossings of thr (v[i-1] < thr <= v[i])
 *   1 first latency    first crossing time (linear interpolation) - t_on; NaN if none
 *   2 AP amplitude     mean over spikes of (peak - v_base), v_base = mean v for t < t_on
 *   3 AP width         mean time above thr of spikes that fall back below it; NaN if none
 *   4 mean ISI         mean interval between consecutive crossings; NaN if < 2 spikes
 */
#include <hip/hip_runtime.h>

#define NG_NFEAT 5

__global__ void ngpu_features_kernel(const double *__restrict__ vhot, int nrec, int nt, int nruns, int nstim,
                                     const double *__restrict__ times, const double *__restrict__ t_on,
                                     const double *__restrict__ t_off, double thr, double *__restrict__ out)
{
    const int run = blockIdx.x * blockDim.x + threadIdx.x;
    if (run >= nruns) return;
    const double *v = vhot + (size_t)run * nrec * nt;
    const int stim = run % nstim;
    const double on = t_on[stim];
    (void)t_off;
    double vb = 0.0;
    int nb = 0;
    for (int i = 0; i < nt && times[i] < on; i++) { vb += v[i]; nb++; }
    vb = nb > 0 ? vb / nb : v[0];
    const double nan = __builtin_nan("");
    int cnt = 0, namp = 0, nw = 0;
    double first = nan, last_up = 0.0, sum_isi = 0.0, sum_amp = 0.0, sum_w = 0.0, t_up = 0.0, peak = 0.0;
    bool in_spike = false;
    for (int i = 1; i < nt; i++) {
        const double a = v[i - 1], b = v[i];
        const double ta = times[i - 1], tb = times[i];
        if (a < thr && b >= thr) {
            const double tc = ta + (thr - a) / (b - a) * (tb - ta);
            cnt++;
            if (cnt == 1) first = tc - on;
            else sum_isi += tc - last_up;
            last_up = tc;
            t_up = tc;
            peak = b;
            in_spike = true;
        } else if (in_spike) {
            if (b > peak) peak = b;
            if (a >= thr && b < thr) {
                const double td = ta + (a - thr) / (a - b) * (tb - ta);
                sum_w += td - t_up;
                nw++;
                sum_amp += peak - vb;
                namp++;
                in_spike = false;
            }
        }
    }
    if (in_spike) { sum_amp += peak - vb; namp++; }
    double *o = out + (size_t)run * NG_NFEAT;
    o[0] = (double)cnt;
    o[1] = first;
    o[2] = namp > 0 ? sum_amp / namp : nan;
    o[3] = nw > 0 ? sum_w / nw : nan;
    o[4] = cnt > 1 ? sum_isi / (cnt - 1) : nan;
}

hipError_t ngpu_launch_features(const double *vhot, int nrec, int nt, int nruns, int nstim, const double *times,
                                const double *t_on, const double *t_off, double thr, double *out, hipStream_t s)
{
    const int bs = 64;
    hipLaunchKernelGGL(ngpu_features_kernel, dim3((nruns + bs - 1) / bs), dim3(bs), 0, s, vhot, nrec, nt, nruns,
                       nstim, times, t_on, t_off, thr, out);
    return hipGetLastError();
}
