/*
 * features.hip -- spike features of every recorded trace, computed on the
 * device right after the simulation (no trace round trip to the host).
 *
 * Stands in for the per-individual eFEL pass of the GA evaluators
 * (PYS/efel_ext.py:23-97, threshold -30 mV at :15).  The features are this
 * build's own definitions (eFEL is not available to pin them), per trace of
 * record site 0 with t_i the start time of step i.  A spike is an upward
 * crossing of thr (v[i-1] < thr <= v[i]) whose linearly interpolated time tc
 * lies in the stimulus window t_on <= tc < t_off; crossings outside the
 * window are ignored.
 *   0 spike count      spikes in the window
 *   1 first latency    first spike's tc - t_on; NaN if none
 *   2 AP amplitude     mean over spikes of (peak - v_base), v_base = mean v for t < t_on
 *   3 AP width         mean time above thr of spikes that fall back below it; NaN if none
 *   4 mean ISI         mean interval between consecutive spikes; NaN if < 2 spikes
 * A trace with a NaN sample (a diverged run) has all five NaN, the count
 * included (the evaluator scores such traces apart, PYS/efel_ext.py:47-49).
 *
 * Layout: one 64-lane workgroup takes 64 runs.  The traces are read in tiles
 * of 64 runs x 64 steps: for each run of the tile the 64 lanes read 64
 * consecutive samples (one 512-byte coalesced load), the tile is written to
 * LDS, and each lane then scans its own run's 64 samples from LDS.
 */
#include <hip/hip_runtime.h>

#include <mutex>

#define NG_NFEAT 5
#define NG_FT 64

/* t_i, the start time of step i, is the running sum dt_0 + ... + dt_{i-1}
 * of the protocol's steps, accumulated by each lane in step order. */
__global__ void __launch_bounds__(64) ngpu_features_kernel(const double *__restrict__ vhot, int nrec, int nt,
                                                           int nruns, int nstim, const double *__restrict__ dts,
                                                           const double *__restrict__ t_on,
                                                           const double *__restrict__ t_off, double thr,
                                                           double *__restrict__ out)
{
    __shared__ double tile[64][NG_FT + 1];
    const int lane = threadIdx.x;
    const int run0 = blockIdx.x * 64;
    const int run = run0 + lane;
    const bool live = run < nruns;
    const int stim = (live ? run : run0) % nstim;
    const double on = t_on[stim], off = t_off[stim];
    const double nan = __builtin_nan("");
    /* baseline: mean of the samples before the first step with t >= on */
    double vsum = 0.0, vb = 0.0;
    int nb = 0;
    bool base_open = true;
    int cnt = 0, namp = 0, nw = 0;
    double first = nan, last_up = 0.0, sum_isi = 0.0, sum_amp = 0.0, sum_w = 0.0, t_up = 0.0, peak = 0.0;
    double prev = 0.0, v0 = 0.0, tnow = 0.0, ta = 0.0;
    bool in_spike = false, diverged = false;
    for (int c0 = 0; c0 < nt; c0 += NG_FT) {
        const int n = min(NG_FT, nt - c0);
        __syncthreads();
        for (int j = 0; j < 64; j++) {
            const int r = run0 + j;
            if (r < nruns && lane < n) tile[j][lane] = vhot[(size_t)r * nrec * nt + c0 + lane];
        }
        __syncthreads();
        if (!live) continue;
        for (int k = 0; k < n; k++) {
            const int i = c0 + k;
            const double b = tile[lane][k];
            diverged |= b != b;
            const double tb = tnow;
            tnow += dts[i];
            if (i == 0) v0 = b;
            if (base_open) {
                if (tb < on) {
                    vsum += b;
                    nb++;
                } else {
                    base_open = false;
                }
                if (!base_open) vb = nb > 0 ? vsum / nb : v0;
            }
            if (i > 0) {
                const double a = prev;
                if (a < thr && b >= thr) {
                    const double tc = ta + (thr - a) / (b - a) * (tb - ta);
                    if (tc >= on && tc < off) {
                        cnt++;
                        if (cnt == 1) first = tc - on;
                        else sum_isi += tc - last_up;
                        last_up = tc;
                        t_up = tc;
                        peak = b;
                        in_spike = true;
                    }
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
            prev = b;
            ta = tb;
        }
    }
    if (!live) return;
    if (in_spike) { sum_amp += peak - vb; namp++; }
    double *o = out + (size_t)run * NG_NFEAT;
    if (diverged) {                       /* a NaN sample: every feature undefined */
        for (int k = 0; k < NG_NFEAT; k++) o[k] = nan;
        return;
    }
    o[0] = (double)cnt;
    o[1] = first;
    o[2] = namp > 0 ? sum_amp / namp : nan;
    o[3] = nw > 0 ? sum_w / nw : nan;
    o[4] = cnt > 1 ? sum_isi / (cnt - 1) : nan;
}

hipError_t ngpu_launch_features(const double *vhot, int nrec, int nt, int nruns, int nstim, const double *dts,
                                const double *t_on, const double *t_off, double thr, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(ngpu_features_kernel, dim3((nruns + 63) / 64), dim3(64), 0, s, vhot, nrec, nt, nruns, nstim,
                       dts, t_on, t_off, thr, out);
    return hipGetLastError();
}

/* Device scratch of the no-allocation feature entry point: the stimulus
 * windows and the features, staged on their way in and out.  Each device
 * holds its own copy for the life of the process; every call on a device
 * shares it, so calls that overlap on one device race. */
#define NG_FEAT_SCRATCH_DOUBLES (1 << 20)
static __device__ __align__(256) double scratch[NG_FEAT_SCRATCH_DOUBLES];

#define NG_MAX_DEVICES 64

/* The device address of `scratch` on the current device (`device`), looked
 * up once per device with hipGetSymbolAddress; cap = its size in doubles. */
hipError_t ngpu_features_scratch(int device, double **ptr, size_t *cap)
{
    static std::mutex mu;
    static double *addr[NG_MAX_DEVICES];
    if (device < 0 || device >= NG_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = (double *)p;
    }
    *ptr = addr[device];
    *cap = NG_FEAT_SCRATCH_DOUBLES;
    return hipSuccess;
}
