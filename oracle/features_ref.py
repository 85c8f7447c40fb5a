"""CPU restatement of the spike features (TEST INFRASTRUCTURE ONLY: the checker
for libneurogpu's on-device feature kernel, neurogpu_amd/csrc/features.hip).

Features replace the GA evaluator's eFEL pass (PYS/efel_ext.py:23-97,
threshold -30 mV at :15); eFEL itself is not available, so their definitions
are this build's and they are pinned by known-answer traces in the tests.
The arithmetic follows the same operation order as the device code, so the two
agree bit for bit.

A spike is an upward crossing of thr whose linearly interpolated time tc lies
in the stimulus window t_on <= tc < t_off; crossings outside it are ignored.
Width and amplitude of a spike that starts in the window are measured in full,
also when it falls back after t_off.  A trace with a NaN sample (a diverged
run) has every feature NaN, the spike count included."""
import math

import numpy as np

NFEAT = 5


def trace_features(v, times, t_on, t_off=math.inf, thr=-30.0):
    v = [float(x) for x in v]
    nt = len(v)
    if any(x != x for x in v):
        # a trace with a NaN sample (a diverged run): every feature undefined
        return np.full(NFEAT, math.nan)
    vb, nb = 0.0, 0
    for i in range(nt):
        if not times[i] < t_on:
            break
        vb += v[i]
        nb += 1
    vb = vb / nb if nb > 0 else v[0]
    cnt = namp = nw = 0
    first = math.nan
    last_up = sum_isi = sum_amp = sum_w = t_up = peak = 0.0
    in_spike = False
    for i in range(1, nt):
        a, b = v[i - 1], v[i]
        ta, tb = float(times[i - 1]), float(times[i])
        if a < thr and b >= thr:
            tc = ta + (thr - a) / (b - a) * (tb - ta)
            if t_on <= tc < t_off:
                cnt += 1
                if cnt == 1:
                    first = tc - t_on
                else:
                    sum_isi += tc - last_up
                last_up = tc
                t_up = tc
                peak = b
                in_spike = True
        elif in_spike:
            if b > peak:
                peak = b
            if a >= thr and b < thr:
                td = ta + (a - thr) / (a - b) * (tb - ta)
                sum_w += td - t_up
                nw += 1
                sum_amp += peak - vb
                namp += 1
                in_spike = False
    if in_spike:
        sum_amp += peak - vb
        namp += 1
    return np.array([float(cnt), first, sum_amp / namp if namp else math.nan,
                     sum_w / nw if nw else math.nan, sum_isi / (cnt - 1) if cnt > 1 else math.nan])


def features(vhot, dts, t_on, t_off=math.inf, thr=-30.0):
    """vhot [nsets, nstim, nrec, nt] -> [nsets, nstim, 5] for record site 0."""
    vhot = np.asarray(vhot)
    # step start times as a running sum, as the device computes them
    tt = np.empty(len(dts))
    acc = 0.0
    for i, d in enumerate(dts):
        tt[i] = acc
        acc += float(d)
    times = tt
    ons = np.broadcast_to(np.asarray(t_on, np.float64), (vhot.shape[1],))
    offs = np.broadcast_to(np.asarray(t_off, np.float64), (vhot.shape[1],))
    out = np.empty(vhot.shape[:2] + (NFEAT,))
    for s in range(vhot.shape[0]):
        for k in range(vhot.shape[1]):
            out[s, k] = trace_features(vhot[s, k, 0], times, float(ons[k]), float(offs[k]), thr)
    return out
