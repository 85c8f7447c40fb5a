/*
 * schedule_host.cpp -- the mechanism-pass schedule of
 * neurogpu_amd/csrc/schedule.h compiled for the host, so that
 * tests/test_schedule_host.py can check it on the CPU.  Test infrastructure
 * only.
 */
#include "schedule.h"

/* out receives the schedule (the segment count, then the packed segments);
 * returns its length, or -1 when cap is too small. */
extern "C" int host_mech_schedule(const int *qmask, int N, int NQ, int mechset, int R, int *out, int cap)
{
    std::vector<int> m(qmask, qmask + N);
    const std::vector<int> s = ng_mech_schedule(m, N, NQ, mechset, R);
    if ((int)s.size() > cap) return -1;
    for (size_t i = 0; i < s.size(); i++) out[i] = s[i];
    return (int)s.size();
}

/* ng_choose_shape with groups that fit while G*R*run_kib + table_kib <= 160
 * (G at most 8/W waves, one group for W >= 8); out = {W, R}. */
extern "C" void host_choose_shape(long nruns, int ncu, const int *ws, int nws, int r_only, double run_kib,
                                  double table_kib, int *out)
{
    const std::vector<int> w(ws, ws + nws);
    auto fit = [&](int W, int R) {
        if (R * run_kib + table_kib > 160.0) return 0;
        int G = 1;
        while ((G + 1) * W <= 8 && (G + 1) * R * run_kib + table_kib <= 160.0) G++;
        return G;
    };
    auto rmax = [](int W) { return W == 1 ? 4 : 6; };
    ng_choose_shape(nruns, ncu, w, r_only, fit, rmax, out[0], out[1]);
}
