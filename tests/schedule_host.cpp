/*
 * schedule_host.cpp -- the mechanism-pass schedule of
 * neurogpu_amd/csrc/schedule.h compiled for the host, so that
 * tests/test_schedule_host.py can check it on the CPU.  Test infrastructure
 * only.
 */
#include "schedule.h"

/* out receives the schedule (offsets[W+1] then packed segments); returns its
 * length, or -1 when cap is too small. */
extern "C" int host_mech_schedule(const int *qmask, int N, int NQ, int mechset, int W, int R, int *out, int cap)
{
    std::vector<int> m(qmask, qmask + N);
    const std::vector<int> s = ng_mech_schedule(m, N, NQ, mechset, W, R);
    if ((int)s.size() > cap) return -1;
    for (size_t i = 0; i < s.size(); i++) out[i] = s[i];
    return (int)s.size();
}
