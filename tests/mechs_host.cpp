/*
 * mechs_host.cpp -- the device mechanism code of neurogpu_amd/csrc/mechs.h,
 * compiled for the host by tests/test_mechs_host.py so that its rearranged
 * rate functions (steady state and rate 1/tau) and gate update can be
 * compared with the oracle's restatement of the reference's forms on the CPU.
 * Test infrastructure only.
 */
#include "mechs.h"

struct HostParams {
    const double *p;
    double operator()(int j) const { return p[j]; }
};

/* out = (inf0, rate0, inf1, rate1); mech follows the set's mechanism order
 * (1 = TTPC, 2 = Mainen), as the oracle's ref_rates. */
extern "C" int host_rates(int ms, int mech, double x, const double *P, double *out)
{
    out[0] = out[1] = out[2] = out[3] = 0.0;
    if (ms == 1) {
        using namespace ttpc;
        switch (mech) {
        case 0: ca_hva(x, out[0], out[1], out[2], out[3]); return 0;
        case 1: ca_lvast(x, out[0], out[1], out[2], out[3]); return 0;
        case 3: ih(x, out[0], out[1]); return 0;
        case 4: im(x, out[0], out[1]); return 0;
        case 5: k_pst(x, out[0], out[1], out[2], out[3]); return 0;
        case 6: k_tst(x, out[0], out[1], out[2], out[3]); return 0;
        case 7: nap_et2(x, out[0], out[1], out[2], out[3]); return 0;
        case 8: nata_t(x, out[0], out[1], out[2], out[3]); return 0;
        case 9: nats2_t(x, out[0], out[1], out[2], out[3]); return 0;
        case 11: out[0] = sk_e2(x); return 0;
        case 12: skv3_1(x, out[0], out[1]); return 0;
        }
    } else if (ms == 2) {
        using namespace mainen;
        HostParams hp{P};
        switch (mech) {
        case 0: ca(x, out[0], out[1], out[2], out[3]); return 0;
        case 2: kca(x, P[4], P[5], P[3], out[0], out[1]); return 0;
        case 3: kx(x, P[7], P[8], P[9], P[10], out[0], out[1]); return 0;
        case 4: kx(x, P[12], P[13], P[14], P[15], out[0], out[1]); return 0;
        case 5: na(x, hp, out[0], out[1], out[2], out[3]); return 0;
        }
    }
    return -1;
}

extern "C" double host_gate(double y, double yinf, double rate, double dt) { return ng_gate(y, yinf, rate, dt); }
extern "C" double host_expc(double x) { return ng_expc(x); }
extern "C" double host_logc(double x) { return ng_logc(x); }
