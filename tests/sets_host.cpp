/*
 * sets_host.cpp -- the hand-written TTPC mechanism set (neurogpu_amd/csrc/
 * mechs.h, SetTTPC) and the one generated from the model's NMODL files
 * (neurogpu_amd/csrc/gen_ttpc.h, SetGen_ttpc), compiled for the host by
 * tests/test_nmodl.py so that the two can be compared compartment by
 * compartment on the CPU.  Test infrastructure only.
 */
#include "gen_ttpc.h"

struct HostParams {
    const double *p;
    double operator()(int j) const { return p[j]; }
};

template <class S> static int run(unsigned m, int op, double v, double dt, double *st, const double *P, double *io)
{
    HostParams hp{P};
    switch (op) {
    case 0: S::init(m, v, st, hp); return 0;
    case 1: S::cnexp(m, v, dt, st, io[1], hp); S::eca_after_cnexp(m, st); return 0;
    case 2: S::current(m, v, st, hp, io[0], io[1], true); return 0;
    }
    return -1;
}

/* which: 0 hand-written, 1 generated.  op 0 init, 1 cnexp (ica = io[1]) then
 * the reversal update, 2 current (io[0] += i, io[1] += ica). */
extern "C" int set_run(int which, unsigned m, int op, double v, double dt, double *st, const double *P, double *io)
{
    return which ? run<SetGen_ttpc>(m, op, v, dt, st, P, io) : run<SetTTPC>(m, op, v, dt, st, P, io);
}

extern "C" unsigned set_state_bits(int which, unsigned m)
{
    return which ? SetGen_ttpc::state_bits(m) : SetTTPC::state_bits(m);
}

extern "C" unsigned set_ica_bits(int which, unsigned m)
{
    return which ? SetGen_ttpc::ica_bits(m) : SetTTPC::ica_bits(m);
}

extern "C" int set_shape(int which, int *out)
{
    if (which) {
        out[0] = SetGen_ttpc::NMECH; out[1] = SetGen_ttpc::NPARAM; out[2] = SetGen_ttpc::NSTATE;
        out[3] = SetGen_ttpc::CAI; out[4] = SetGen_ttpc::ECA; out[5] = SetGen_ttpc::ICA;
        for (int k = 0; k <= SetGen_ttpc::NMECH; k++) out[6 + k] = SetGen_ttpc::PBASE[k];
    } else {
        out[0] = SetTTPC::NMECH; out[1] = SetTTPC::NPARAM; out[2] = SetTTPC::NSTATE;
        out[3] = SetTTPC::CAI; out[4] = SetTTPC::ECA; out[5] = SetTTPC::ICA;
    }
    return 0;
}
