/*
 * sets_host.cpp -- the hand-written mechanism sets (neurogpu_amd/csrc/mechs.h:
 * SetTTPC, SetMainen) and the ones generated from the models' NMODL files
 * (neurogpu_amd/csrc/gen_ttpc.h, gen_mainen.h: SetGen_ttpc, SetGen_mainen),
 * compiled for the host by tests/test_nmodl.py so that each generated set
 * can be compared with its hand-written counterpart compartment by
 * compartment on the CPU.  Test infrastructure only.
 */
#include "gen_ttpc.h"
#include "gen_mainen.h"

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

template <class S> static void shape(int *out)
{
    out[0] = S::NMECH; out[1] = S::NPARAM; out[2] = S::NSTATE;
    out[3] = S::CAI; out[4] = S::ECA; out[5] = S::ICA;
}

/* which: 0 SetTTPC, 1 SetGen_ttpc, 2 SetMainen, 3 SetGen_mainen.  op 0 init,
 * 1 cnexp (ica = io[1]) then the reversal update, 2 current (io[0] += i,
 * io[1] += ica). */
extern "C" int set_run(int which, unsigned m, int op, double v, double dt, double *st, const double *P, double *io)
{
    switch (which) {
    case 0: return run<SetTTPC>(m, op, v, dt, st, P, io);
    case 1: return run<SetGen_ttpc>(m, op, v, dt, st, P, io);
    case 2: return run<SetMainen>(m, op, v, dt, st, P, io);
    case 3: return run<SetGen_mainen>(m, op, v, dt, st, P, io);
    }
    return -1;
}

extern "C" unsigned set_state_bits(int which, unsigned m)
{
    switch (which) {
    case 0: return SetTTPC::state_bits(m);
    case 1: return SetGen_ttpc::state_bits(m);
    case 2: return SetMainen::state_bits(m);
    default: return SetGen_mainen::state_bits(m);
    }
}

extern "C" unsigned set_ica_bits(int which, unsigned m)
{
    switch (which) {
    case 0: return SetTTPC::ica_bits(m);
    case 1: return SetGen_ttpc::ica_bits(m);
    case 2: return SetMainen::ica_bits(m);
    default: return SetGen_mainen::ica_bits(m);
    }
}

/* counts and ion slots; the generated sets also their parameter bases */
extern "C" int set_shape(int which, int *out)
{
    switch (which) {
    case 0: shape<SetTTPC>(out); return 0;
    case 2: shape<SetMainen>(out); return 0;
    case 1:
        shape<SetGen_ttpc>(out);
        for (int k = 0; k <= SetGen_ttpc::NMECH; k++) out[6 + k] = SetGen_ttpc::PBASE[k];
        return 0;
    case 3:
        shape<SetGen_mainen>(out);
        for (int k = 0; k <= SetGen_mainen::NMECH; k++) out[6 + k] = SetGen_mainen::PBASE[k];
        return 0;
    }
    return -1;
}
