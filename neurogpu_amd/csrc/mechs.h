/*
 * mechs.h -- membrane mechanisms on the device (the nrn_cur / nrn_state
 * equivalents of the reference's generated AllModels.cu), for three mechanism
 * sets:
 *
 *   SetPassive : pas                                     (params g, e)
 *   SetTTPC    : the 13 BBP TTPC mechanisms              (Code/OneKernel/src/AllModels.cu)
 *   SetMainen  : ca, cad, kca, km, kv, na, pas           (Figures/FigureS4_mainen/CParsed/AllModels.cu)
 *
 * Every expression keeps the operation order of the generated code so that,
 * compiled with -ffp-contract=off and the shared ngmath.h transcendental
 * functions, the device produces the same bits as the CPU restatement.
 *
 * Interface of a set S (all per compartment, lane-local):
 *   S::NMECH, S::NPARAM, S::NSTATE        counts
 *   S::state_bits(mech_mask)              which state slots the present mechanisms use
 *   S::init(mask, v, st, P)               INITIAL blocks in mechanism order
 *   S::cnexp(mask, v, dt, st, ica, P)     DERIVATIVE (cnexp) blocks at the new v; the
 *                                         reversal eca is not touched (see S::eca_after_cnexp)
 *   S::eca_after_cnexp(mask, st)          eca after a cnexp pass.  Every DERIVATIVE block
 *                                         that changes cai ends with eca = f(cai), and the
 *                                         blocks that follow read cai only, so one
 *                                         evaluation at the end gives the same bits; where
 *                                         no block changes cai it is constant and eca =
 *                                         f(cai) already held.  Valid after the first step.
 *   S::current(mask, v, st, P, sum, ica)  BREAKPOINT currents; ica accumulated when asked
 * st[] holds the state slots (ion concentration cai and reversal eca included)
 * and P(j) returns parameter j (the set's parameter-matrix index).
 *   S::ICA, S::ica_bits(mask)             slot carrying the calcium current of the last
 *                                         BREAKPOINT to the next DERIVATIVE pass, present
 *                                         where a mechanism reads it
 */
#ifndef NGPU_MECHS_H
#define NGPU_MECHS_H

#include "ngmath.h"

#define NG_DEV __device__ __forceinline__

#define NG_DEF_CAI 5.e-5
#define NG_DEF_CAO 2.
#define NG_FARADAY 96485.309

NG_DEV double ng_cnexp(double y, double yinf, double tau, double dt)
{
    return y + (1. - ng_exp(dt * ((-1.0) / tau))) * ((-(yinf / tau)) / ((-1.0) / tau) - y);
}

/* ===================================================================== pas */
struct SetPassive {
    static constexpr int NMECH = 1, NPARAM = 2, NSTATE = 1;
    static constexpr int ECA = 31, ICA = 0;
    NG_DEV static unsigned ica_bits(unsigned) { return 0u; }
    NG_DEV static void eca_after_cnexp(unsigned, double *) {}
    NG_DEV static unsigned state_bits(unsigned) { return 0u; }
    template <class PF> NG_DEV static void init(unsigned, double, double *, const PF &) {}
    template <class PF> NG_DEV static void cnexp(unsigned, double, double, double *, double, const PF &) {}
    template <class PF>
    NG_DEV static void current(unsigned mask, double v, const double *, const PF &P, double &sum, double &, bool)
    {
        if (mask & 1u) sum += P(0) * (v - P(1));
    }
};

/* ==================================================================== TTPC */
/* Mechanism bits (alphabetical, case-insensitive): 0 Ca_HVA 1 Ca_LVAst
 * 2 CaDynamics_E2 3 Ih 4 Im 5 K_Pst 6 K_Tst 7 Nap_Et2 8 NaTa_t 9 NaTs2_t 10 pas
 * 11 SK_E2 12 SKv3_1.  State slots: 0,1 Ca_HVA m,h | 2,3 Ca_LVAst m,h | 4 Ih m |
 * 5 Im m | 6,7 K_Pst m,h | 8,9 K_Tst m,h | 10,11 Nap_Et2 m,h | 12,13 NaTa_t m,h |
 * 14,15 NaTs2_t m,h | 16 SKv3_1 m | 17 SK_E2 z | 18 cai | 19 eca | 20 ica. */
namespace ttpc {
#define TT_KTF (1000. * 8.3134 * (34.0 + 273.15) / NG_FARADAY)
#define TT_ENA 50.0
#define TT_EK (-85.0)
NG_DEV double qt() { return ng_pow(2.3, ((34.0 - 21.0) / 10.0)); }
NG_DEV double eca_of(double cai) { return TT_KTF / 2 * ng_log(NG_DEF_CAO / cai); }

NG_DEV void ca_hva(double v, double &mi, double &mt, double &hi, double &ht)
{
    if (v == -27.0) v = v + 0.0001;
    double mA = (0.055 * (-27.0 - v)) / (ng_exp((-27.0 - v) / 3.8) - 1.0);
    double mB = (0.94 * ng_exp((-75.0 - v) / 17.0));
    mi = mA / (mA + mB); mt = 1.0 / (mA + mB);
    double hA = (0.000457 * ng_exp((-13.0 - v) / 50.0));
    double hB = (0.0065 / (ng_exp((-v - 15.0) / 28.0) + 1.0));
    hi = hA / (hA + hB); ht = 1.0 / (hA + hB);
}
NG_DEV void ca_lvast(double v, double &mi, double &mt, double &hi, double &ht)
{
    double q = qt();
    v = v + 10.0;
    mi = 1.0000 / (1.0 + ng_exp((v - -30.000) / -6.0));
    mt = (5.0000 + 20.0000 / (1.0 + ng_exp((v - -25.000) / 5.0))) / q;
    hi = 1.0000 / (1.0 + ng_exp((v - -80.000) / 6.4));
    ht = (20.0000 + 50.0000 / (1.0 + ng_exp((v - -40.000) / 7.0))) / q;
}
NG_DEV void ih(double v, double &mi, double &mt)
{
    if (v == -154.9) v = v + 0.0001;
    double mA = 0.001 * 6.43 * (v + 154.9) / (ng_exp((v + 154.9) / 11.9) - 1.0);
    double mB = 0.001 * 193.0 * ng_exp(v / 33.1);
    mi = mA / (mA + mB); mt = 1.0 / (mA + mB);
}
NG_DEV void im(double v, double &mi, double &mt)
{
    double q = qt();
    double mA = 3.3e-3 * ng_exp(2.5 * 0.04 * (v - -35.0));
    double mB = 3.3e-3 * ng_exp(-2.5 * 0.04 * (v - -35.0));
    mi = mA / (mA + mB); mt = (1.0 / (mA + mB)) / q;
}
NG_DEV void k_pst(double v, double &mi, double &mt, double &hi, double &ht)
{
    double q = qt();
    v = v + 10.0;
    mi = (1.0 / (1.0 + ng_exp(-(v + 1.0) / 12.0)));
    if (v < -50.0) mt = (1.25 + 175.03 * ng_exp(-v * -0.026)) / q;
    else           mt = ((1.25 + 13.0 * ng_exp(-v * 0.026))) / q;
    hi = 1.0 / (1.0 + ng_exp(-(v + 54.0) / -11.0));
    ht = (360.0 + (1010.0 + 24.0 * (v + 55.0)) * ng_exp(-ng_pow(((v + 75.0) / 48.0), 2.0))) / q;
}
NG_DEV void k_tst(double v, double &mi, double &mt, double &hi, double &ht)
{
    double q = qt();
    v = v + 10.0;
    mi = 1.0 / (1.0 + ng_exp(-(v + 0.0) / 19.0));
    mt = (0.34 + 0.92 * ng_exp(-ng_pow(((v + 71.0) / 59.0), 2.0))) / q;
    hi = 1.0 / (1.0 + ng_exp(-(v + 66.0) / -10.0));
    ht = (8.0 + 49.0 * ng_exp(-ng_pow(((v + 73.0) / 23.0), 2.0))) / q;
}
NG_DEV void nap_et2(double v, double &mi, double &mt, double &hi, double &ht)
{
    double q = qt();
    mi = 1.0 / (1.0 + ng_exp((v - -52.6) / -4.6));
    if (v == -38.0) v = v + 0.0001;
    double mA = (0.182 * (v - -38.0)) / (1.0 - (ng_exp(-(v - -38.0) / 6.0)));
    double mB = (0.124 * (-v - 38.0)) / (1.0 - (ng_exp(-(-v - 38.0) / 6.0)));
    mt = 6.0 * (1.0 / (mA + mB)) / q;
    if (v == -17.0) v = v + 0.0001;
    if (v == -64.4) v = v + 0.0001;
    hi = 1.0 / (1.0 + ng_exp((v - -48.8) / 10.0));
    double hA = -2.88e-6 * (v + 17.0) / (1.0 - ng_exp((v + 17.0) / 4.63));
    double hB = 6.94e-6 * (v + 64.4) / (1.0 - ng_exp(-(v + 64.4) / 2.63));
    ht = (1.0 / (hA + hB)) / q;
}
NG_DEV void nata_t(double v, double &mi, double &mt, double &hi, double &ht)
{
    double q = qt();
    if (v == -38.0) v = v + 0.0001;
    double mA = (0.182 * (v - -38.0)) / (1.0 - (ng_exp(-(v - -38.0) / 6.0)));
    double mB = (0.124 * (-v - 38.0)) / (1.0 - (ng_exp(-(-v - 38.0) / 6.0)));
    mt = (1.0 / (mA + mB)) / q;
    mi = mA / (mA + mB);
    if (v == -66.0) v = v + 0.0001;
    double hA = (-0.015 * (v - -66.0)) / (1.0 - (ng_exp((v - -66.0) / 6.0)));
    double hB = (-0.015 * (-v - 66.0)) / (1.0 - (ng_exp((-v - 66.0) / 6.0)));
    ht = (1.0 / (hA + hB)) / q;
    hi = hA / (hA + hB);
}
NG_DEV void nats2_t(double v, double &mi, double &mt, double &hi, double &ht)
{
    double q = qt();
    if (v == -32.0) v = v + 0.0001;
    double mA = (0.182 * (v - -32.0)) / (1.0 - (ng_exp(-(v - -32.0) / 6.0)));
    double mB = (0.124 * (-v - 32.0)) / (1.0 - (ng_exp(-(-v - 32.0) / 6.0)));
    mi = mA / (mA + mB);
    mt = (1.0 / (mA + mB)) / q;
    if (v == -60.0) v = v + 0.0001;
    double hA = (-0.015 * (v - -60.0)) / (1.0 - (ng_exp((v - -60.0) / 6.0)));
    double hB = (-0.015 * (-v - 60.0)) / (1.0 - (ng_exp((-v - 60.0) / 6.0)));
    hi = hA / (hA + hB);
    ht = (1.0 / (hA + hB)) / q;
}
NG_DEV double sk_e2(double ca)
{
    if (ca < 1e-7) ca = ca + 1e-07;
    return 1.0 / (1.0 + ng_pow((0.00043 / ca), 4.8));
}
NG_DEV void skv3_1(double v, double &mi, double &mt)
{
    mi = 1.0 / (1.0 + ng_exp(((v - (18.700)) / (-9.700))));
    mt = 0.2 * 20.000 / (1.0 + ng_exp(((v - (-46.560)) / (-44.140))));
}
} // namespace ttpc

struct SetTTPC {
    static constexpr int NMECH = 13, NPARAM = 19, NSTATE = 21;
    static constexpr int CAI = 18, ECA = 19, ICA = 20;
    /* CaDynamics_E2 reads ica */
    NG_DEV static unsigned ica_bits(unsigned m) { return (m & (1u << 2)) ? 1u << ICA : 0u; }
    NG_DEV static void eca_after_cnexp(unsigned m, double *st)
    {
        if (m & ((1u << 0) | (1u << 1) | (1u << 2) | (1u << 11))) st[ECA] = ttpc::eca_of(st[CAI]);
    }
    NG_DEV static unsigned state_bits(unsigned m)
    {
        unsigned s = 0;
        if (m & (1u << 0)) s |= 0x3u;
        if (m & (1u << 1)) s |= 0xcu;
        if (m & (1u << 3)) s |= 1u << 4;
        if (m & (1u << 4)) s |= 1u << 5;
        if (m & (1u << 5)) s |= 0xc0u;
        if (m & (1u << 6)) s |= 0x300u;
        if (m & (1u << 7)) s |= 0xc00u;
        if (m & (1u << 8)) s |= 0x3000u;
        if (m & (1u << 9)) s |= 0xc000u;
        if (m & (1u << 12)) s |= 1u << 16;
        if (m & (1u << 11)) s |= 1u << 17;
        if (m & ((1u << 0) | (1u << 1) | (1u << 2) | (1u << 11))) s |= (1u << CAI) | (1u << ECA);
        return s;
    }
    template <class PF> NG_DEV static void init(unsigned m, double v, double *st, const PF &)
    {
        using namespace ttpc;
        double a, b, c, d;
        if (m & (1u << 0)) { st[CAI] = NG_DEF_CAI; st[ECA] = eca_of(st[CAI]); ca_hva(v, a, b, c, d); st[0] = a; st[1] = c; }
        if (m & (1u << 1)) { st[CAI] = NG_DEF_CAI; st[ECA] = eca_of(st[CAI]); ca_lvast(v, a, b, c, d); st[2] = a; st[3] = c; }
        if (m & (1u << 2)) { st[CAI] = NG_DEF_CAI; st[ECA] = eca_of(st[CAI]); }
        if (m & (1u << 3)) { ih(v, a, b); st[4] = a; }
        if (m & (1u << 4)) { im(v, a, b); st[5] = a; }
        if (m & (1u << 5)) { k_pst(v, a, b, c, d); st[6] = a; st[7] = c; }
        if (m & (1u << 6)) { k_tst(v, a, b, c, d); st[8] = a; st[9] = c; }
        if (m & (1u << 7)) { nap_et2(v, a, b, c, d); st[10] = a; st[11] = c; }
        if (m & (1u << 8)) { nata_t(v, a, b, c, d); st[12] = a; st[13] = c; }
        if (m & (1u << 9)) { nats2_t(v, a, b, c, d); st[14] = a; st[15] = c; }
        if (m & (1u << 11)) { st[ECA] = eca_of(st[CAI]); st[17] = sk_e2(st[CAI]); }
        if (m & (1u << 12)) { skv3_1(v, a, b); st[16] = a; }
    }
    template <class PF> NG_DEV static void cnexp(unsigned m, double v, double dt, double *st, double ica, const PF &P)
    {
        using namespace ttpc;
        double a, b, c, d;
        if (m & (1u << 0)) { ca_hva(v, a, b, c, d); st[0] = ng_cnexp(st[0], a, b, dt); st[1] = ng_cnexp(st[1], c, d, dt); }
        if (m & (1u << 1)) { ca_lvast(v, a, b, c, d); st[2] = ng_cnexp(st[2], a, b, dt); st[3] = ng_cnexp(st[3], c, d, dt); }
        if (m & (1u << 2)) {
            double gamma = P(2), decay = P(3), depth = P(4), minCai = P(5);
            double A = (-(10000.0)) * (((ica) * (gamma)) / (2.0 * NG_FARADAY * depth));
            double B = ((-minCai)) / decay;
            double C = (-(1.0)) / decay;
            st[CAI] = st[CAI] + (1. - ng_exp(dt * ((-(1.0)) / decay))) * ((-(A - B)) / C - st[CAI]);
        }
        if (m & (1u << 3)) { ih(v, a, b); st[4] = ng_cnexp(st[4], a, b, dt); }
        if (m & (1u << 4)) { im(v, a, b); st[5] = ng_cnexp(st[5], a, b, dt); }
        if (m & (1u << 5)) { k_pst(v, a, b, c, d); st[6] = ng_cnexp(st[6], a, b, dt); st[7] = ng_cnexp(st[7], c, d, dt); }
        if (m & (1u << 6)) { k_tst(v, a, b, c, d); st[8] = ng_cnexp(st[8], a, b, dt); st[9] = ng_cnexp(st[9], c, d, dt); }
        if (m & (1u << 7)) { nap_et2(v, a, b, c, d); st[10] = ng_cnexp(st[10], a, b, dt); st[11] = ng_cnexp(st[11], c, d, dt); }
        if (m & (1u << 8)) { nata_t(v, a, b, c, d); st[12] = ng_cnexp(st[12], a, b, dt); st[13] = ng_cnexp(st[13], c, d, dt); }
        if (m & (1u << 9)) { nats2_t(v, a, b, c, d); st[14] = ng_cnexp(st[14], a, b, dt); st[15] = ng_cnexp(st[15], c, d, dt); }
        if (m & (1u << 11)) { double zi = sk_e2(st[CAI]); st[17] = ng_cnexp(st[17], zi, P(17), dt); }
        if (m & (1u << 12)) { skv3_1(v, a, b); st[16] = ng_cnexp(st[16], a, b, dt); }
    }
    template <class PF>
    NG_DEV static void current(unsigned m, double v, const double *st, const PF &P, double &sum, double &ica, bool add_ica)
    {
        double s = sum, g, i;
        if (m & (1u << 0)) { g = P(0) * st[0] * st[0] * st[1]; i = g * (v - st[ECA]); s += i; if (add_ica) ica += i; }
        if (m & (1u << 1)) { g = P(1) * st[2] * st[2] * st[3]; i = g * (v - st[ECA]); s += i; if (add_ica) ica += i; }
        if (m & (1u << 3)) { g = P(6) * st[4]; i = g * (v - P(7)); s += i; }
        if (m & (1u << 4)) { g = P(8) * st[5]; i = g * (v - TT_EK); s += i; }
        if (m & (1u << 5)) { g = P(9) * st[6] * st[6] * st[7]; i = g * (v - TT_EK); s += i; }
        if (m & (1u << 6)) { g = P(10) * ng_pow(st[8], 4.0) * st[9]; i = g * (v - TT_EK); s += i; }
        if (m & (1u << 7)) { g = P(11) * st[10] * st[10] * st[10] * st[11]; i = g * (v - TT_ENA); s += i; }
        if (m & (1u << 8)) { g = P(12) * st[12] * st[12] * st[12] * st[13]; i = g * (v - TT_ENA); s += i; }
        if (m & (1u << 9)) { g = P(13) * st[14] * st[14] * st[14] * st[15]; i = g * (v - TT_ENA); s += i; }
        if (m & (1u << 10)) { i = P(14) * (v - P(15)); s += i; }
        if (m & (1u << 11)) { g = P(16) * st[17]; i = g * (v - TT_EK); s += i; }
        if (m & (1u << 12)) { g = P(18) * st[16]; i = g * (v - TT_EK); s += i; }
        sum = s;
    }
};

/* ================================================================== Mainen */
/* Mechanism bits: 0 ca 1 cad 2 kca 3 km 4 kv 5 na 6 pas.  Parameters:
 * ca p0 gbar p1 cao | kca p2 gbar p3 caix p4 Ra p5 Rb | km p6..p10 | kv p11..p15
 * (gbar tha qa Ra Rb) | na p16 gbar p17 tha p18 qa p19 Ra p20 Rb p21 thi1
 * p22 thi2 p23 qi p24 thinf p25 qinf p26 Rg p27 Rd | pas p28 g p29 e.
 * State slots: 0,1 ca m,h | 2 cad ca | 3 kca n | 4 km n | 5 kv n | 6,7 na m,h |
 * 8 cai | 9 eca | 10 ica.  Globals of the generated model: celsius 37, tadj
 * 3.2093639532679714, vshift_na -5, depth_cad 0.09334562733124982,
 * cainf 1e-4, taur 200, ena 60, ek -90. */
namespace mainen {
#define MN_KTF (1000. * 8.3134 * (37.0 + 273.15) / NG_FARADAY)
#define MN_TADJ 3.2093639532679714
#define MN_VSHIFT_CA 0.0
#define MN_VSHIFT_NA (-5.0)
#define MN_DEPTH_CAD 0.09334562733124982
#define MN_CAINF 0.0001
#define MN_TAUR 200.0
#define MN_ENA 60.0
#define MN_EK (-90.0)
NG_DEV double eca_of(double cai) { return MN_KTF / 2 * ng_log(NG_DEF_CAO / cai); }
NG_DEV double efun(double z)
{
    if (fabs(z) < 1e-4) return 1 - z / 2;
    return z / (ng_exp(z) - 1);
}
NG_DEV double trap0(double v, double th, double a, double q)
{
    if (fabs((v - th) / q) > 1e-6) return a * (v - th) / (1 - ng_exp(-(v - th) / q));
    return a * q;
}
NG_DEV void ca(double vm, double &mi, double &mt, double &hi, double &ht)
{
    double a = 0.209 * efun(-(27.0 + vm) / 3.8);
    double b = 0.94 * ng_exp((-75.0 - vm) / 17.0);
    mt = 1.0 / MN_TADJ / (a + b); mi = a / (a + b);
    a = 0.000457 * ng_exp((-13.0 - vm) / 50.0);
    b = 0.0065 / (ng_exp((-vm - 15.0) / 28.0) + 1.0);
    ht = 1.0 / MN_TADJ / (a + b); hi = a / (a + b);
}
NG_DEV void kca(double cai, double Ra, double Rb, double caix, double &ni, double &nt)
{
    double a = Ra * ng_pow(cai, caix), b = Rb;
    nt = 1.0 / MN_TADJ / (a + b); ni = a / (a + b);
}
NG_DEV void kx(double v, double tha, double qa, double Ra, double Rb, double &ni, double &nt)
{
    double a = Ra * qa * efun(-(v - tha) / qa);
    double b = Rb * qa * efun((v - tha) / qa);
    nt = 1.0 / MN_TADJ / (a + b); ni = a / (a + b);
}
template <class PF>
NG_DEV void na(double vm, const PF &P, double &mi, double &mt, double &hi, double &ht)
{
    double a = trap0(vm, P(17), P(19), P(18));
    double b = trap0(-vm, -P(17), P(20), P(18));
    mt = 1.0 / MN_TADJ / (a + b); mi = a / (a + b);
    a = trap0(vm, P(21), P(27), P(23));
    b = trap0(-vm, -P(22), P(26), P(23));
    ht = 1.0 / MN_TADJ / (a + b);
    hi = 1.0 / (1.0 + ng_exp((vm - P(24)) / P(25)));
}
} // namespace mainen

struct SetMainen {
    static constexpr int NMECH = 7, NPARAM = 30, NSTATE = 11;
    static constexpr int CAI = 8, ECA = 9, ICA = 10;
    /* cad reads ica */
    NG_DEV static unsigned ica_bits(unsigned m) { return (m & 2u) ? 1u << ICA : 0u; }
    NG_DEV static void eca_after_cnexp(unsigned m, double *st)
    {
        if (m & 7u) st[ECA] = mainen::eca_of(st[CAI]);
    }
    NG_DEV static unsigned state_bits(unsigned m)
    {
        unsigned s = 0;
        if (m & 1u) s |= 0x3u;
        if (m & 2u) s |= 0x4u;
        if (m & 4u) s |= 0x8u;
        if (m & 8u) s |= 0x10u;
        if (m & 16u) s |= 0x20u;
        if (m & 32u) s |= 0xc0u;
        if (m & 7u) s |= (1u << CAI) | (1u << ECA);
        return s;
    }
    template <class PF> NG_DEV static void init(unsigned m, double v, double *st, const PF &P)
    {
        using namespace mainen;
        double a, b, c, d;
        if (m & 1u) { st[ECA] = eca_of(st[CAI]); st[ECA] = eca_of(st[CAI]); ca(v + MN_VSHIFT_CA, a, b, c, d); st[0] = a; st[1] = c; }
        if (m & 2u) { st[CAI] = NG_DEF_CAI; st[ECA] = eca_of(st[CAI]); st[2] = MN_CAINF; st[CAI] = st[2]; }
        if (m & 4u) { st[ECA] = eca_of(st[CAI]); kca(st[CAI], P(4), P(5), P(3), a, b); st[3] = a; }
        if (m & 8u) { kx(v, P(7), P(8), P(9), P(10), a, b); st[4] = a; }
        if (m & 16u) { kx(v, P(12), P(13), P(14), P(15), a, b); st[5] = a; }
        if (m & 32u) { na(v + MN_VSHIFT_NA, P, a, b, c, d); st[6] = a; st[7] = c; }
    }
    template <class PF> NG_DEV static void cnexp(unsigned m, double v, double dt, double *st, double ica, const PF &P)
    {
        using namespace mainen;
        double a, b, c, d;
        if (m & 1u) { ca(v + MN_VSHIFT_CA, a, b, c, d); st[0] = ng_cnexp(st[0], a, b, dt); st[1] = ng_cnexp(st[1], c, d, dt); }
        if (m & 2u) {
            double drive = -(10000.0) * ica / (2.0 * NG_FARADAY * MN_DEPTH_CAD);
            if (drive <= 0.) drive = 0.;
            st[2] = st[2] + (1. - ng_exp(dt * ((-1.0) / MN_TAUR))) * ((-(drive + (MN_CAINF) / MN_TAUR)) / ((-1.0) / MN_TAUR) - st[2]);
            st[CAI] = st[2];
        }
        if (m & 4u) { kca(st[CAI], P(4), P(5), P(3), a, b); st[3] = ng_cnexp(st[3], a, b, dt); }
        if (m & 8u) { kx(v, P(7), P(8), P(9), P(10), a, b); st[4] = ng_cnexp(st[4], a, b, dt); }
        if (m & 16u) { kx(v, P(12), P(13), P(14), P(15), a, b); st[5] = ng_cnexp(st[5], a, b, dt); }
        if (m & 32u) { na(v + MN_VSHIFT_NA, P, a, b, c, d); st[6] = ng_cnexp(st[6], a, b, dt); st[7] = ng_cnexp(st[7], c, d, dt); }
    }
    template <class PF>
    NG_DEV static void current(unsigned m, double v, const double *st, const PF &P, double &sum, double &ica, bool add_ica)
    {
        double s = sum, g, i;
        if (m & 1u) { g = MN_TADJ * P(0) * st[0] * st[0] * st[1]; i = (1e-4) * g * (v - st[ECA]); s += i; if (add_ica) ica += i; }
        if (m & 4u) { g = MN_TADJ * P(2) * st[3]; i = (1e-4) * g * (v - MN_EK); s += i; }
        if (m & 8u) { g = MN_TADJ * P(6) * st[4]; i = (1e-4) * g * (v - MN_EK); s += i; }
        if (m & 16u) { g = MN_TADJ * P(11) * st[5]; i = (1e-4) * g * (v - MN_EK); s += i; }
        if (m & 32u) { g = MN_TADJ * P(16) * st[6] * st[6] * st[6] * st[7]; i = (1e-4) * g * (v - MN_ENA); s += i; }
        if (m & 64u) { i = P(28) * (v - P(29)); s += i; }
        sum = s;
    }
};

#endif
