/*
 * mechs.h -- membrane mechanisms on the device (the nrn_cur / nrn_state
 * equivalents of the reference's generated AllModels.cu), for three mechanism
 * sets:
 *
 *   SetPassive : pas                                     (params g, e)
 *   SetTTPC    : the 13 BBP TTPC mechanisms              (Code/OneKernel/src/AllModels.cu)
 *   SetMainen  : ca, cad, kca, km, kv, na, pas           (Figures/FigureS4_mainen/CParsed/AllModels.cu)
 *
 * The equations are the reference's; their evaluation is arranged for the
 * FP64 VALU, where an IEEE division costs about ten instructions and an exp
 * about twenty:
 *   - a gate is advanced from its rate 1/tau instead of tau:
 *       y += (1 - exp(-dt * rate)) * (yinf - y)
 *     is the reference's cnexp y + (1 - exp(dt*(-1/tau))) * ((-(yinf/tau))/(-1/tau) - y)
 *     (AllModels.cu:302) with the two quotients that cancel taken out; for
 *     alpha/beta gates rate = (alpha + beta) * q10 needs no division at all;
 *   - divisions by literal constants are products with the reciprocal, and
 *     the other quotients are products with the hardware reciprocal refined
 *     by two Newton steps (ng_rcp, ng_div: within about one ulp);
 *   - exp and log (ng_expc, ng_logc) drop the special-case branches and
 *     evaluate their polynomials by Estrin's scheme (short dependent chains);
 *   - exp(-x) next to exp(x) of the same argument is the reciprocal of exp(x);
 *     for the x/(1 - exp(-x/k)) pairs of the Na gates,
 *       beta/alpha = (b/a) exp(-x/k),  x_inf = 1 / (1 + (b/a) exp(-x/k));
 *   - q10 = 2.3^((34-21)/10) is a literal.
 * The singularity bumps (v == -27, -38, ... -> v + 1e-4) and the branches
 * (K_Pst tau at v < -50, Mainen efun/trap0 thresholds) are kept as written.
 * The results differ from the reference's operation order by rounding only;
 * the simulation is checked against the FP64 CPU oracle of the tests at
 * 1e-6 mV max-abs with identical spike counts.
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
 *                                         evaluation at the end gives the same value; where
 *                                         no block changes cai it is constant and eca =
 *                                         f(cai) already held.  Valid after the first step.
 *   S::current(mask, v, st, P, sum, ica)  BREAKPOINT currents; ica accumulated when asked
 * st[] holds the state slots (ion concentration cai and reversal eca included)
 * and P(j) returns parameter j (the set's parameter-matrix index).
 *   S::ICA, S::ica_bits(mask)             slot carrying the calcium current of the last
 *                                         BREAKPOINT to the next DERIVATIVE pass, present
 *                                         where a mechanism reads it
 * The rate functions are host- and device-callable so that a CPU test can
 * compare them with the oracle's restatement of the reference's forms.
 */
#ifndef NGPU_MECHS_H
#define NGPU_MECHS_H

#include "ngmath.h"
#include <type_traits>

#ifdef __HIPCC__
#define NG_DEV __host__ __device__ __forceinline__
#else
#include <math.h>
#define NG_DEV inline
#endif

/* Call f with the mechanism mask m: as a compile-time constant when m is one
 * of the listed masks (the mechanism code then compiles to straight-line code
 * in which the independent gates of different mechanisms interleave), else as
 * the run-time value. */
template <unsigned... MS> struct NgMasks {};
template <class F>
NG_DEV void ng_dispatch(unsigned m, F &&f, NgMasks<>) { f(m); }
template <class F, unsigned M0, unsigned... MS>
NG_DEV void ng_dispatch(unsigned m, F &&f, NgMasks<M0, MS...>)
{
    if (m == M0) f(std::integral_constant<unsigned, M0>{});
    else ng_dispatch(m, f, NgMasks<MS...>{});
}

#define NG_DEF_CAI 5.e-5
#define NG_DEF_CAO 2.
#define NG_FARADAY 96485.309

/* exp(x) for the rate functions: ng_exp's Cody-Waite reduction x = k ln2 + r
 * and its degree-13 Taylor polynomial, with the argument clamped to
 * [-708, 708] instead of branching on the range (the result stays a normal
 * number; physiological arguments are far inside).  The polynomial is
 *   1 + r (1 + r Q(r)),   Q(r) = sum_{n=2..13} r^(n-2) / n!,
 * with Q by Estrin's scheme (coefficient pairs, then r^2, r^4, r^8), so the
 * dependent chain is six multiply-adds instead of thirteen; the two leading
 * terms are added last as in Horner's rule, so the rounding stays that of
 * ng_exp's final steps (checked against ng_exp in the host tests). */
NG_DEV double ng_expc(double x)
{
    x = fmin(fmax(x, -708.0), 708.0);
    const double kf = rint(x * NG_LOG2E);
    const double r = fma(-kf, NG_LN2_LO, fma(-kf, NG_LN2_HI, x));
    const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    const double q2 = fma(1.6666666666666666667e-01, r, 0.5);                      /* 1/2! + r/3!   */
    const double q4 = fma(8.3333333333333333333e-03, r, 4.1666666666666666667e-02); /* 1/4! + r/5!   */
    const double q6 = fma(1.9841269841269841253e-04, r, 1.3888888888888888889e-03); /* 1/6! + r/7!   */
    const double q8 = fma(2.7557319223985890653e-06, r, 2.4801587301587301566e-05); /* 1/8! + r/9!   */
    const double q10 = fma(2.5052108385441718775e-08, r, 2.7557319223985890653e-07); /* 1/10! + r/11! */
    const double q12 = fma(1.6059043836821614599e-10, r, 2.0876756987868098979e-09); /* 1/12! + r/13! */
    const double s2 = fma(q4, r2, q2), s6 = fma(q8, r2, q6), s10 = fma(q12, r2, q10);
    const double Q = fma(s10, r8, fma(s6, r4, s2));
    return ldexp(fma(fma(Q, r, 1.0), r, 1.0), (int)kf);
}

/* log(x) for positive normal finite x (concentration ratios): ng_log's
 * reduction x = 2^k m, m in [sqrt(1/2), sqrt(2)), f = m - 1, s = f / (2 + f)
 * and its atanh series R(z), z = s^2, to the same degree, without the
 * special-case branches, with s from the refined reciprocal and R by
 * Estrin's scheme. */
NG_DEV double ng_logc(double x);

/* 1/y: the hardware estimate refined by two Newton steps (normal, finite y). */
NG_DEV double ng_rcp(double y)
{
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(y);
    double e = fma(-y, r, 1.0);
    r = fma(r, e, r);
    e = fma(-y, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / y;
#endif
}

/* a / b as a times the refined reciprocal of b (within about one ulp of the
 * quotient; normal, finite b). */
NG_DEV double ng_div(double a, double b) { return a * ng_rcp(b); }

NG_DEV double ng_logc(double x)
{
    int k;
    double m = frexp(x, &k);
    const bool lo = m < NG_SQRT_HALF;
    m = lo ? m * 2.0 : m;
    const double kd = (double)(lo ? k - 1 : k);
    const double f = m - 1.0;
    const double s = ng_div(f, 2.0 + f);
    const double z = s * s, z2 = z * z, z4 = z2 * z2, z8 = z4 * z4;
    const double a1 = fma(0.4, z, 0.6666666666666666666667);                          /* 2/3 + z 2/5   */
    const double a3 = fma(0.2222222222222222222222, z, 0.2857142857142857142857);     /* 2/7 + z 2/9   */
    const double a5 = fma(0.1538461538461538461538, z, 0.1818181818181818181818);     /* 2/11 + z 2/13 */
    const double a7 = fma(0.1176470588235294117647, z, 0.1333333333333333333333);     /* 2/15 + z 2/17 */
    const double a9 = fma(0.0952380952380952380952, z, 0.1052631578947368421053);     /* 2/19 + z 2/21 */
    const double a11 = 0.0869565217391304347826;                                       /* 2/23          */
    const double R = z * fma(fma(a11, z2, a9), z8, fma(fma(a7, z2, a5), z4, fma(a3, z2, a1)));
    const double hfsq = 0.5 * f * f;
    return kd * NG_LN2_HI + (f - (hfsq - (s * (hfsq + R) + kd * NG_LN2_LO)));
}

/* one cnexp step of a gate from its steady state and rate 1/tau */
NG_DEV double ng_gate(double y, double yinf, double rate, double dt)
{
    return y + (1. - ng_expc(-(dt * rate))) * (yinf - y);
}

/* ===================================================================== pas */
struct SetPassive {
    static constexpr int NMECH = 1, NPARAM = 2, NSTATE = 1;
    static constexpr int ECA = 31, ICA = 0;
    template <class F> NG_DEV static void dispatch(unsigned m, F &&f) { ng_dispatch(m, f, NgMasks<0u, 1u>{}); }
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
 * 14,15 NaTs2_t m,h | 16 SKv3_1 m | 17 SK_E2 z | 18 cai | 19 eca | 20 ica.
 * Each rate function returns (inf, rate) per gate; rate = 1/tau. */
namespace ttpc {
#define TT_KTF (1000. * 8.3134 * (34.0 + 273.15) / NG_FARADAY)
#define TT_ENA 50.0
#define TT_EK (-85.0)
#define TT_QT 2.952882641412121                  /* 2.3^((34-21)/10) */
NG_DEV double eca_of(double cai) { return TT_KTF / 2 * ng_logc(ng_div(NG_DEF_CAO, cai)); }

/* AllModels.cu:66-78 */
NG_DEV void ca_hva(double v, double &mi, double &mr, double &hi, double &hr)
{
    if (v == -27.0) v = v + 0.0001;
    const double x = -27.0 - v;
    const double mA = ng_div(0.055 * x, ng_expc(x * (1.0 / 3.8)) - 1.0);
    const double mB = 0.94 * ng_expc((-75.0 - v) * (1.0 / 17.0));
    mr = mA + mB; mi = ng_div(mA, mr);
    const double hA = 0.000457 * ng_expc((-13.0 - v) * (1.0 / 50.0));
    const double hB = 0.0065 * ng_rcp(ng_expc((-v - 15.0) * (1.0 / 28.0)) + 1.0);
    hr = hA + hB; hi = ng_div(hA, hr);
}
/* AllModels.cu:79-88; with E = e^{(v+25)/5}, q / (5 + 20 / (1 + E)) = q (1 + E) / (25 + 5 E),
 * and likewise for h */
NG_DEV void ca_lvast(double v, double &mi, double &mr, double &hi, double &hr)
{
    v = v + 10.0;
    mi = ng_rcp(1.0 + ng_expc((v + 30.0) * (-1.0 / 6.0)));
    const double Em = ng_expc((v + 25.0) * (1.0 / 5.0));
    mr = ng_div(TT_QT * (1.0 + Em), 25.0 + 5.0 * Em);
    hi = ng_rcp(1.0 + ng_expc((v + 80.0) * (1.0 / 6.4)));
    const double Eh = ng_expc((v + 40.0) * (1.0 / 7.0));
    hr = ng_div(TT_QT * (1.0 + Eh), 70.0 + 20.0 * Eh);
}
/* AllModels.cu:89-97 */
NG_DEV void ih(double v, double &mi, double &mr)
{
    if (v == -154.9) v = v + 0.0001;
    const double x = v + 154.9;
    const double mA = ng_div((0.001 * 6.43) * x, ng_expc(x * (1.0 / 11.9)) - 1.0);
    const double mB = (0.001 * 193.0) * ng_expc(v * (1.0 / 33.1));
    mr = mA + mB; mi = ng_div(mA, mr);
}
/* AllModels.cu:98-105: mA = k e^{x}, mB = k e^{-x} */
NG_DEV void im(double v, double &mi, double &mr)
{
    const double E = ng_expc((2.5 * 0.04) * (v + 35.0));
    const double Ei = ng_rcp(E);
    mi = ng_div(E, E + Ei);
    mr = (3.3e-3 * E + 3.3e-3 * Ei) * TT_QT;
}
/* AllModels.cu:106-120 */
NG_DEV void k_pst(double v, double &mi, double &mr, double &hi, double &hr)
{
    v = v + 10.0;
    mi = ng_rcp(1.0 + ng_expc(-(v + 1.0) * (1.0 / 12.0)));
    const bool lo = v < -50.0;
    mr = TT_QT * ng_rcp(1.25 + (lo ? 175.03 : 13.0) * ng_expc(v * (lo ? 0.026 : -0.026)));
    hi = ng_rcp(1.0 + ng_expc((v + 54.0) * (1.0 / 11.0)));
    const double z = (v + 75.0) * (1.0 / 48.0);
    hr = TT_QT * ng_rcp(360.0 + (1010.0 + 24.0 * (v + 55.0)) * ng_expc(-(z * z)));
}
/* AllModels.cu:121-130 */
NG_DEV void k_tst(double v, double &mi, double &mr, double &hi, double &hr)
{
    v = v + 10.0;
    mi = ng_rcp(1.0 + ng_expc(v * (-1.0 / 19.0)));
    const double zm = (v + 71.0) * (1.0 / 59.0);
    mr = TT_QT * ng_rcp(0.34 + 0.92 * ng_expc(-(zm * zm)));
    hi = ng_rcp(1.0 + ng_expc((v + 66.0) * (1.0 / 10.0)));
    const double zh = (v + 73.0) * (1.0 / 23.0);
    hr = TT_QT * ng_rcp(8.0 + 49.0 * ng_expc(-(zh * zh)));
}
/* The Na activation pair alpha = a x / (1 - e^{-x/k}), beta = b (-x) / (1 - e^{x/k}):
 * with E = e^{-x/k} and R = x / (1 - E), alpha = a R and beta = b R E. */
NG_DEV void na_pair(double x, double a, double b, double &alpha, double &beta, double &E)
{
    E = ng_expc(x * (-1.0 / 6.0));
    const double R = ng_div(x, 1.0 - E);
    alpha = a * R;
    beta = b * R * E;
}
/* AllModels.cu:131-151 */
NG_DEV void nap_et2(double v, double &mi, double &mr, double &hi, double &hr)
{
    mi = ng_rcp(1.0 + ng_expc((v + 52.6) * (-1.0 / 4.6)));
    if (v == -38.0) v = v + 0.0001;
    double mA, mB, E;
    na_pair(v + 38.0, 0.182, 0.124, mA, mB, E);
    mr = (mA + mB) * (TT_QT / 6.0);
    if (v == -17.0) v = v + 0.0001;
    if (v == -64.4) v = v + 0.0001;
    hi = ng_rcp(1.0 + ng_expc((v + 48.8) * (1.0 / 10.0)));
    const double hA = ng_div(-2.88e-6 * (v + 17.0), 1.0 - ng_expc((v + 17.0) * (1.0 / 4.63)));
    const double hB = ng_div(6.94e-6 * (v + 64.4), 1.0 - ng_expc(-(v + 64.4) * (1.0 / 2.63)));
    hr = (hA + hB) * TT_QT;
}
/* The Na inactivation pair alpha = -c y / (1 - e^{y/6}), beta = -c (-y) / (1 - e^{-y/6}):
 * with F = e^{y/6} and R = y / (1 - F), alpha = -c R, beta = -c R F and
 * alpha / (alpha + beta) = 1 / (1 + F). */
NG_DEV void na_inact(double y, double &hi, double &hr)
{
    const double F = ng_expc(y * (1.0 / 6.0));
    const double R = ng_div(y, 1.0 - F);
    hi = ng_rcp(1.0 + F);
    hr = (-0.015 * R) * (1.0 + F) * TT_QT;
}
/* AllModels.cu:152-169 */
NG_DEV void nata_t(double v, double &mi, double &mr, double &hi, double &hr)
{
    if (v == -38.0) v = v + 0.0001;
    double mA, mB, E;
    na_pair(v + 38.0, 0.182, 0.124, mA, mB, E);
    mr = (mA + mB) * TT_QT;
    mi = 0.182 * ng_rcp(0.182 + 0.124 * E);
    if (v == -66.0) v = v + 0.0001;
    na_inact(v + 66.0, hi, hr);
}
/* AllModels.cu:170-187 */
NG_DEV void nats2_t(double v, double &mi, double &mr, double &hi, double &hr)
{
    if (v == -32.0) v = v + 0.0001;
    double mA, mB, E;
    na_pair(v + 32.0, 0.182, 0.124, mA, mB, E);
    mi = 0.182 * ng_rcp(0.182 + 0.124 * E);
    mr = (mA + mB) * TT_QT;
    if (v == -60.0) v = v + 0.0001;
    na_inact(v + 60.0, hi, hr);
}
/* AllModels.cu:188-193 */
NG_DEV double sk_e2(double ca)
{
    if (ca < 1e-7) ca = ca + 1e-07;
    return ng_rcp(1.0 + ng_expc(4.8 * ng_logc(ng_div(0.00043, ca))));
}
/* AllModels.cu:194-197: tau = 4 / (1 + e^{...}) */
NG_DEV void skv3_1(double v, double &mi, double &mr)
{
    mi = ng_rcp(1.0 + ng_expc((v - 18.7) * (-1.0 / 9.7)));
    mr = (1.0 + ng_expc((v + 46.56) * (-1.0 / 44.14))) * 0.25;
}
} // namespace ttpc

struct SetTTPC {
    static constexpr int NMECH = 13, NPARAM = 19, NSTATE = 21;
    static constexpr int CAI = 18, ECA = 19, ICA = 20;
    /* the BBP cADpyr placement (Figures/Figure2_BBP_TTPC): basal {Ih, pas} and
     * apical {Ih, Im, NaTs2_t, pas, SKv3_1} hold nearly every segment; the few
     * soma and axon segments take the run-time mask code, whose register
     * demand is far below a specialisation with all thirteen mechanisms */
    template <class F> NG_DEV static void dispatch(unsigned m, F &&f)
    {
        ng_dispatch(m, f, NgMasks<0x408u, 0x1618u>{});
    }
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
    /* INITIAL blocks (AllModels.cu:202-296): gates at their steady state */
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
    /* DERIVATIVE blocks (AllModels.cu:299-377) */
    template <class PF> NG_DEV static void cnexp(unsigned m, double v, double dt, double *st, double ica, const PF &P)
    {
        using namespace ttpc;
        double a, b, c, d;
        if (m & (1u << 0)) { ca_hva(v, a, b, c, d); st[0] = ng_gate(st[0], a, b, dt); st[1] = ng_gate(st[1], c, d, dt); }
        if (m & (1u << 1)) { ca_lvast(v, a, b, c, d); st[2] = ng_gate(st[2], a, b, dt); st[3] = ng_gate(st[3], c, d, dt); }
        if (m & (1u << 2)) {
            /* cai += (1 - e^{-dt/decay}) (decay A + minCai - cai), A = -1e4 ica gamma / (2 F depth) */
            const double gamma = P(2), decay = P(3), depth = P(4), minCai = P(5);
            const double A = (-(10000.0)) * ng_div(ica * gamma, (2.0 * NG_FARADAY) * depth);
            st[CAI] = st[CAI] + (1. - ng_expc(-(dt * ng_rcp(decay)))) * ((A * decay + minCai) - st[CAI]);
        }
        if (m & (1u << 3)) { ih(v, a, b); st[4] = ng_gate(st[4], a, b, dt); }
        if (m & (1u << 4)) { im(v, a, b); st[5] = ng_gate(st[5], a, b, dt); }
        if (m & (1u << 5)) { k_pst(v, a, b, c, d); st[6] = ng_gate(st[6], a, b, dt); st[7] = ng_gate(st[7], c, d, dt); }
        if (m & (1u << 6)) { k_tst(v, a, b, c, d); st[8] = ng_gate(st[8], a, b, dt); st[9] = ng_gate(st[9], c, d, dt); }
        if (m & (1u << 7)) { nap_et2(v, a, b, c, d); st[10] = ng_gate(st[10], a, b, dt); st[11] = ng_gate(st[11], c, d, dt); }
        if (m & (1u << 8)) { nata_t(v, a, b, c, d); st[12] = ng_gate(st[12], a, b, dt); st[13] = ng_gate(st[13], c, d, dt); }
        if (m & (1u << 9)) { nats2_t(v, a, b, c, d); st[14] = ng_gate(st[14], a, b, dt); st[15] = ng_gate(st[15], c, d, dt); }
        if (m & (1u << 11)) { const double zi = sk_e2(st[CAI]); st[17] = ng_gate(st[17], zi, ng_rcp(P(17)), dt); }
        if (m & (1u << 12)) { skv3_1(v, a, b); st[16] = ng_gate(st[16], a, b, dt); }
    }
    /* BREAKPOINT blocks (AllModels.cu:382-516) */
    template <class PF>
    NG_DEV static void current(unsigned m, double v, const double *st, const PF &P, double &sum, double &ica, bool add_ica)
    {
        double s = sum, g, i;
        if (m & (1u << 0)) { g = P(0) * st[0] * st[0] * st[1]; i = g * (v - st[ECA]); s += i; if (add_ica) ica += i; }
        if (m & (1u << 1)) { g = P(1) * st[2] * st[2] * st[3]; i = g * (v - st[ECA]); s += i; if (add_ica) ica += i; }
        if (m & (1u << 3)) { g = P(6) * st[4]; i = g * (v - P(7)); s += i; }
        if (m & (1u << 4)) { g = P(8) * st[5]; i = g * (v - TT_EK); s += i; }
        if (m & (1u << 5)) { g = P(9) * st[6] * st[6] * st[7]; i = g * (v - TT_EK); s += i; }
        if (m & (1u << 6)) { const double m2 = st[8] * st[8]; g = P(10) * (m2 * m2) * st[9]; i = g * (v - TT_EK); s += i; }
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
 * cainf 1e-4, taur 200, ena 60, ek -90.  Rate functions return (inf, 1/tau). */
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
NG_DEV double eca_of(double cai) { return MN_KTF / 2 * ng_logc(ng_div(NG_DEF_CAO, cai)); }
NG_DEV double efun(double z)
{
    if (fabs(z) < 1e-4) return 1 - z * 0.5;
    return ng_div(z, ng_expc(z) - 1);
}
NG_DEV void ca(double vm, double &mi, double &mr, double &hi, double &hr)
{
    double a = 0.209 * efun(-(27.0 + vm) * (1.0 / 3.8));
    double b = 0.94 * ng_expc((-75.0 - vm) * (1.0 / 17.0));
    mr = (a + b) * MN_TADJ; mi = ng_div(a, a + b);
    a = 0.000457 * ng_expc((-13.0 - vm) * (1.0 / 50.0));
    b = 0.0065 * ng_rcp(ng_expc((-vm - 15.0) * (1.0 / 28.0)) + 1.0);
    hr = (a + b) * MN_TADJ; hi = ng_div(a, a + b);
}
NG_DEV void kca(double cai, double Ra, double Rb, double caix, double &ni, double &nr)
{
    const double a = Ra * ng_pow(cai, caix), b = Rb;
    nr = (a + b) * MN_TADJ; ni = ng_div(a, a + b);
}
/* alpha = Ra qa efun(-z), beta = Rb qa efun(z), z = (v - tha)/qa; away from
 * z = 0, efun(-z) = e^{z} efun(z) */
NG_DEV void kx(double v, double tha, double qa, double Ra, double Rb, double &ni, double &nr)
{
    const double z = ng_div(v - tha, qa);
    double ep, em;
    if (fabs(z) < 1e-4) { ep = 1 - z * 0.5; em = 1 + z * 0.5; }
    else { const double E = ng_expc(z); ep = ng_div(z, E - 1); em = E * ep; }
    const double a = Ra * qa * em, b = Rb * qa * ep;
    nr = (a + b) * MN_TADJ; ni = ng_div(a, a + b);
}
/* trap0(v, th, a, q) = a (v - th) / (1 - e^{-(v - th)/q}), or a q near v = th;
 * the m pair trap0(vm, th, Ra, q), trap0(-vm, -th, Rb, q) shares the exp:
 * with x = vm - th, E = e^{-x/q}, R = x / (1 - E): alpha = Ra R, beta = Rb R E */
template <class PF>
NG_DEV void na(double vm, const PF &P, double &mi, double &mr, double &hi, double &hr)
{
    const double tha = P(17), qa = P(18);
    const double x = vm - tha, z = ng_div(x, qa);
    double a, b;
    if (fabs(z) > 1e-6) {
        const double E = ng_expc(-z);
        const double R = ng_div(x, 1 - E);
        a = P(19) * R; b = P(20) * R * E;
    } else { a = P(19) * qa; b = P(20) * qa; }
    mr = (a + b) * MN_TADJ; mi = ng_div(a, a + b);
    const double qi = P(23), rqi = ng_rcp(qi);
    const double x1 = vm - P(21), x2 = -vm + P(22), z1 = x1 * rqi, z2 = x2 * rqi;
    a = (fabs(z1) > 1e-6) ? ng_div(P(27) * x1, 1 - ng_expc(-z1)) : P(27) * qi;
    b = (fabs(z2) > 1e-6) ? ng_div(P(26) * x2, 1 - ng_expc(-z2)) : P(26) * qi;
    hr = (a + b) * MN_TADJ;
    hi = ng_rcp(1.0 + ng_expc(ng_div(vm - P(24), P(25))));
}
} // namespace mainen

struct SetMainen {
    static constexpr int NMECH = 7, NPARAM = 30, NSTATE = 11;
    static constexpr int CAI = 8, ECA = 9, ICA = 10;
    /* dendrites {ca, cad, kca, km, na, pas} hold nearly every segment; soma
     * (all) and axon {kv, na, pas} take the run-time mask code */
    template <class F> NG_DEV static void dispatch(unsigned m, F &&f)
    {
        ng_dispatch(m, f, NgMasks<0x6fu>{});
    }
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
        if (m & 1u) { st[ECA] = eca_of(st[CAI]); ca(v + MN_VSHIFT_CA, a, b, c, d); st[0] = a; st[1] = c; }
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
        if (m & 1u) { ca(v + MN_VSHIFT_CA, a, b, c, d); st[0] = ng_gate(st[0], a, b, dt); st[1] = ng_gate(st[1], c, d, dt); }
        if (m & 2u) {
            /* ca += (1 - e^{-dt/taur}) (drive taur + cainf - ca) */
            double drive = ica * (-(10000.0) / (2.0 * NG_FARADAY * MN_DEPTH_CAD));
            if (drive <= 0.) drive = 0.;
            st[2] = st[2] + (1. - ng_expc(dt * (-1.0 / MN_TAUR))) * ((drive * MN_TAUR + MN_CAINF) - st[2]);
            st[CAI] = st[2];
        }
        if (m & 4u) { kca(st[CAI], P(4), P(5), P(3), a, b); st[3] = ng_gate(st[3], a, b, dt); }
        if (m & 8u) { kx(v, P(7), P(8), P(9), P(10), a, b); st[4] = ng_gate(st[4], a, b, dt); }
        if (m & 16u) { kx(v, P(12), P(13), P(14), P(15), a, b); st[5] = ng_gate(st[5], a, b, dt); }
        if (m & 32u) { na(v + MN_VSHIFT_NA, P, a, b, c, d); st[6] = ng_gate(st[6], a, b, dt); st[7] = ng_gate(st[7], c, d, dt); }
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
