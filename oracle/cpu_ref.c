/*
 * cpu_ref.c -- FP64 CPU oracle for the NeuroGPU compartmental simulation loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the
 * smoke check and the CPU-baseline leg of the benchmark call it; the product
 * path (the HIP library) never links, loads or falls back to it.
 *
 * It restates, in plain C and in double precision, what the reference kernel
 * does for one parameter set x one stimulus:
 *   - step order, assembly and record:  Code/OneKernel/src/CudaStuff.cu:142-349
 *     (same algorithm, FP32 build: GUI/BBP_TTPC_EXAMPLE/pyNeuroGPU_unix/src/CudaStuff.cu:142-335)
 *   - triangularisation BeforeLU:        Code/OneKernel/src/CudaStuff.cu:31-69
 *   - pointer-jumping back-substitution: Code/OneKernel/src/CudaStuff.cu:71-103 (BKSUB1)
 *   - mechanisms of the BBP TTPC model:  Code/OneKernel/src/AllModels.cu:66-516
 *   - mechanisms of the Mainen model:    Figures/FigureS4_mainen/CParsed/AllModels.cu
 *   - tree-map meaning:                  NeuroGPU_Base/scripts/pyscripts/create_auxilliary_data_3.py:22-315
 *
 * Decisions where the reference is ambiguous or racy (all deliberate):
 *   - padded duplicate entries of CompByLevel32 / CompByFLevel32 are processed
 *     once (they are idempotent only under lock-step SIMT);
 *   - sibling scatter into a parent is a sequential sum in segment order;
 *   - the previous-step solution x is 0 before step 0 (dv starts at 0);
 *   - every (set, stimulus) run owns its mechanism state;
 *   - arithmetic is pure FP64: float-suffixed constants are taken at their
 *     decimal double value and powf is pow;
 *   - the trace is unrotated: vhot[t] is v at the record site before step t.
 *
 * How each part is pinned (tests/test_oracle.py, tests/test_formats_tree.py):
 *   - the solve: dense linear solves of generated trees, the one-compartment
 *     backward-Euler closed form, rest and identical-instance invariants;
 *   - the rate functions and currents: the mechanisms' NMODL rates
 *     procedures and BREAKPOINT blocks evaluated from the .mod sources
 *     (tests/golden/mod_rates_*.npz, mod_currents.npz), to 1e-10 and 1e-12;
 *   - the state updates: the exact solution over dt of each mechanism's
 *     NMODL DERIVATIVE block (tests/golden/mod_derivs_{ttpc,mainen}.npz);
 *   - cnexp: the exact per-step contraction at fixed v; the Nernst constant;
 *   - the tree maps: integer equality with four reference BasicConstSegP.csv;
 *   - end to end: the reference's TTPC fixture within Code/test_output.py's
 *     bounds.
 *
 * Build: see oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp -shared).
 */
#include <math.h>
#include "ngmath.h"
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EPS_V 0.001           /* CANON/Util.h:22 */
#define DEF_CAI 5.e-5         /* Code/OneKernel/src/AllModels.cu:22 */
#define DEF_CAO 2.            /* Code/OneKernel/src/AllModels.cu:23 */
#define FARADAY 96485.309     /* AllModels.cu:12 (float literal taken as its decimal value) */

enum { MS_PASSIVE = 0, MS_TTPC = 1, MS_MAINEN = 2 };

/* Tree data as held in BasicConstSegP.csv (Appendix of the reference formats;
 * read by CANON/MainC.cu:335-551).  All index arrays keep the file's 1-based
 * conventions. */
typedef struct {
    int N, ncomps, mechset, nmodels;
    const double *e, *f, *cm;        /* e,f in Hines-row order; cm per segment */
    const int *ks;                   /* [N]   parent row (1-based) of row i   */
    const int *seg_to_comp;          /* [N]   */
    const int *bool_model;           /* [nmodels*N] mechanism-major           */
    const int *son_no;               /* [N]   SonNoVec                         */
    int depth, logn_depth, n_fathers, n_call;
    const int *rel_starts, *rel_ends, *rel_vec;   /* [nFathers],[nFathers],[nCall]    */
    const int *seg_start, *seg_end, *fathers;     /* [nCall+1],[nCall+1],[nFathers]   */
    const int *fidxs;                              /* [logn_depth*N]                   */
    int comp_depth, comp_fdepth;
    const int *comp_by_level32, *comp_by_flevel32; /* [comp_depth*32],[comp_fdepth*32] */
    int n_lrel, n_flrel;
    const int *lrel_starts, *lrel_ends, *flrel_starts, *flrel_ends;
} ref_tree;

/* ------------------------------------------------------------------------ */
/* Mechanisms.  Each mechanism restates the generated device code with the
 * same operation order.  States are indexed as the reference's state_macro
 * slots; ion values (cai, eca, ica) are per-segment registers as in the
 * reference kernel. */

#define MAXST 19

static double cnexp1(double y, double yinf, double tau, double dt)
{
    /* y + (1 - exp(dt*(-1/tau))) * (-(yinf/tau)/(-1/tau) - y)  (AllModels.cu:302) */
    return y + (1. - ng_exp(dt * ((-1.0) / tau))) * ((-(yinf / tau)) / ((-1.0) / tau) - y);
}

/* ---- BBP TTPC set (Code/OneKernel/src/AllModels.cu) ---- */
#define TT_CELSIUS 34.0
#define TT_KTF (1000. * 8.3134 * (TT_CELSIUS + 273.15) / FARADAY)
#define TT_ENA 50.0
#define TT_EK (-85.0)
static double tt_qt(void) { return ng_pow(2.3, ((34.0 - 21.0) / 10.0)); }
static double tt_eca(double cai) { return TT_KTF / 2 * ng_log(DEF_CAO / cai); }

static void r_ca_hva(double v, double *mi, double *mt, double *hi, double *ht)
{   /* AllModels.cu:66-78 */
    if (v == -27.0) v = v + 0.0001;
    double mA = (0.055 * (-27.0 - v)) / (ng_exp((-27.0 - v) / 3.8) - 1.0);
    double mB = (0.94 * ng_exp((-75.0 - v) / 17.0));
    *mi = mA / (mA + mB); *mt = 1.0 / (mA + mB);
    double hA = (0.000457 * ng_exp((-13.0 - v) / 50.0));
    double hB = (0.0065 / (ng_exp((-v - 15.0) / 28.0) + 1.0));
    *hi = hA / (hA + hB); *ht = 1.0 / (hA + hB);
}
static void r_ca_lvast(double v, double *mi, double *mt, double *hi, double *ht)
{   /* AllModels.cu:79-88 */
    double qt = tt_qt();
    v = v + 10.0;
    *mi = 1.0000 / (1.0 + ng_exp((v - -30.000) / -6.0));
    *mt = (5.0000 + 20.0000 / (1.0 + ng_exp((v - -25.000) / 5.0))) / qt;
    *hi = 1.0000 / (1.0 + ng_exp((v - -80.000) / 6.4));
    *ht = (20.0000 + 50.0000 / (1.0 + ng_exp((v - -40.000) / 7.0))) / qt;
}
static void r_ih(double v, double *mi, double *mt)
{   /* AllModels.cu:89-97 */
    if (v == -154.9) v = v + 0.0001;
    double mA = 0.001 * 6.43 * (v + 154.9) / (ng_exp((v + 154.9) / 11.9) - 1.0);
    double mB = 0.001 * 193.0 * ng_exp(v / 33.1);
    *mi = mA / (mA + mB); *mt = 1.0 / (mA + mB);
}
static void r_im(double v, double *mi, double *mt)
{   /* AllModels.cu:98-105 */
    double qt = tt_qt();
    double mA = 3.3e-3 * ng_exp(2.5 * 0.04 * (v - -35.0));
    double mB = 3.3e-3 * ng_exp(-2.5 * 0.04 * (v - -35.0));
    *mi = mA / (mA + mB); *mt = (1.0 / (mA + mB)) / qt;
}
static void r_k_pst(double v, double *mi, double *mt, double *hi, double *ht)
{   /* AllModels.cu:106-120 */
    double qt = tt_qt();
    v = v + 10.0;
    *mi = (1.0 / (1.0 + ng_exp(-(v + 1.0) / 12.0)));
    if (v < -50.0) *mt = (1.25 + 175.03 * ng_exp(-v * -0.026)) / qt;
    else           *mt = ((1.25 + 13.0 * ng_exp(-v * 0.026))) / qt;
    *hi = 1.0 / (1.0 + ng_exp(-(v + 54.0) / -11.0));
    *ht = (360.0 + (1010.0 + 24.0 * (v + 55.0)) * ng_exp(-ng_pow(((v + 75.0) / 48.0), 2.0))) / qt;
}
static void r_k_tst(double v, double *mi, double *mt, double *hi, double *ht)
{   /* AllModels.cu:121-130 */
    double qt = tt_qt();
    v = v + 10.0;
    *mi = 1.0 / (1.0 + ng_exp(-(v + 0.0) / 19.0));
    *mt = (0.34 + 0.92 * ng_exp(-ng_pow(((v + 71.0) / 59.0), 2.0))) / qt;
    *hi = 1.0 / (1.0 + ng_exp(-(v + 66.0) / -10.0));
    *ht = (8.0 + 49.0 * ng_exp(-ng_pow(((v + 73.0) / 23.0), 2.0))) / qt;
}
static void r_nap_et2(double v, double *mi, double *mt, double *hi, double *ht)
{   /* AllModels.cu:131-151 */
    double qt = tt_qt();
    *mi = 1.0 / (1.0 + ng_exp((v - -52.6) / -4.6));
    if (v == -38.0) v = v + 0.0001;
    double mA = (0.182 * (v - -38.0)) / (1.0 - (ng_exp(-(v - -38.0) / 6.0)));
    double mB = (0.124 * (-v - 38.0)) / (1.0 - (ng_exp(-(-v - 38.0) / 6.0)));
    *mt = 6.0 * (1.0 / (mA + mB)) / qt;
    if (v == -17.0) v = v + 0.0001;
    if (v == -64.4) v = v + 0.0001;
    *hi = 1.0 / (1.0 + ng_exp((v - -48.8) / 10.0));
    double hA = -2.88e-6 * (v + 17.0) / (1.0 - ng_exp((v + 17.0) / 4.63));
    double hB = 6.94e-6 * (v + 64.4) / (1.0 - ng_exp(-(v + 64.4) / 2.63));
    *ht = (1.0 / (hA + hB)) / qt;
}
static void r_nata_t(double v, double *mi, double *mt, double *hi, double *ht)
{   /* AllModels.cu:152-169 */
    double qt = tt_qt();
    if (v == -38.0) v = v + 0.0001;
    double mA = (0.182 * (v - -38.0)) / (1.0 - (ng_exp(-(v - -38.0) / 6.0)));
    double mB = (0.124 * (-v - 38.0)) / (1.0 - (ng_exp(-(-v - 38.0) / 6.0)));
    *mt = (1.0 / (mA + mB)) / qt;
    *mi = mA / (mA + mB);
    if (v == -66.0) v = v + 0.0001;
    double hA = (-0.015 * (v - -66.0)) / (1.0 - (ng_exp((v - -66.0) / 6.0)));
    double hB = (-0.015 * (-v - 66.0)) / (1.0 - (ng_exp((-v - 66.0) / 6.0)));
    *ht = (1.0 / (hA + hB)) / qt;
    *hi = hA / (hA + hB);
}
static void r_nats2_t(double v, double *mi, double *mt, double *hi, double *ht)
{   /* AllModels.cu:170-187 */
    double qt = tt_qt();
    if (v == -32.0) v = v + 0.0001;
    double mA = (0.182 * (v - -32.0)) / (1.0 - (ng_exp(-(v - -32.0) / 6.0)));
    double mB = (0.124 * (-v - 32.0)) / (1.0 - (ng_exp(-(-v - 32.0) / 6.0)));
    *mi = mA / (mA + mB);
    *mt = (1.0 / (mA + mB)) / qt;
    if (v == -60.0) v = v + 0.0001;
    double hA = (-0.015 * (v - -60.0)) / (1.0 - (ng_exp((v - -60.0) / 6.0)));
    double hB = (-0.015 * (-v - 60.0)) / (1.0 - (ng_exp((-v - 60.0) / 6.0)));
    *hi = hA / (hA + hB);
    *ht = (1.0 / (hA + hB)) / qt;
}
static double r_sk_e2(double ca)
{   /* AllModels.cu:188-193 */
    if (ca < 1e-7) ca = ca + 1e-07;
    return 1.0 / (1.0 + ng_pow((0.00043 / ca), 4.8));
}
static void r_skv3_1(double v, double *mi, double *mt)
{   /* AllModels.cu:194-197 */
    *mi = 1.0 / (1.0 + ng_exp(((v - (18.700)) / (-9.700))));
    *mt = 0.2 * 20.000 / (1.0 + ng_exp(((v - (-46.560)) / (-44.140))));
}

/* TTPC mechanism order (alphabetical, case-insensitive; PYS/extractModel.py:327)
 * and parameter layout (CudaStuff.cu:234 param_macro indices):
 *  0 Ca_HVA(p0)  1 Ca_LVAst(p1)  2 CaDynamics_E2(p2..p5: gamma,decay,depth,minCai)
 *  3 Ih(p6,p7: gIhbar,ehcn) 4 Im(p8) 5 K_Pst(p9) 6 K_Tst(p10) 7 Nap_Et2(p11)
 *  8 NaTa_t(p12) 9 NaTs2_t(p13) 10 pas(p14,p15: g,e) 11 SK_E2(p16,p17: gbar,zTau)
 * 12 SKv3_1(p18).  State slots as state_macro(k): 1,2 Ca_HVA; 3,4 Ca_LVAst; 5 Ih;
 *  6 Im; 7,8 K_Pst; 9,10 K_Tst; 11,12 Nap_Et2; 13,14 NaTa_t; 15,16 NaTs2_t;
 * 17 SKv3_1; 18 SK_E2. */
typedef struct { double cai, eca, ica; } ionreg;

static void tt_init(const int *has, const double *P, double v, double *st, ionreg *io)
{   /* CudaStuff.cu:234 (CALL_TO_INIT_STATES_CU), AllModels.cu:202-296 */
    double a, b, c, d;
    if (has[0]) { io->cai = DEF_CAI; io->eca = tt_eca(io->cai); r_ca_hva(v, &a, &b, &c, &d); st[1] = a; st[2] = c; }
    if (has[1]) { io->cai = DEF_CAI; io->eca = tt_eca(io->cai); r_ca_lvast(v, &a, &b, &c, &d); st[3] = a; st[4] = c; }
    if (has[2]) { io->cai = DEF_CAI; io->eca = tt_eca(io->cai); }
    if (has[3]) { r_ih(v, &a, &b); st[5] = a; }
    if (has[4]) { r_im(v, &a, &b); st[6] = a; }
    if (has[5]) { r_k_pst(v, &a, &b, &c, &d); st[7] = a; st[8] = c; }
    if (has[6]) { r_k_tst(v, &a, &b, &c, &d); st[9] = a; st[10] = c; }
    if (has[7]) { r_nap_et2(v, &a, &b, &c, &d); st[11] = a; st[12] = c; }
    if (has[8]) { r_nata_t(v, &a, &b, &c, &d); st[13] = a; st[14] = c; }
    if (has[9]) { r_nats2_t(v, &a, &b, &c, &d); st[15] = a; st[16] = c; }
    if (has[11]) { io->eca = tt_eca(io->cai); st[18] = r_sk_e2(io->cai); }
    if (has[12]) { r_skv3_1(v, &a, &b); st[17] = a; }
    (void)P;
}

static void tt_cur(const int *has, const double *P, double v, const double *st,
                   ionreg *io, int add_ica, double *sum)
{   /* CudaStuff.cu:277-278, AllModels.cu:382-516 */
    double s = *sum, g, i;
    if (has[0]) { g = P[0] * st[1] * st[1] * st[2]; i = g * (v - io->eca); s += i; if (add_ica) io->ica += i; }
    if (has[1]) { g = P[1] * st[3] * st[3] * st[4]; i = g * (v - io->eca); s += i; if (add_ica) io->ica += i; }
    if (has[3]) { g = P[6] * st[5]; i = g * (v - P[7]); s += i; }
    if (has[4]) { g = P[8] * st[6]; i = g * (v - TT_EK); s += i; }
    if (has[5]) { g = P[9] * st[7] * st[7] * st[8]; i = g * (v - TT_EK); s += i; }
    if (has[6]) { g = P[10] * ng_pow(st[9], 4.0) * st[10]; i = g * (v - TT_EK); s += i; }
    if (has[7]) { g = P[11] * st[11] * st[11] * st[11] * st[12]; i = g * (v - TT_ENA); s += i; }
    if (has[8]) { g = P[12] * st[13] * st[13] * st[13] * st[14]; i = g * (v - TT_ENA); s += i; }
    if (has[9]) { g = P[13] * st[15] * st[15] * st[15] * st[16]; i = g * (v - TT_ENA); s += i; }
    if (has[10]) { i = P[14] * (v - P[15]); s += i; }
    if (has[11]) { g = P[16] * st[18]; i = g * (v - TT_EK); s += i; }
    if (has[12]) { g = P[18] * st[17]; i = g * (v - TT_EK); s += i; }
    *sum = s;
}

static void tt_deriv(const int *has, const double *P, double dt, double v, double *st, ionreg *io)
{   /* CudaStuff.cu:339, AllModels.cu:299-377 */
    double a, b, c, d;
    if (has[0]) { r_ca_hva(v, &a, &b, &c, &d); st[1] = cnexp1(st[1], a, b, dt); st[2] = cnexp1(st[2], c, d, dt); }
    if (has[1]) { r_ca_lvast(v, &a, &b, &c, &d); st[3] = cnexp1(st[3], a, b, dt); st[4] = cnexp1(st[4], c, d, dt); }
    if (has[2]) {
        double gamma = P[2], decay = P[3], depth = P[4], minCai = P[5];
        double A = (-(10000.0)) * (((io->ica) * (gamma)) / (2.0 * FARADAY * depth));
        double B = ((-minCai)) / decay;
        double C = (-(1.0)) / decay;
        io->cai = io->cai + (1. - ng_exp(dt * ((-(1.0)) / decay))) * ((-(A - B)) / C - io->cai);
        io->eca = tt_eca(io->cai);
    }
    if (has[3]) { r_ih(v, &a, &b); st[5] = cnexp1(st[5], a, b, dt); }
    if (has[4]) { r_im(v, &a, &b); st[6] = cnexp1(st[6], a, b, dt); }
    if (has[5]) { r_k_pst(v, &a, &b, &c, &d); st[7] = cnexp1(st[7], a, b, dt); st[8] = cnexp1(st[8], c, d, dt); }
    if (has[6]) { r_k_tst(v, &a, &b, &c, &d); st[9] = cnexp1(st[9], a, b, dt); st[10] = cnexp1(st[10], c, d, dt); }
    if (has[7]) { r_nap_et2(v, &a, &b, &c, &d); st[11] = cnexp1(st[11], a, b, dt); st[12] = cnexp1(st[12], c, d, dt); }
    if (has[8]) { r_nata_t(v, &a, &b, &c, &d); st[13] = cnexp1(st[13], a, b, dt); st[14] = cnexp1(st[14], c, d, dt); }
    if (has[9]) { r_nats2_t(v, &a, &b, &c, &d); st[15] = cnexp1(st[15], a, b, dt); st[16] = cnexp1(st[16], c, d, dt); }
    if (has[11]) { double zi = r_sk_e2(io->cai); st[18] = cnexp1(st[18], zi, P[17], dt); io->eca = tt_eca(io->cai); }
    if (has[12]) { r_skv3_1(v, &a, &b); st[17] = cnexp1(st[17], a, b, dt); }
}

/* ---- Mainen & Sejnowski set (Figures/FigureS4_mainen/CParsed/AllModels.cu) ----
 * Order ca, cad, kca, km, kv, na, pas.  Parameter layout:
 *  ca: p0 gbar, p1 cao | kca: p2 gbar, p3 caix, p4 Ra, p5 Rb |
 *  km: p6 gbar, p7 tha, p8 qa, p9 Ra, p10 Rb | kv: p11..p15 (same) |
 *  na: p16 gbar, p17 tha, p18 qa, p19 Ra, p20 Rb, p21 thi1, p22 thi2, p23 qi,
 *      p24 thinf, p25 qinf, p26 Rg, p27 Rd | pas: p28 g, p29 e.
 * States: 1,2 ca(m,h); 3 cad(ca); 4 kca(n); 5 km(n); 6 kv(n); 7,8 na(m,h). */
#define MN_CELSIUS 37.0
#define MN_KTF (1000. * 8.3134 * (MN_CELSIUS + 273.15) / FARADAY)
#define MN_TADJ 3.2093639532679714
#define MN_VSHIFT_CA 0.0
#define MN_VSHIFT_NA (-5.0)
#define MN_DEPTH_CAD 0.09334562733124982
#define MN_CAINF 0.0001
#define MN_TAUR 200.0
#define MN_ENA 60.0
#define MN_EK (-90.0)
static double mn_eca(double cai) { return MN_KTF / 2 * ng_log(DEF_CAO / cai); }
static double mn_efun(double z)
{
    if (fabs(z) < 1e-4) return 1 - z / 2;
    return z / (ng_exp(z) - 1);
}
static double mn_trap0(double v, double th, double a, double q)
{
    if (fabs((v - th) / q) > 1e-6) return a * (v - th) / (1 - ng_exp(-(v - th) / q));
    return a * q;
}
static void r_mn_ca(double vm, double *mi, double *mt, double *hi, double *ht)
{
    double a = 0.209 * mn_efun(-(27.0 + vm) / 3.8);
    double b = 0.94 * ng_exp((-75.0 - vm) / 17.0);
    *mt = 1.0 / MN_TADJ / (a + b); *mi = a / (a + b);
    a = 0.000457 * ng_exp((-13.0 - vm) / 50.0);
    b = 0.0065 / (ng_exp((-vm - 15.0) / 28.0) + 1.0);
    *ht = 1.0 / MN_TADJ / (a + b); *hi = a / (a + b);
}
static void r_mn_kca(double cai, double Ra, double Rb, double caix, double *ni, double *nt)
{
    double a = Ra * ng_pow(cai, caix), b = Rb;
    *nt = 1.0 / MN_TADJ / (a + b); *ni = a / (a + b);
}
static void r_mn_kx(double v, double tha, double qa, double Ra, double Rb, double *ni, double *nt)
{
    double a = Ra * qa * mn_efun(-(v - tha) / qa);
    double b = Rb * qa * mn_efun((v - tha) / qa);
    *nt = 1.0 / MN_TADJ / (a + b); *ni = a / (a + b);
}
static void r_mn_na(double vm, const double *P, double *mi, double *mt, double *hi, double *ht)
{
    double a = mn_trap0(vm, P[17], P[19], P[18]);
    double b = mn_trap0(-vm, -P[17], P[20], P[18]);
    *mt = 1.0 / MN_TADJ / (a + b); *mi = a / (a + b);
    a = mn_trap0(vm, P[21], P[27], P[23]);
    b = mn_trap0(-vm, -P[22], P[26], P[23]);
    *ht = 1.0 / MN_TADJ / (a + b);
    *hi = 1.0 / (1.0 + ng_exp((vm - P[24]) / P[25]));
}

static void mn_init(const int *has, const double *P, double v, double *st, ionreg *io)
{
    double a, b, c, d;
    if (has[0]) { io->eca = mn_eca(io->cai); io->eca = mn_eca(io->cai); r_mn_ca(v + MN_VSHIFT_CA, &a, &b, &c, &d); st[1] = a; st[2] = c; }
    if (has[1]) { io->cai = DEF_CAI; io->eca = mn_eca(io->cai); st[3] = MN_CAINF; io->cai = st[3]; }
    if (has[2]) { io->eca = mn_eca(io->cai); r_mn_kca(io->cai, P[4], P[5], P[3], &a, &b); st[4] = a; }
    if (has[3]) { r_mn_kx(v, P[7], P[8], P[9], P[10], &a, &b); st[5] = a; }
    if (has[4]) { r_mn_kx(v, P[12], P[13], P[14], P[15], &a, &b); st[6] = a; }
    if (has[5]) { r_mn_na(v + MN_VSHIFT_NA, P, &a, &b, &c, &d); st[7] = a; st[8] = c; }
}

static void mn_cur(const int *has, const double *P, double v, const double *st,
                   ionreg *io, int add_ica, double *sum)
{
    double s = *sum, g, i;
    if (has[0]) { g = MN_TADJ * P[0] * st[1] * st[1] * st[2]; i = (1e-4) * g * (v - io->eca); s += i; if (add_ica) io->ica += i; }
    if (has[2]) { g = MN_TADJ * P[2] * st[4]; i = (1e-4) * g * (v - MN_EK); s += i; }
    if (has[3]) { g = MN_TADJ * P[6] * st[5]; i = (1e-4) * g * (v - MN_EK); s += i; }
    if (has[4]) { g = MN_TADJ * P[11] * st[6]; i = (1e-4) * g * (v - MN_EK); s += i; }
    if (has[5]) { g = MN_TADJ * P[16] * st[7] * st[7] * st[7] * st[8]; i = (1e-4) * g * (v - MN_ENA); s += i; }
    if (has[6]) { i = P[28] * (v - P[29]); s += i; }
    *sum = s;
}

static void mn_deriv(const int *has, const double *P, double dt, double v, double *st, ionreg *io)
{
    double a, b, c, d;
    if (has[0]) { r_mn_ca(v + MN_VSHIFT_CA, &a, &b, &c, &d); st[1] = cnexp1(st[1], a, b, dt); st[2] = cnexp1(st[2], c, d, dt); io->eca = mn_eca(io->cai); }
    if (has[1]) {
        double drive = -(10000.0) * io->ica / (2.0 * FARADAY * MN_DEPTH_CAD);
        if (drive <= 0.) drive = 0.;
        st[3] = st[3] + (1. - ng_exp(dt * ((-1.0) / MN_TAUR))) * ((-(drive + (MN_CAINF) / MN_TAUR)) / ((-1.0) / MN_TAUR) - st[3]);
        io->cai = st[3];
        io->eca = mn_eca(io->cai);
    }
    if (has[2]) { r_mn_kca(io->cai, P[4], P[5], P[3], &a, &b); st[4] = cnexp1(st[4], a, b, dt); io->eca = mn_eca(io->cai); }
    if (has[3]) { r_mn_kx(v, P[7], P[8], P[9], P[10], &a, &b); st[5] = cnexp1(st[5], a, b, dt); }
    if (has[4]) { r_mn_kx(v, P[12], P[13], P[14], P[15], &a, &b); st[6] = cnexp1(st[6], a, b, dt); }
    if (has[5]) { r_mn_na(v + MN_VSHIFT_NA, P, &a, &b, &c, &d); st[7] = cnexp1(st[7], a, b, dt); st[8] = cnexp1(st[8], c, d, dt); }
}

/* ---- passive set: pas only, params p0 g, p1 e ---- */
static void ps_cur(const int *has, const double *P, double v, double *sum)
{
    if (has[0]) *sum += P[0] * (v - P[1]);
}

int ref_mechset_nmodels(int ms) { return ms == MS_TTPC ? 13 : ms == MS_MAINEN ? 7 : ms == MS_PASSIVE ? 1 : -1; }
int ref_mechset_nparams(int ms) { return ms == MS_TTPC ? 19 : ms == MS_MAINEN ? 30 : ms == MS_PASSIVE ? 2 : -1; }
static const int TT_PBASE[14] = {0, 1, 2, 6, 8, 9, 10, 11, 12, 13, 14, 16, 18, 19};
static const int MN_PBASE[8] = {0, 2, 2, 6, 11, 16, 28, 30};
static const int PS_PBASE[2] = {0, 2};

/* ------------------------------------------------------------------------ */
/* Branched Hines solve on (u, b) in Hines-row order, as the reference does it. */

/* BeforeLU (CudaStuff.cu:31-69): level by level, each run eliminates its rows
 * serially; after each level but the last, each father at FLevel = level+1
 * absorbs its related run ends.  Duplicate padded lanes are visited once. */
static void before_lu(const ref_tree *T, double *u, double *b)
{
    const double *E = T->e, *F = T->f;
    for (int L = 0; L <= T->depth; L++) {
        int lo = T->lrel_starts[L], hi = T->lrel_ends[L];
        for (int r = lo; r <= hi; r++) {
            for (int lane = 0; lane < 32; lane++) {
                int J = T->comp_by_level32[r * 32 + lane] - 1;
                int dup = 0;
                for (int q = lo * 32; q < r * 32 + lane; q++)
                    if (T->comp_by_level32[q] - 1 == J) { dup = 1; break; }
                if (dup) continue;
                for (int i = T->seg_start[J] - 1; i < T->seg_end[J]; i++) {
                    double um1 = u[i - 1];
                    u[i] = u[i] - F[i - 1] * (E[i - 1] / um1);
                    um1 = u[i - 1];
                    double bm1 = b[i - 1];
                    b[i] = b[i] - bm1 * E[i - 1] / um1;
                }
            }
        }
        if (L < T->depth) {
            int flo = T->flrel_starts[L], fhi = T->flrel_ends[L];
            for (int r = flo; r <= fhi; r++) {
                for (int lane = 0; lane < 32; lane++) {
                    int CurB = T->comp_by_flevel32[r * 32 + lane] - 1;
                    int dup = 0;
                    for (int q = flo * 32; q < r * 32 + lane; q++)
                        if (T->comp_by_flevel32[q] - 1 == CurB) { dup = 1; break; }
                    if (dup) continue;
                    int J = T->fathers[CurB] - 1;
                    for (int j = T->rel_starts[CurB]; j <= T->rel_ends[CurB]; j++) {
                        int t = T->rel_vec[j - 1] - 1;
                        double um1 = u[t - 1];
                        u[J] -= F[t - 1] * (E[t - 1] / um1);
                        um1 = u[t - 1];
                        double bm1 = b[t - 1];
                        b[J] -= bm1 * E[t - 1] / um1;
                    }
                }
            }
        }
    }
}

/* BkSub BKSUB1 (CudaStuff.cu:71-103): PX = b/u, PF = -f/u, sentinel row N,
 * then LognDepth rounds of pointer jumping.  In the reference every read in a
 * round sees the value from before the round (ancestors have higher rows and
 * are updated later or in the same lock-step instruction), so it is written
 * here as a Jacobi update.  Result x is left in b[0..N-1]. */
static void bk_sub(const ref_tree *T, double *u, double *b, double *tmp)
{
    int N = T->N;
    double *PX = b, *PF = u;
    for (int j = 0; j < N; j++) {
        PX[j] = PX[j] / PF[j];
        PF[j] = -T->f[j] / PF[j];
    }
    PX[N] = 0; PF[N] = 1;
    double *oPX = tmp, *oPF = tmp + (N + 1);
    for (int r = 0; r < T->logn_depth; r++) {
        memcpy(oPX, PX, sizeof(double) * (N + 1));
        memcpy(oPF, PF, sizeof(double) * (N + 1));
        for (int j = 0; j < N; j++) {
            int a = T->fidxs[r * N + j] - 1;
            PX[j] = oPX[j] + oPX[a] * oPF[j];
        }
        for (int j = 0; j < N; j++) {
            int a = T->fidxs[r * N + j] - 1;
            PF[j] = oPF[j] * oPF[a];
        }
    }
}

/* Plain serial Hines solve used only to cross-check the two routines above:
 * eliminate every row into its parent in ascending row order, then substitute
 * from the root down.  mode 1 = elimination only (back-substitution by
 * pointer jumping), mode 2 = both serial. */
static void serial_solve(const ref_tree *T, double *u, double *b, int mode, double *tmp)
{
    int N = T->N;
    for (int h = 0; h < N - 1; h++) {
        int p = T->ks[h + 1] - 1;
        double um1 = u[h];
        u[p] = u[p] - T->f[h] * (T->e[h] / um1);
        b[p] = b[p] - b[h] * T->e[h] / um1;
    }
    if (mode == 1) { bk_sub(T, u, b, tmp); return; }
    b[N - 1] = b[N - 1] / u[N - 1];
    for (int h = N - 2; h >= 0; h--) {
        int p = T->ks[h + 1] - 1;
        b[h] = b[h] / u[h] + (-T->f[h] / u[h]) * b[p];
    }
}

/* Exposed for tests: solve the Hines system in place; x returned in b[0..N-1].
 * mode 0 = reference (BeforeLU + BKSUB1), 1 = serial elimination + BKSUB1,
 * 2 = fully serial.  u and b must hold N+1 entries. */
int ref_solve(const ref_tree *T, double *u, double *b, int mode)
{
    double *tmp = (double *)malloc(sizeof(double) * 2 * (T->N + 1));
    if (!tmp) return -1;
    if (mode == 0) { before_lu(T, u, b); bk_sub(T, u, b, tmp); }
    else serial_solve(T, u, b, mode, tmp);
    free(tmp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* One run: parameter set `set`, stimulus `stim`. */
typedef struct {
    const ref_tree *T;
    const double *params;   /* [nsets][nparams*ncomps] */
    const double *amps;     /* [nstim][nt_amps]        */
    const double *dts;      /* [nt]                    */
    int nsets, nstim, nt, nt_amps, stim_loc, nrec, solve_mode;
    double stim_area, v_init;
    const int *rec_sites;
    double *vhot;           /* [nsets][nstim][nrec][nt] */
    double *vfinal;         /* optional [nsets*nstim][N] */
} run_args;

static int run_one(const run_args *A, int set, int stim)
{
    const ref_tree *T = A->T;
    int N = T->N, NC = T->ncomps, ms = T->mechset;
    int nm = ref_mechset_nmodels(ms), np = ref_mechset_nparams(ms);
    const int *pbase = ms == MS_TTPC ? TT_PBASE : ms == MS_MAINEN ? MN_PBASE : PS_PBASE;
    if (nm != T->nmodels) return -2;
    double *v = calloc(N, sizeof(double)), *dv = calloc(N, sizeof(double));
    double *st = calloc((size_t)N * MAXST, sizeof(double));
    ionreg *io = calloc(N, sizeof(ionreg));
    double *u = calloc(N + 1, sizeof(double)), *b = calloc(N + 1, sizeof(double));
    double *rhs = calloc(N, sizeof(double)), *D = calloc(N, sizeof(double));
    double *tmp = calloc(2 * (N + 1), sizeof(double));
    double *P = calloc((size_t)N * np, sizeof(double));
    int *has = calloc((size_t)N * nm, sizeof(int)), *par = calloc(N, sizeof(int));
    if (!v || !dv || !st || !io || !u || !b || !rhs || !D || !tmp || !P || !has || !par) return -1;

    /* param_macro(p, s) = ParamsM[set][p*NComps + SegToComp[s]] (CudaStuff.cu:226) */
    const double *PM = A->params + (size_t)set * np * NC;
    for (int s = 0; s < N; s++) {
        for (int p = 0; p < np; p++) P[(size_t)s * np + p] = PM[(size_t)p * NC + T->seg_to_comp[s]];
        for (int m = 0; m < nm; m++) has[s * nm + m] = T->bool_model[m * N + s] != 0;
        par[s] = s == 0 ? 0 : N - T->ks[N - s];      /* CudaStuff.cu:209-215 */
        v[s] = A->v_init;
    }
    (void)pbase;
    for (int s = 0; s < N; s++) {
        if (ms == MS_TTPC) tt_init(has + s * nm, P + (size_t)s * np, v[s], st + (size_t)s * MAXST, io + s);
        else if (ms == MS_MAINEN) mn_init(has + s * nm, P + (size_t)s * np, v[s], st + (size_t)s * MAXST, io + s);
    }
    const double *amp = A->amps + (size_t)stim * A->nt_amps;
    double *out = A->vhot + ((size_t)set * A->nstim + stim) * A->nrec * A->nt;
    for (int i = 0; i < A->nt; i++) {
        double dt = A->dts[i];
        for (int r = 0; r < A->nrec; r++) out[(size_t)r * A->nt + i] = v[A->rec_sites[r]];
        for (int s = 0; s < N; s++) {
            double *Ps = P + (size_t)s * np, *sts = st + (size_t)s * MAXST;
            const int *hs = has + s * nm;
            double sumDv = 0, sum = 0, istim = 0;
            io[s].ica = 0;
            if (s == A->stim_loc) istim = 100 * amp[i] / A->stim_area;
            if (ms == MS_TTPC) {
                ionreg tmpio = io[s];
                tt_cur(hs, Ps, v[s] + 0.001, sts, &tmpio, 0, &sumDv);
                tt_cur(hs, Ps, v[s], sts, io + s, 1, &sum);
            } else if (ms == MS_MAINEN) {
                ionreg tmpio = io[s];
                mn_cur(hs, Ps, v[s] + 0.001, sts, &tmpio, 0, &sumDv);
                mn_cur(hs, Ps, v[s], sts, io + s, 1, &sum);
            } else {
                ps_cur(hs, Ps, v[s] + 0.001, &sumDv);
                ps_cur(hs, Ps, v[s], &sum);
            }
            int h = N - s - 1;
            double gModel = (sumDv - sum) / EPS_V;
            rhs[s] = istim - sum;
            D[s] = gModel + T->cm[s] / (dt * 1000);
            D[s] -= T->f[h];
            dv[s] += b[N - par[s] - 1] - b[h];
        }
        for (int s = 0; s < N; s++) {
            int h = N - s - 1;
            rhs[s] -= T->f[h] * dv[s];
            b[h] = rhs[s];
            u[h] = D[s];
        }
        for (int s = 0; s < N; s++) {          /* CudaStuff.cu:296-305, sequential in s */
            int sn = T->son_no[s];
            if (sn == 1 || sn == 2) {
                int hp = N - par[s] - 1, h = N - s - 1;
                b[hp] += T->e[h] * dv[s];
                u[hp] -= T->e[h];
            }
        }
        if (A->solve_mode == 0) { before_lu(T, u, b); bk_sub(T, u, b, tmp); }
        else serial_solve(T, u, b, A->solve_mode, tmp);
        for (int s = 0; s < N; s++) v[s] += b[N - s - 1];
        for (int s = 0; s < N; s++) {
            if (ms == MS_TTPC) tt_deriv(has + s * nm, P + (size_t)s * np, dt, v[s], st + (size_t)s * MAXST, io + s);
            else if (ms == MS_MAINEN) mn_deriv(has + s * nm, P + (size_t)s * np, dt, v[s], st + (size_t)s * MAXST, io + s);
        }
    }
    if (A->vfinal) memcpy(A->vfinal + ((size_t)set * A->nstim + stim) * N, v, sizeof(double) * N);
    free(v); free(dv); free(st); free(io); free(u); free(b); free(rhs); free(D);
    free(tmp); free(P); free(has); free(par);
    return 0;
}

/* Simulate every (set, stim) run.  Returns 0 on success, <0 on error. */
int ref_simulate(const ref_tree *T, const double *params, int nsets,
                 const double *amps, int nstim, int nt_amps,
                 const double *dts, int nt, int stim_loc, double stim_area,
                 double v_init, int nrec, const int *rec_sites,
                 double *vhot, double *vfinal, int solve_mode, int nthreads)
{
    if (!T || T->N < 1 || nsets < 0 || nstim < 1 || nt < 0 || nt_amps < nt) return -3;
    if (stim_loc < -1 || stim_loc >= T->N) return -4;
    for (int r = 0; r < nrec; r++) if (rec_sites[r] < 0 || rec_sites[r] >= T->N) return -5;
    run_args A = {T, params, amps, dts, nsets, nstim, nt, nt_amps, stim_loc, nrec, solve_mode,
                  stim_area, v_init, rec_sites, vhot, vfinal};
    int total = nsets * nstim, err = 0;
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
    #pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads) reduction(min:err)
#endif
    for (int k = 0; k < total; k++) {
        int rc = run_one(&A, k / nstim, k % nstim);
        if (rc < err) err = rc;
    }
    (void)nthreads;
    return err;
}

/* Rate functions exposed for closed-form tests.  mech follows the set's order;
 * out receives (inf0, tau0, inf1, tau1). */
int ref_rates(int ms, int mech, double x, const double *P, double *out)
{
    out[0] = out[1] = out[2] = out[3] = 0;
    if (ms == MS_TTPC) {
        switch (mech) {
        case 0: r_ca_hva(x, out, out + 1, out + 2, out + 3); return 0;
        case 1: r_ca_lvast(x, out, out + 1, out + 2, out + 3); return 0;
        case 3: r_ih(x, out, out + 1); return 0;
        case 4: r_im(x, out, out + 1); return 0;
        case 5: r_k_pst(x, out, out + 1, out + 2, out + 3); return 0;
        case 6: r_k_tst(x, out, out + 1, out + 2, out + 3); return 0;
        case 7: r_nap_et2(x, out, out + 1, out + 2, out + 3); return 0;
        case 8: r_nata_t(x, out, out + 1, out + 2, out + 3); return 0;
        case 9: r_nats2_t(x, out, out + 1, out + 2, out + 3); return 0;
        case 11: out[0] = r_sk_e2(x); return 0;
        case 12: r_skv3_1(x, out, out + 1); return 0;
        }
    } else if (ms == MS_MAINEN) {
        switch (mech) {
        case 0: r_mn_ca(x, out, out + 1, out + 2, out + 3); return 0;
        case 2: r_mn_kca(x, P[4], P[5], P[3], out, out + 1); return 0;
        case 3: r_mn_kx(x, P[7], P[8], P[9], P[10], out, out + 1); return 0;
        case 4: r_mn_kx(x, P[12], P[13], P[14], P[15], out, out + 1); return 0;
        case 5: r_mn_na(x, P, out, out + 1, out + 2, out + 3); return 0;
        }
    }
    return -1;
}

/* The membrane current of one mechanism alone (the BREAKPOINT terms of
 * tt_cur / mn_cur with only that mechanism present) at v, for states st
 * (state_macro slots) and compartment parameters P, with eca given:
 * out[0] = the current, out[1] = its calcium part.  Returns -1 for a
 * mechanism without a current term. */
int ref_current(int ms, int mech, double v, const double *st, const double *P, double eca, double *out)
{
    int has[16] = {0};
    if (mech < 0 || mech >= 16) return -1;
    has[mech] = 1;
    ionreg io = {5e-5, eca, 0.0};
    double s = 0.0;
    if (ms == MS_TTPC && mech != 2) tt_cur(has, P, v, st, &io, 1, &s);
    else if (ms == MS_MAINEN && mech != 1) mn_cur(has, P, v, st, &io, 1, &s);
    else return -1;
    out[0] = s;
    out[1] = io.ica;
    return 0;
}

/* One mechanism's state update alone (its DERIVATIVE terms of tt_deriv /
 * mn_deriv) over dt at v: st updated in place, *cai in and out, ica the
 * calcium current of the step.  Returns -1 for a mechanism without states. */
int ref_deriv(int ms, int mech, double dt, double v, double *st, const double *P, double *cai, double ica)
{
    int has[16] = {0};
    if (mech < 0 || mech >= 16) return -1;
    has[mech] = 1;
    if (ms == MS_TTPC && mech != 10) {
        ionreg io = {*cai, tt_eca(*cai), ica};
        tt_deriv(has, P, dt, v, st, &io);
        *cai = io.cai;
    } else if (ms == MS_MAINEN && mech != 6) {
        ionreg io = {*cai, mn_eca(*cai), ica};
        mn_deriv(has, P, dt, v, st, &io);
        *cai = io.cai;
    } else {
        return -1;
    }
    return 0;
}

double ref_cnexp(double y, double yinf, double tau, double dt) { return cnexp1(y, yinf, tau, dt); }
double ref_ktf(int ms) { return ms == MS_MAINEN ? MN_KTF : TT_KTF; }

/* The shared transcendental functions, exposed for the accuracy tests. */
double ref_math(int which, double x, double y)
{
    return which == 0 ? ng_exp(x) : which == 1 ? ng_log(x) : ng_pow(x, y);
}
