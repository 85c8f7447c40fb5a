/* Generated by neurogpu_amd.nmodl -- do not edit. */
#pragma once
#include "mechs.h"
#include <math.h>
#ifdef __HIPCC__
#define NG_HD_M __host__ __device__ inline
#else
#define NG_HD_M inline
#endif


namespace ng_gen_mainen {

/* ca: generated by neurogpu_amd.nmodl from ca.mod */
struct Mech_ca {
    double v = 0.0, dt = 0.0;
    double gbar = 0.1;
    double vshift = 0.0;
    double cao = 2.5;
    double cai = 0.0;
    double temp = 23.0;
    double q10 = 2.3;
    double vmin = -120.0;
    double vmax = 100.0;
    static constexpr double FARADAY = 96485.309;
    static constexpr double R = 8.313424;
    double m = 0.0;
    double h = 0.0;
    double celsius = 0.0;
    double ica = 0.0;
    double gca = 0.0;
    double eca = 0.0;
    double minf = 0.0;
    double hinf = 0.0;
    double mtau = 0.0;
    double htau = 0.0;
    double tadj = 3.2093639532679714;
    NG_HD_M double efun(double z)
    {
        double ret_ = 0.0;
        if ((fabs(z) < 0.0001)) {
            ret_ = (1.0 - (z * 0.5));
        } else {
            const double exp_0 = ng_expc(z);
            ret_ = ng_div(z, (exp_0 - 1.0));
        }
        return ret_;
    }
    NG_HD_M void trates(double v)
    {
        double a_l0 = 0.0;
        double b_l0 = 0.0;
        tadj = 3.2093639532679714;
        double efun_v1 = 0.0;
        if ((fabs(((-(27.0 + v)) * 0.2631578947368421)) < 0.0001)) {
            efun_v1 = (1.0 - (((-(27.0 + v)) * 0.2631578947368421) * 0.5));
        } else {
            const double exp_0 = ng_expc(((-(27.0 + v)) * 0.2631578947368421));
            efun_v1 = ng_div(((-(27.0 + v)) * 0.2631578947368421), (exp_0 - 1.0));
        }
        a_l0 = (0.209 * efun_v1);
        const double exp_1 = ng_expc(((-75.0 - v) * 0.058823529411764705));
        b_l0 = (0.94 * exp_1);
        mtau = ng_div(0.3115882195229801, (a_l0 + b_l0));
        minf = ng_div(a_l0, (a_l0 + b_l0));
        const double exp_2 = ng_expc(((-13.0 - v) * 0.02));
        a_l0 = (0.000457 * exp_2);
        const double exp_3 = ng_expc((((-v) - 15.0) * 0.03571428571428571));
        b_l0 = ng_div(0.0065, (exp_3 + 1.0));
        htau = ng_div(0.3115882195229801, (a_l0 + b_l0));
        hinf = ng_div(a_l0, (a_l0 + b_l0));
    }
    NG_HD_M void rates(double vm)
    {
        double a = 0.0;
        double b = 0.0;
        tadj = 3.2093639532679714;
        double efun_v0 = 0.0;
        if ((fabs(((-(27.0 + vm)) * 0.2631578947368421)) < 0.0001)) {
            efun_v0 = (1.0 - (((-(27.0 + vm)) * 0.2631578947368421) * 0.5));
        } else {
            const double exp_0 = ng_expc(((-(27.0 + vm)) * 0.2631578947368421));
            efun_v0 = ng_div(((-(27.0 + vm)) * 0.2631578947368421), (exp_0 - 1.0));
        }
        a = (0.209 * efun_v0);
        const double exp_1 = ng_expc(((-75.0 - vm) * 0.058823529411764705));
        b = (0.94 * exp_1);
        mtau = ng_div(0.3115882195229801, (a + b));
        minf = ng_div(a, (a + b));
        const double exp_2 = ng_expc(((-13.0 - vm) * 0.02));
        a = (0.000457 * exp_2);
        const double exp_3 = ng_expc((((-vm) - 15.0) * 0.03571428571428571));
        b = ng_div(0.0065, (exp_3 + 1.0));
        htau = ng_div(0.3115882195229801, (a + b));
        hinf = ng_div(a, (a + b));
    }
    NG_HD_M void initial()
    {
        tadj = 3.2093639532679714;
        double a_l1 = 0.0;
        double b_l1 = 0.0;
        tadj = 3.2093639532679714;
        double efun_v2 = 0.0;
        if ((fabs(((-(27.0 + (v + 0.0))) * 0.2631578947368421)) < 0.0001)) {
            efun_v2 = (1.0 - (((-(27.0 + (v + 0.0))) * 0.2631578947368421) * 0.5));
        } else {
            const double exp_0 = ng_expc(((-(27.0 + (v + 0.0))) * 0.2631578947368421));
            efun_v2 = ng_div(((-(27.0 + (v + 0.0))) * 0.2631578947368421), (exp_0 - 1.0));
        }
        a_l1 = (0.209 * efun_v2);
        const double exp_1 = ng_expc(((-75.0 - (v + 0.0)) * 0.058823529411764705));
        b_l1 = (0.94 * exp_1);
        mtau = ng_div(0.3115882195229801, (a_l1 + b_l1));
        minf = ng_div(a_l1, (a_l1 + b_l1));
        const double exp_2 = ng_expc(((-13.0 - (v + 0.0)) * 0.02));
        a_l1 = (0.000457 * exp_2);
        const double exp_3 = ng_expc((((-(v + 0.0)) - 15.0) * 0.03571428571428571));
        b_l1 = ng_div(0.0065, (exp_3 + 1.0));
        htau = ng_div(0.3115882195229801, (a_l1 + b_l1));
        hinf = ng_div(a_l1, (a_l1 + b_l1));
        m = minf;
        h = hinf;
    }
    NG_HD_M void states()
    {
        double a_l1 = 0.0;
        double b_l1 = 0.0;
        tadj = 3.2093639532679714;
        double efun_v2 = 0.0;
        if ((fabs(((-(27.0 + (v + 0.0))) * 0.2631578947368421)) < 0.0001)) {
            efun_v2 = (1.0 - (((-(27.0 + (v + 0.0))) * 0.2631578947368421) * 0.5));
        } else {
            const double exp_0 = ng_expc(((-(27.0 + (v + 0.0))) * 0.2631578947368421));
            efun_v2 = ng_div(((-(27.0 + (v + 0.0))) * 0.2631578947368421), (exp_0 - 1.0));
        }
        a_l1 = (0.209 * efun_v2);
        const double exp_1 = ng_expc(((-75.0 - (v + 0.0)) * 0.058823529411764705));
        b_l1 = (0.94 * exp_1);
        mtau = ng_div(0.3115882195229801, (a_l1 + b_l1));
        minf = ng_div(a_l1, (a_l1 + b_l1));
        const double exp_2 = ng_expc(((-13.0 - (v + 0.0)) * 0.02));
        a_l1 = (0.000457 * exp_2);
        const double exp_3 = ng_expc((((-(v + 0.0)) - 15.0) * 0.03571428571428571));
        b_l1 = ng_div(0.0065, (exp_3 + 1.0));
        htau = ng_div(0.3115882195229801, (a_l1 + b_l1));
        hinf = ng_div(a_l1, (a_l1 + b_l1));
        m = m + (1.0 - ng_expc(-(dt * ng_rcp(mtau)))) * (minf - m);
        h = h + (1.0 - ng_expc(-(dt * ((a_l1 + b_l1) * 3.2093639532679714)))) * (hinf - h);
    }
    NG_HD_M void breakpoint()
    {
        gca = ((((3.2093639532679714 * gbar) * m) * m) * h);
        ica = ((0.0001 * gca) * (v - eca));
    }
};

/* cad: generated by neurogpu_amd.nmodl from cad.mod */
struct Mech_cad {
    double v = 0.0, dt = 0.0;
    double depth = 0.1;
    double taur = 200.0;
    double cainf = 0.0001;
    double cai = 0.0;
    static constexpr double FARADAY = 96485.309;
    double ca = 0.0;
    double ica = 0.0;
    double drive_channel = 0.0;
    NG_HD_M void initial()
    {
        ca = 0.0001;
        cai = ca;
    }
    NG_HD_M void states()
    {
        drive_channel = ((-10000.0 * ica) * 5.551557387549293e-05);
        if ((drive_channel <= 0.0)) {
            drive_channel = 0.0;
        }
        ca = ca + (1.0 - ng_expc(-(dt * 0.005))) * (((drive_channel + 5e-07) * 200.0) - ca);
        cai = ca;
    }
    NG_HD_M void breakpoint()
    {
    }
};

/* kca: generated by neurogpu_amd.nmodl from kca.mod */
struct Mech_kca {
    double v = 0.0, dt = 0.0;
    double gbar = 10.0;
    double cai = 0.0;
    double caix = 1.0;
    double Ra = 0.01;
    double Rb = 0.02;
    double celsius = 0.0;
    double temp = 23.0;
    double q10 = 2.3;
    double vmin = -120.0;
    double vmax = 100.0;
    double n = 0.0;
    double a = 0.0;
    double b = 0.0;
    double ik = 0.0;
    double gk = 0.0;
    double ek = 0.0;
    double ninf = 0.0;
    double ntau = 0.0;
    double tadj = 3.2093639532679714;
    NG_HD_M void rates(double cai)
    {
        a = (Ra * ng_pow(cai, caix));
        b = Rb;
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
    }
    NG_HD_M void initial()
    {
        a = (Ra * ng_pow(cai, caix));
        b = Rb;
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
        n = ninf;
    }
    NG_HD_M void states()
    {
        a = (Ra * ng_pow(cai, caix));
        b = Rb;
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
        n = n + (1.0 - ng_expc(-(dt * ((a + b) * 3.2093639532679714)))) * (ninf - n);
    }
    NG_HD_M void breakpoint()
    {
        gk = ((3.2093639532679714 * gbar) * n);
        ik = ((0.0001 * gk) * (v - ek));
    }
};

/* km: generated by neurogpu_amd.nmodl from km.mod */
struct Mech_km {
    double v = 0.0, dt = 0.0;
    double gbar = 10.0;
    double tha = -30.0;
    double qa = 9.0;
    double Ra = 0.001;
    double Rb = 0.001;
    double temp = 23.0;
    double q10 = 2.3;
    double vmin = -120.0;
    double vmax = 100.0;
    double n = 0.0;
    double celsius = 0.0;
    double a = 0.0;
    double b = 0.0;
    double ik = 0.0;
    double gk = 0.0;
    double ek = 0.0;
    double ninf = 0.0;
    double ntau = 0.0;
    double tadj = 3.2093639532679714;
    NG_HD_M double efun(double z)
    {
        double ret_ = 0.0;
        if ((fabs(z) < 0.0001)) {
            ret_ = (1.0 - (z * 0.5));
        } else {
            const double exp_0 = ng_expc(z);
            ret_ = ng_div(z, (exp_0 - 1.0));
        }
        return ret_;
    }
    NG_HD_M void trates(double v)
    {
        double efun_v1 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((-(v - tha)), qa)) < 0.0001)) {
            efun_v1 = (1.0 - (ng_div((-(v - tha)), qa) * 0.5));
        } else {
            efun_v1 = ng_div(ng_div((-(v - tha)), qa), (exp_0 - 1.0));
        }
        a = ((Ra * qa) * efun_v1);
        double efun_v2 = 0.0;
        if ((fabs(ng_div((v - tha), qa)) < 0.0001)) {
            efun_v2 = (1.0 - (ng_div((v - tha), qa) * 0.5));
        } else {
            efun_v2 = ng_div(ng_div((v - tha), qa), (ng_rcp(exp_0) - 1.0));
        }
        b = ((Rb * qa) * efun_v2);
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
    }
    NG_HD_M void rates(double v)
    {
        double efun_v0 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((-(v - tha)), qa)) < 0.0001)) {
            efun_v0 = (1.0 - (ng_div((-(v - tha)), qa) * 0.5));
        } else {
            efun_v0 = ng_div(ng_div((-(v - tha)), qa), (exp_0 - 1.0));
        }
        a = ((Ra * qa) * efun_v0);
        double efun_v1 = 0.0;
        if ((fabs(ng_div((v - tha), qa)) < 0.0001)) {
            efun_v1 = (1.0 - (ng_div((v - tha), qa) * 0.5));
        } else {
            efun_v1 = ng_div(ng_div((v - tha), qa), (ng_rcp(exp_0) - 1.0));
        }
        b = ((Rb * qa) * efun_v1);
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
    }
    NG_HD_M void initial()
    {
        tadj = 3.2093639532679714;
        double efun_v2 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((-(v - tha)), qa)) < 0.0001)) {
            efun_v2 = (1.0 - (ng_div((-(v - tha)), qa) * 0.5));
        } else {
            efun_v2 = ng_div(ng_div((-(v - tha)), qa), (exp_0 - 1.0));
        }
        a = ((Ra * qa) * efun_v2);
        double efun_v3 = 0.0;
        if ((fabs(ng_div((v - tha), qa)) < 0.0001)) {
            efun_v3 = (1.0 - (ng_div((v - tha), qa) * 0.5));
        } else {
            efun_v3 = ng_div(ng_div((v - tha), qa), (ng_rcp(exp_0) - 1.0));
        }
        b = ((Rb * qa) * efun_v3);
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
        n = ninf;
    }
    NG_HD_M void states()
    {
        double efun_v2 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((-(v - tha)), qa)) < 0.0001)) {
            efun_v2 = (1.0 - (ng_div((-(v - tha)), qa) * 0.5));
        } else {
            efun_v2 = ng_div(ng_div((-(v - tha)), qa), (exp_0 - 1.0));
        }
        a = ((Ra * qa) * efun_v2);
        double efun_v3 = 0.0;
        if ((fabs(ng_div((v - tha), qa)) < 0.0001)) {
            efun_v3 = (1.0 - (ng_div((v - tha), qa) * 0.5));
        } else {
            efun_v3 = ng_div(ng_div((v - tha), qa), (ng_rcp(exp_0) - 1.0));
        }
        b = ((Rb * qa) * efun_v3);
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
        n = n + (1.0 - ng_expc(-(dt * ((a + b) * 3.2093639532679714)))) * (ninf - n);
    }
    NG_HD_M void breakpoint()
    {
        gk = ((3.2093639532679714 * gbar) * n);
        ik = ((0.0001 * gk) * (v - ek));
    }
};

/* kv: generated by neurogpu_amd.nmodl from kv.mod */
struct Mech_kv {
    double v = 0.0, dt = 0.0;
    double gbar = 5.0;
    double tha = 25.0;
    double qa = 9.0;
    double Ra = 0.02;
    double Rb = 0.002;
    double temp = 23.0;
    double q10 = 2.3;
    double vmin = -120.0;
    double vmax = 100.0;
    double n = 0.0;
    double celsius = 0.0;
    double a = 0.0;
    double b = 0.0;
    double ik = 0.0;
    double gk = 0.0;
    double ek = 0.0;
    double ninf = 0.0;
    double ntau = 0.0;
    double tadj = 3.2093639532679714;
    NG_HD_M double efun(double z)
    {
        double ret_ = 0.0;
        if ((fabs(z) < 0.0001)) {
            ret_ = (1.0 - (z * 0.5));
        } else {
            const double exp_0 = ng_expc(z);
            ret_ = ng_div(z, (exp_0 - 1.0));
        }
        return ret_;
    }
    NG_HD_M void trates(double v)
    {
        double efun_v1 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((-(v - tha)), qa)) < 0.0001)) {
            efun_v1 = (1.0 - (ng_div((-(v - tha)), qa) * 0.5));
        } else {
            efun_v1 = ng_div(ng_div((-(v - tha)), qa), (exp_0 - 1.0));
        }
        a = ((Ra * qa) * efun_v1);
        double efun_v2 = 0.0;
        if ((fabs(ng_div((v - tha), qa)) < 0.0001)) {
            efun_v2 = (1.0 - (ng_div((v - tha), qa) * 0.5));
        } else {
            efun_v2 = ng_div(ng_div((v - tha), qa), (ng_rcp(exp_0) - 1.0));
        }
        b = ((Rb * qa) * efun_v2);
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
    }
    NG_HD_M void rates(double v)
    {
        double efun_v0 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((-(v - tha)), qa)) < 0.0001)) {
            efun_v0 = (1.0 - (ng_div((-(v - tha)), qa) * 0.5));
        } else {
            efun_v0 = ng_div(ng_div((-(v - tha)), qa), (exp_0 - 1.0));
        }
        a = ((Ra * qa) * efun_v0);
        double efun_v1 = 0.0;
        if ((fabs(ng_div((v - tha), qa)) < 0.0001)) {
            efun_v1 = (1.0 - (ng_div((v - tha), qa) * 0.5));
        } else {
            efun_v1 = ng_div(ng_div((v - tha), qa), (ng_rcp(exp_0) - 1.0));
        }
        b = ((Rb * qa) * efun_v1);
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
    }
    NG_HD_M void initial()
    {
        tadj = 3.2093639532679714;
        double efun_v2 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((-(v - tha)), qa)) < 0.0001)) {
            efun_v2 = (1.0 - (ng_div((-(v - tha)), qa) * 0.5));
        } else {
            efun_v2 = ng_div(ng_div((-(v - tha)), qa), (exp_0 - 1.0));
        }
        a = ((Ra * qa) * efun_v2);
        double efun_v3 = 0.0;
        if ((fabs(ng_div((v - tha), qa)) < 0.0001)) {
            efun_v3 = (1.0 - (ng_div((v - tha), qa) * 0.5));
        } else {
            efun_v3 = ng_div(ng_div((v - tha), qa), (ng_rcp(exp_0) - 1.0));
        }
        b = ((Rb * qa) * efun_v3);
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
        n = ninf;
    }
    NG_HD_M void states()
    {
        double efun_v2 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((-(v - tha)), qa)) < 0.0001)) {
            efun_v2 = (1.0 - (ng_div((-(v - tha)), qa) * 0.5));
        } else {
            efun_v2 = ng_div(ng_div((-(v - tha)), qa), (exp_0 - 1.0));
        }
        a = ((Ra * qa) * efun_v2);
        double efun_v3 = 0.0;
        if ((fabs(ng_div((v - tha), qa)) < 0.0001)) {
            efun_v3 = (1.0 - (ng_div((v - tha), qa) * 0.5));
        } else {
            efun_v3 = ng_div(ng_div((v - tha), qa), (ng_rcp(exp_0) - 1.0));
        }
        b = ((Rb * qa) * efun_v3);
        tadj = 3.2093639532679714;
        ntau = ng_div(0.3115882195229801, (a + b));
        ninf = ng_div(a, (a + b));
        n = n + (1.0 - ng_expc(-(dt * ((a + b) * 3.2093639532679714)))) * (ninf - n);
    }
    NG_HD_M void breakpoint()
    {
        gk = ((3.2093639532679714 * gbar) * n);
        ik = ((0.0001 * gk) * (v - ek));
    }
};

/* na: generated by neurogpu_amd.nmodl from na.mod */
struct Mech_na {
    double v = 0.0, dt = 0.0;
    double gbar = 1000.0;
    double vshift = -10.0;
    double tha = -35.0;
    double qa = 9.0;
    double Ra = 0.182;
    double Rb = 0.124;
    double thi1 = -50.0;
    double thi2 = -75.0;
    double qi = 5.0;
    double thinf = -65.0;
    double qinf = 6.2;
    double Rg = 0.0091;
    double Rd = 0.024;
    double temp = 23.0;
    double q10 = 2.3;
    double celsius = 0.0;
    double vmin = -120.0;
    double vmax = 100.0;
    double m = 0.0;
    double h = 0.0;
    double ina = 0.0;
    double gna = 0.0;
    double ena = 0.0;
    double minf = 0.0;
    double hinf = 0.0;
    double mtau = 0.0;
    double htau = 0.0;
    double tadj = 3.2093639532679714;
    NG_HD_M double trap0(double v, double th, double a, double q)
    {
        double ret_ = 0.0;
        if ((fabs(ng_div((v - th), q)) > 1e-06)) {
            const double exp_0 = ng_expc(ng_div((-(v - th)), q));
            ret_ = ng_div((a * (v - th)), (1.0 - exp_0));
        } else {
            ret_ = (a * q);
        }
        return ret_;
    }
    NG_HD_M void trates(double v)
    {
        double a_l0 = 0.0;
        double b_l0 = 0.0;
        double trap0_v1 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(v - tha)), qa));
        if ((fabs(ng_div((v - tha), qa)) > 1e-06)) {
            trap0_v1 = ng_div((Ra * (v - tha)), (1.0 - exp_0));
        } else {
            trap0_v1 = (Ra * qa);
        }
        a_l0 = trap0_v1;
        double trap0_v2 = 0.0;
        if ((fabs(ng_div(((-v) - (-tha)), qa)) > 1e-06)) {
            trap0_v2 = ng_div((Rb * ((-v) - (-tha))), (1.0 - ng_rcp(exp_0)));
        } else {
            trap0_v2 = (Rb * qa);
        }
        b_l0 = trap0_v2;
        tadj = 3.2093639532679714;
        mtau = ng_div(0.3115882195229801, (a_l0 + b_l0));
        minf = ng_div(a_l0, (a_l0 + b_l0));
        double trap0_v3 = 0.0;
        if ((fabs(ng_div((v - thi1), qi)) > 1e-06)) {
            const double exp_1 = ng_expc(ng_div((-(v - thi1)), qi));
            trap0_v3 = ng_div((Rd * (v - thi1)), (1.0 - exp_1));
        } else {
            trap0_v3 = (Rd * qi);
        }
        a_l0 = trap0_v3;
        double trap0_v4 = 0.0;
        if ((fabs(ng_div(((-v) - (-thi2)), qi)) > 1e-06)) {
            const double exp_2 = ng_expc(ng_div((-((-v) - (-thi2))), qi));
            trap0_v4 = ng_div((Rg * ((-v) - (-thi2))), (1.0 - exp_2));
        } else {
            trap0_v4 = (Rg * qi);
        }
        b_l0 = trap0_v4;
        htau = ng_div(0.3115882195229801, (a_l0 + b_l0));
        const double exp_3 = ng_expc(ng_div((v - thinf), qinf));
        hinf = ng_rcp((1.0 + exp_3));
    }
    NG_HD_M void rates(double vm)
    {
        double a = 0.0;
        double b = 0.0;
        double trap0_v0 = 0.0;
        const double exp_0 = ng_expc(ng_div((-(vm - tha)), qa));
        if ((fabs(ng_div((vm - tha), qa)) > 1e-06)) {
            trap0_v0 = ng_div((Ra * (vm - tha)), (1.0 - exp_0));
        } else {
            trap0_v0 = (Ra * qa);
        }
        a = trap0_v0;
        double trap0_v1 = 0.0;
        if ((fabs(ng_div(((-vm) - (-tha)), qa)) > 1e-06)) {
            trap0_v1 = ng_div((Rb * ((-vm) - (-tha))), (1.0 - ng_rcp(exp_0)));
        } else {
            trap0_v1 = (Rb * qa);
        }
        b = trap0_v1;
        tadj = 3.2093639532679714;
        mtau = ng_div(0.3115882195229801, (a + b));
        minf = ng_div(a, (a + b));
        double trap0_v2 = 0.0;
        if ((fabs(ng_div((vm - thi1), qi)) > 1e-06)) {
            const double exp_1 = ng_expc(ng_div((-(vm - thi1)), qi));
            trap0_v2 = ng_div((Rd * (vm - thi1)), (1.0 - exp_1));
        } else {
            trap0_v2 = (Rd * qi);
        }
        a = trap0_v2;
        double trap0_v3 = 0.0;
        if ((fabs(ng_div(((-vm) - (-thi2)), qi)) > 1e-06)) {
            const double exp_2 = ng_expc(ng_div((-((-vm) - (-thi2))), qi));
            trap0_v3 = ng_div((Rg * ((-vm) - (-thi2))), (1.0 - exp_2));
        } else {
            trap0_v3 = (Rg * qi);
        }
        b = trap0_v3;
        htau = ng_div(0.3115882195229801, (a + b));
        const double exp_3 = ng_expc(ng_div((vm - thinf), qinf));
        hinf = ng_rcp((1.0 + exp_3));
    }
    NG_HD_M void initial()
    {
        tadj = 3.2093639532679714;
        double a_l1 = 0.0;
        double b_l1 = 0.0;
        double trap0_v2 = 0.0;
        const double exp_0 = ng_expc(ng_div((-((v + -5.0) - tha)), qa));
        if ((fabs(ng_div(((v + -5.0) - tha), qa)) > 1e-06)) {
            trap0_v2 = ng_div((Ra * ((v + -5.0) - tha)), (1.0 - exp_0));
        } else {
            trap0_v2 = (Ra * qa);
        }
        a_l1 = trap0_v2;
        double trap0_v3 = 0.0;
        if ((fabs(ng_div(((-(v + -5.0)) - (-tha)), qa)) > 1e-06)) {
            trap0_v3 = ng_div((Rb * ((-(v + -5.0)) - (-tha))), (1.0 - ng_rcp(exp_0)));
        } else {
            trap0_v3 = (Rb * qa);
        }
        b_l1 = trap0_v3;
        tadj = 3.2093639532679714;
        mtau = ng_div(0.3115882195229801, (a_l1 + b_l1));
        minf = ng_div(a_l1, (a_l1 + b_l1));
        double trap0_v4 = 0.0;
        if ((fabs(ng_div(((v + -5.0) - thi1), qi)) > 1e-06)) {
            const double exp_1 = ng_expc(ng_div((-((v + -5.0) - thi1)), qi));
            trap0_v4 = ng_div((Rd * ((v + -5.0) - thi1)), (1.0 - exp_1));
        } else {
            trap0_v4 = (Rd * qi);
        }
        a_l1 = trap0_v4;
        double trap0_v5 = 0.0;
        if ((fabs(ng_div(((-(v + -5.0)) - (-thi2)), qi)) > 1e-06)) {
            const double exp_2 = ng_expc(ng_div((-((-(v + -5.0)) - (-thi2))), qi));
            trap0_v5 = ng_div((Rg * ((-(v + -5.0)) - (-thi2))), (1.0 - exp_2));
        } else {
            trap0_v5 = (Rg * qi);
        }
        b_l1 = trap0_v5;
        htau = ng_div(0.3115882195229801, (a_l1 + b_l1));
        const double exp_3 = ng_expc(ng_div(((v + -5.0) - thinf), qinf));
        hinf = ng_rcp((1.0 + exp_3));
        m = minf;
        h = hinf;
    }
    NG_HD_M void states()
    {
        double a_l1 = 0.0;
        double b_l1 = 0.0;
        double trap0_v2 = 0.0;
        const double exp_0 = ng_expc(ng_div((-((v + -5.0) - tha)), qa));
        if ((fabs(ng_div(((v + -5.0) - tha), qa)) > 1e-06)) {
            trap0_v2 = ng_div((Ra * ((v + -5.0) - tha)), (1.0 - exp_0));
        } else {
            trap0_v2 = (Ra * qa);
        }
        a_l1 = trap0_v2;
        double trap0_v3 = 0.0;
        if ((fabs(ng_div(((-(v + -5.0)) - (-tha)), qa)) > 1e-06)) {
            trap0_v3 = ng_div((Rb * ((-(v + -5.0)) - (-tha))), (1.0 - ng_rcp(exp_0)));
        } else {
            trap0_v3 = (Rb * qa);
        }
        b_l1 = trap0_v3;
        tadj = 3.2093639532679714;
        mtau = ng_div(0.3115882195229801, (a_l1 + b_l1));
        minf = ng_div(a_l1, (a_l1 + b_l1));
        double trap0_v4 = 0.0;
        if ((fabs(ng_div(((v + -5.0) - thi1), qi)) > 1e-06)) {
            const double exp_1 = ng_expc(ng_div((-((v + -5.0) - thi1)), qi));
            trap0_v4 = ng_div((Rd * ((v + -5.0) - thi1)), (1.0 - exp_1));
        } else {
            trap0_v4 = (Rd * qi);
        }
        a_l1 = trap0_v4;
        double trap0_v5 = 0.0;
        if ((fabs(ng_div(((-(v + -5.0)) - (-thi2)), qi)) > 1e-06)) {
            const double exp_2 = ng_expc(ng_div((-((-(v + -5.0)) - (-thi2))), qi));
            trap0_v5 = ng_div((Rg * ((-(v + -5.0)) - (-thi2))), (1.0 - exp_2));
        } else {
            trap0_v5 = (Rg * qi);
        }
        b_l1 = trap0_v5;
        htau = ng_div(0.3115882195229801, (a_l1 + b_l1));
        const double exp_3 = ng_expc(ng_div(((v + -5.0) - thinf), qinf));
        hinf = ng_rcp((1.0 + exp_3));
        m = m + (1.0 - ng_expc(-(dt * ng_rcp(mtau)))) * (minf - m);
        h = h + (1.0 - ng_expc(-(dt * ((a_l1 + b_l1) * 3.2093639532679714)))) * (hinf - h);
    }
    NG_HD_M void breakpoint()
    {
        gna = (((((3.2093639532679714 * gbar) * m) * m) * m) * h);
        ina = ((0.0001 * gna) * (v - ena));
    }
};

/* pas: generated by neurogpu_amd.nmodl from pas (built in) */
struct Mech_pas {
    double v = 0.0, dt = 0.0;
    double g = 0.001;
    double e = -70.0;
    double i = 0.0;
    NG_HD_M void initial()
    {
    }
    NG_HD_M void states()
    {
    }
    NG_HD_M void breakpoint()
    {
        i = (g * (v - e));
    }
};

/* SetGen_mainen: the simulation kernel's mechanism set over the generated mechanisms
 * ca, cad, kca, km, kv, na, pas (bit k = mechanism k); generated by neurogpu_amd.nmodl. */
struct SetGen_mainen {
    static constexpr int NMECH = 7, NPARAM = 30, NSTATE = 11;
    static constexpr int CAI = 8, ECA = 9, ICA = 10;
    /* first parameter-matrix column of each mechanism, then the total */
    static constexpr int PBASE[8] = {0, 2, 2, 6, 11, 16, 28, 30};
    template <class F> NG_DEV static void dispatch(unsigned m, F &&f)
    {
        ng_dispatch(m, f, NgMasks<0x6fu>{});
    }
    NG_DEV static unsigned ica_bits(unsigned m) { return (m & 0x2u) ? 1u << ICA : 0u; }
    NG_DEV static double eca_of(double cai) { return 13.361624877005887 * ng_logc(ng_div(2.0, cai)); }
    NG_DEV static void eca_after_cnexp(unsigned m, double *st)
    {
        if (m & 0x7u) st[ECA] = eca_of(st[CAI]);
    }
    NG_DEV static unsigned state_bits(unsigned m)
    {
        unsigned s = 0;
        if (m & (1u << 0)) s |= 0x3u;
        if (m & (1u << 1)) s |= 0x4u;
        if (m & (1u << 2)) s |= 0x8u;
        if (m & (1u << 3)) s |= 0x10u;
        if (m & (1u << 4)) s |= 0x20u;
        if (m & (1u << 5)) s |= 0xc0u;
        if (m & 0x7u) s |= (1u << CAI) | (1u << ECA);
        return s;
    }
    /* INITIAL blocks in mechanism order.  A calcium mechanism that writes cai
     * or does not read it first sets it to its resting value; every calcium
     * mechanism sets eca from cai (CuInitModel_*, Code/OneKernel/src/AllModels.cu:202-296,
     * Figures/FigureS4_mainen/CParsed/AllModels.cu) */
    template <class PF> NG_DEV static void init(unsigned m, double v, double *st, const PF &P)
    {
        if (m & (1u << 0)) {
            st[ECA] = eca_of(st[CAI]);
            Mech_ca M;
            M.v = v;
            M.celsius = 37.0;
            M.vshift = 0.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(0);
            M.cao = P(1);
            M.m = st[0];
            M.h = st[1];
            M.eca = st[ECA];
            M.cai = st[CAI];
            M.initial();
            st[0] = M.m;
            st[1] = M.h;
        }
        if (m & (1u << 1)) {
            st[CAI] = 5e-05;
            st[ECA] = eca_of(st[CAI]);
            Mech_cad M;
            M.v = v;
            M.depth = 0.09334562733124982;
            M.taur = 200.0;
            M.cainf = 0.0001;
            M.ca = st[2];
            M.ica = 0.0;
            M.cai = st[CAI];
            M.initial();
            st[2] = M.ca;
            st[CAI] = M.cai;
        }
        if (m & (1u << 2)) {
            st[ECA] = eca_of(st[CAI]);
            Mech_kca M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(2);
            M.caix = P(3);
            M.Ra = P(4);
            M.Rb = P(5);
            M.n = st[3];
            M.ek = -90.0;
            M.cai = st[CAI];
            M.initial();
            st[3] = M.n;
        }
        if (m & (1u << 3)) {
            Mech_km M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(6);
            M.tha = P(7);
            M.qa = P(8);
            M.Ra = P(9);
            M.Rb = P(10);
            M.n = st[4];
            M.ek = -90.0;
            M.initial();
            st[4] = M.n;
        }
        if (m & (1u << 4)) {
            Mech_kv M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(11);
            M.tha = P(12);
            M.qa = P(13);
            M.Ra = P(14);
            M.Rb = P(15);
            M.n = st[5];
            M.ek = -90.0;
            M.initial();
            st[5] = M.n;
        }
        if (m & (1u << 5)) {
            Mech_na M;
            M.v = v;
            M.celsius = 37.0;
            M.vshift = -5.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(16);
            M.tha = P(17);
            M.qa = P(18);
            M.Ra = P(19);
            M.Rb = P(20);
            M.thi1 = P(21);
            M.thi2 = P(22);
            M.qi = P(23);
            M.thinf = P(24);
            M.qinf = P(25);
            M.Rg = P(26);
            M.Rd = P(27);
            M.m = st[6];
            M.h = st[7];
            M.ena = 60.0;
            M.initial();
            st[6] = M.m;
            st[7] = M.h;
        }
        if (m & (1u << 6)) {
            Mech_pas M;
            M.v = v;
            M.g = P(28);
            M.e = P(29);
            M.initial();
        }
    }
    /* DERIVATIVE (cnexp) blocks at the new v, in mechanism order; ica is the
     * calcium current of the last BREAKPOINT */
    template <class PF> NG_DEV static void cnexp(unsigned m, double v, double dt, double *st, double ica, const PF &P)
    {
        (void)ica;
        if (m & (1u << 0)) {
            Mech_ca M;
            M.v = v;
            M.celsius = 37.0;
            M.vshift = 0.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(0);
            M.cao = P(1);
            M.m = st[0];
            M.h = st[1];
            M.eca = st[ECA];
            M.cai = st[CAI];
            M.dt = dt;
            M.states();
            st[0] = M.m;
            st[1] = M.h;
        }
        if (m & (1u << 1)) {
            Mech_cad M;
            M.v = v;
            M.depth = 0.09334562733124982;
            M.taur = 200.0;
            M.cainf = 0.0001;
            M.ca = st[2];
            M.ica = ica;
            M.cai = st[CAI];
            M.dt = dt;
            M.states();
            st[2] = M.ca;
            st[CAI] = M.cai;
        }
        if (m & (1u << 2)) {
            Mech_kca M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(2);
            M.caix = P(3);
            M.Ra = P(4);
            M.Rb = P(5);
            M.n = st[3];
            M.ek = -90.0;
            M.cai = st[CAI];
            M.dt = dt;
            M.states();
            st[3] = M.n;
        }
        if (m & (1u << 3)) {
            Mech_km M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(6);
            M.tha = P(7);
            M.qa = P(8);
            M.Ra = P(9);
            M.Rb = P(10);
            M.n = st[4];
            M.ek = -90.0;
            M.dt = dt;
            M.states();
            st[4] = M.n;
        }
        if (m & (1u << 4)) {
            Mech_kv M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(11);
            M.tha = P(12);
            M.qa = P(13);
            M.Ra = P(14);
            M.Rb = P(15);
            M.n = st[5];
            M.ek = -90.0;
            M.dt = dt;
            M.states();
            st[5] = M.n;
        }
        if (m & (1u << 5)) {
            Mech_na M;
            M.v = v;
            M.celsius = 37.0;
            M.vshift = -5.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(16);
            M.tha = P(17);
            M.qa = P(18);
            M.Ra = P(19);
            M.Rb = P(20);
            M.thi1 = P(21);
            M.thi2 = P(22);
            M.qi = P(23);
            M.thinf = P(24);
            M.qinf = P(25);
            M.Rg = P(26);
            M.Rd = P(27);
            M.m = st[6];
            M.h = st[7];
            M.ena = 60.0;
            M.dt = dt;
            M.states();
            st[6] = M.m;
            st[7] = M.h;
        }
    }
    /* BREAKPOINT blocks: the sum of the currents the mechanisms write, and
     * the calcium currents into ica when asked */
    template <class PF>
    NG_DEV static void current(unsigned m, double v, const double *stc, const PF &P, double &sum, double &ica, bool add_ica)
    {
        double *st = const_cast<double *>(stc);   /* read only: no block below stores a slot */
        double s = sum;
        (void)ica; (void)add_ica;
        if (m & (1u << 0)) {
            Mech_ca M;
            M.v = v;
            M.celsius = 37.0;
            M.vshift = 0.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(0);
            M.cao = P(1);
            M.m = st[0];
            M.h = st[1];
            M.eca = st[ECA];
            M.cai = st[CAI];
            M.breakpoint();
            s += M.ica;
            if (add_ica) ica += M.ica;
        }
        if (m & (1u << 2)) {
            Mech_kca M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(2);
            M.caix = P(3);
            M.Ra = P(4);
            M.Rb = P(5);
            M.n = st[3];
            M.ek = -90.0;
            M.cai = st[CAI];
            M.breakpoint();
            s += M.ik;
        }
        if (m & (1u << 3)) {
            Mech_km M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(6);
            M.tha = P(7);
            M.qa = P(8);
            M.Ra = P(9);
            M.Rb = P(10);
            M.n = st[4];
            M.ek = -90.0;
            M.breakpoint();
            s += M.ik;
        }
        if (m & (1u << 4)) {
            Mech_kv M;
            M.v = v;
            M.celsius = 37.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(11);
            M.tha = P(12);
            M.qa = P(13);
            M.Ra = P(14);
            M.Rb = P(15);
            M.n = st[5];
            M.ek = -90.0;
            M.breakpoint();
            s += M.ik;
        }
        if (m & (1u << 5)) {
            Mech_na M;
            M.v = v;
            M.celsius = 37.0;
            M.vshift = -5.0;
            M.temp = 23.0;
            M.q10 = 2.3;
            M.vmin = -120.0;
            M.vmax = 100.0;
            M.gbar = P(16);
            M.tha = P(17);
            M.qa = P(18);
            M.Ra = P(19);
            M.Rb = P(20);
            M.thi1 = P(21);
            M.thi2 = P(22);
            M.qi = P(23);
            M.thinf = P(24);
            M.qinf = P(25);
            M.Rg = P(26);
            M.Rd = P(27);
            M.m = st[6];
            M.h = st[7];
            M.ena = 60.0;
            M.breakpoint();
            s += M.ina;
        }
        if (m & (1u << 6)) {
            Mech_pas M;
            M.v = v;
            M.g = P(28);
            M.e = P(29);
            M.breakpoint();
            s += M.i;
        }
        sum = s;
    }
};

} // namespace ng_gen_mainen

using ng_gen_mainen::SetGen_mainen;
