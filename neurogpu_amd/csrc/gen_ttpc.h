/* Generated by neurogpu_amd.nmodl -- do not edit. */
#pragma once
#include "mechs.h"
#include <math.h>
#ifdef __HIPCC__
#define NG_HD_M __host__ __device__ inline
#else
#define NG_HD_M inline
#endif


namespace ng_gen_ttpc {

/* Ca_HVA: generated by neurogpu_amd.nmodl from Ca_HVA.mod */
struct Mech_Ca_HVA {
    double v = 0.0, dt = 0.0;
    double gCa_HVAbar = 1e-05;
    double m = 0.0;
    double h = 0.0;
    double eca = 0.0;
    double ica = 0.0;
    double gCa = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double mAlpha = 0.0;
    double mBeta = 0.0;
    double hInf = 0.0;
    double hTau = 0.0;
    double hAlpha = 0.0;
    double hBeta = 0.0;
    NG_HD_M void rates()
    {
        if ((v == -27.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-27.0 - v) * 0.2631578947368421));
        mAlpha = ng_div((0.055 * (-27.0 - v)), (exp_0 - 1.0));
        const double exp_1 = ng_expc(((-75.0 - v) * 0.058823529411764705));
        mBeta = (0.94 * exp_1);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = ng_rcp((mAlpha + mBeta));
        const double exp_2 = ng_expc(((-13.0 - v) * 0.02));
        hAlpha = (0.000457 * exp_2);
        const double exp_3 = ng_expc((((-v) - 15.0) * 0.03571428571428571));
        hBeta = ng_div(0.0065, (exp_3 + 1.0));
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
        hTau = ng_rcp((hAlpha + hBeta));
    }
    NG_HD_M void initial()
    {
        if ((v == -27.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-27.0 - v) * 0.2631578947368421));
        mAlpha = ng_div((0.055 * (-27.0 - v)), (exp_0 - 1.0));
        const double exp_1 = ng_expc(((-75.0 - v) * 0.058823529411764705));
        mBeta = (0.94 * exp_1);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = ng_rcp((mAlpha + mBeta));
        const double exp_2 = ng_expc(((-13.0 - v) * 0.02));
        hAlpha = (0.000457 * exp_2);
        const double exp_3 = ng_expc((((-v) - 15.0) * 0.03571428571428571));
        hBeta = ng_div(0.0065, (exp_3 + 1.0));
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
        hTau = ng_rcp((hAlpha + hBeta));
        m = mInf;
        h = hInf;
    }
    NG_HD_M void states()
    {
        if ((v == -27.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-27.0 - v) * 0.2631578947368421));
        mAlpha = ng_div((0.055 * (-27.0 - v)), (exp_0 - 1.0));
        const double exp_1 = ng_expc(((-75.0 - v) * 0.058823529411764705));
        mBeta = (0.94 * exp_1);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = ng_rcp((mAlpha + mBeta));
        const double exp_2 = ng_expc(((-13.0 - v) * 0.02));
        hAlpha = (0.000457 * exp_2);
        const double exp_3 = ng_expc((((-v) - 15.0) * 0.03571428571428571));
        hBeta = ng_div(0.0065, (exp_3 + 1.0));
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
        hTau = ng_rcp((hAlpha + hBeta));
        m = m + (1.0 - ng_expc(-(dt * (mAlpha + mBeta)))) * (mInf - m);
        h = h + (1.0 - ng_expc(-(dt * (hAlpha + hBeta)))) * (hInf - h);
    }
    NG_HD_M void breakpoint()
    {
        gCa = (((gCa_HVAbar * m) * m) * h);
        ica = (gCa * (v - eca));
    }
};

/* Ca_LVAst: generated by neurogpu_amd.nmodl from Ca_LVAst.mod */
struct Mech_Ca_LVAst {
    double v = 0.0, dt = 0.0;
    double gCa_LVAstbar = 1e-05;
    double m = 0.0;
    double h = 0.0;
    double eca = 0.0;
    double ica = 0.0;
    double gCa_LVAst = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double hInf = 0.0;
    double hTau = 0.0;
    NG_HD_M void rates()
    {
        double qt = 0.0;
        qt = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((v - -30.0) * -0.16666666666666666));
        mInf = ng_rcp((1.0 + exp_0));
        const double exp_1 = ng_expc(((v - -25.0) * 0.2));
        mTau = ((5.0 + ng_div(20.0, (1.0 + exp_1))) * 0.3386521313023745);
        const double exp_2 = ng_expc(((v - -80.0) * 0.15625));
        hInf = ng_rcp((1.0 + exp_2));
        const double exp_3 = ng_expc(((v - -40.0) * 0.14285714285714285));
        hTau = ((20.0 + ng_div(50.0, (1.0 + exp_3))) * 0.3386521313023745);
        v = (v - 10.0);
    }
    NG_HD_M void initial()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((v - -30.0) * -0.16666666666666666));
        mInf = ng_rcp((1.0 + exp_0));
        const double exp_1 = ng_expc(((v - -25.0) * 0.2));
        mTau = ((5.0 + ng_div(20.0, (1.0 + exp_1))) * 0.3386521313023745);
        const double exp_2 = ng_expc(((v - -80.0) * 0.15625));
        hInf = ng_rcp((1.0 + exp_2));
        const double exp_3 = ng_expc(((v - -40.0) * 0.14285714285714285));
        hTau = ((20.0 + ng_div(50.0, (1.0 + exp_3))) * 0.3386521313023745);
        v = (v - 10.0);
        m = mInf;
        h = hInf;
    }
    NG_HD_M void states()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((v - -30.0) * -0.16666666666666666));
        mInf = ng_rcp((1.0 + exp_0));
        const double exp_1 = ng_expc(((v - -25.0) * 0.2));
        mTau = ((5.0 + ng_div(20.0, (1.0 + exp_1))) * 0.3386521313023745);
        const double exp_2 = ng_expc(((v - -80.0) * 0.15625));
        hInf = ng_rcp((1.0 + exp_2));
        const double exp_3 = ng_expc(((v - -40.0) * 0.14285714285714285));
        hTau = ((20.0 + ng_div(50.0, (1.0 + exp_3))) * 0.3386521313023745);
        v = (v - 10.0);
        m = m + (1.0 - ng_expc(-(dt * ng_rcp(mTau)))) * (mInf - m);
        h = h + (1.0 - ng_expc(-(dt * ng_rcp(hTau)))) * (hInf - h);
    }
    NG_HD_M void breakpoint()
    {
        gCa_LVAst = (((gCa_LVAstbar * m) * m) * h);
        ica = (gCa_LVAst * (v - eca));
    }
};

/* CaDynamics_E2: generated by neurogpu_amd.nmodl from CaDynamics_E2.mod */
struct Mech_CaDynamics_E2 {
    double v = 0.0, dt = 0.0;
    double gamma = 0.05;
    double decay = 80.0;
    double depth = 0.1;
    double minCai = 0.0001;
    static constexpr double FARADAY = 96485.309;
    double cai = 0.0;
    double ica = 0.0;
    NG_HD_M void initial()
    {
    }
    NG_HD_M void states()
    {
        cai = cai + (1.0 - ng_expc(-(dt * ng_rcp(decay)))) * ((((-10000.0 * ng_div((ica * gamma), (192970.618 * depth))) - ng_div((-minCai), decay)) * decay) - cai);
    }
    NG_HD_M void breakpoint()
    {
    }
};

/* Ih: generated by neurogpu_amd.nmodl from Ih.mod */
struct Mech_Ih {
    double v = 0.0, dt = 0.0;
    double gIhbar = 1e-05;
    double ehcn = -45.0;
    double m = 0.0;
    double ihcn = 0.0;
    double gIh = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double mAlpha = 0.0;
    double mBeta = 0.0;
    NG_HD_M void rates()
    {
        if ((v == -154.9)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((v + 154.9) * 0.08403361344537814));
        mAlpha = ng_div((0.00643 * (v + 154.9)), (exp_0 - 1.0));
        const double exp_1 = ng_expc((v * 0.030211480362537763));
        mBeta = (0.193 * exp_1);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = ng_rcp((mAlpha + mBeta));
    }
    NG_HD_M void initial()
    {
        if ((v == -154.9)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((v + 154.9) * 0.08403361344537814));
        mAlpha = ng_div((0.00643 * (v + 154.9)), (exp_0 - 1.0));
        const double exp_1 = ng_expc((v * 0.030211480362537763));
        mBeta = (0.193 * exp_1);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = ng_rcp((mAlpha + mBeta));
        m = mInf;
    }
    NG_HD_M void states()
    {
        if ((v == -154.9)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((v + 154.9) * 0.08403361344537814));
        mAlpha = ng_div((0.00643 * (v + 154.9)), (exp_0 - 1.0));
        const double exp_1 = ng_expc((v * 0.030211480362537763));
        mBeta = (0.193 * exp_1);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = ng_rcp((mAlpha + mBeta));
        m = m + (1.0 - ng_expc(-(dt * (mAlpha + mBeta)))) * (mInf - m);
    }
    NG_HD_M void breakpoint()
    {
        gIh = (gIhbar * m);
        ihcn = (gIh * (v - ehcn));
    }
};

/* Im: generated by neurogpu_amd.nmodl from Im.mod */
struct Mech_Im {
    double v = 0.0, dt = 0.0;
    double gImbar = 1e-05;
    double m = 0.0;
    double ek = 0.0;
    double ik = 0.0;
    double gIm = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double mAlpha = 0.0;
    double mBeta = 0.0;
    NG_HD_M void rates()
    {
        double qt = 0.0;
        qt = 2.952882641412121;
        const double exp_0 = ng_expc((0.1 * (v - -35.0)));
        mAlpha = (0.0033 * exp_0);
        mBeta = (0.0033 * ng_rcp(exp_0));
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
    }
    NG_HD_M void initial()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        const double exp_0 = ng_expc((0.1 * (v - -35.0)));
        mAlpha = (0.0033 * exp_0);
        mBeta = (0.0033 * ng_rcp(exp_0));
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
        m = mInf;
    }
    NG_HD_M void states()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        const double exp_0 = ng_expc((0.1 * (v - -35.0)));
        mAlpha = (0.0033 * exp_0);
        mBeta = (0.0033 * ng_rcp(exp_0));
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
        m = m + (1.0 - ng_expc(-(dt * ((mAlpha + mBeta) * 2.952882641412121)))) * (mInf - m);
    }
    NG_HD_M void breakpoint()
    {
        gIm = (gImbar * m);
        ik = (gIm * (v - ek));
    }
};

/* K_Pst: generated by neurogpu_amd.nmodl from K_Pst.mod */
struct Mech_K_Pst {
    double v = 0.0, dt = 0.0;
    double gK_Pstbar = 1e-05;
    double m = 0.0;
    double h = 0.0;
    double ek = 0.0;
    double ik = 0.0;
    double gK_Pst = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double hInf = 0.0;
    double hTau = 0.0;
    NG_HD_M void rates()
    {
        double qt = 0.0;
        qt = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((-(v + 1.0)) * 0.08333333333333333));
        mInf = ng_rcp((1.0 + exp_0));
        if ((v < -50.0)) {
            const double exp_1 = ng_expc(((-v) * -0.026));
            mTau = ((1.25 + (175.03 * exp_1)) * 0.3386521313023745);
        } else {
            const double exp_2 = ng_expc(((-v) * 0.026));
            mTau = ((1.25 + (13.0 * exp_2)) * 0.3386521313023745);
        }
        const double exp_3 = ng_expc(((-(v + 54.0)) * -0.09090909090909091));
        hInf = ng_rcp((1.0 + exp_3));
        hTau = ((360.0 + ((1010.0 + (24.0 * (v + 55.0))) * ng_expc((-ng_pow(((v + 75.0) * 0.020833333333333332), 2.0))))) * 0.3386521313023745);
        v = (v - 10.0);
    }
    NG_HD_M void initial()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((-(v + 1.0)) * 0.08333333333333333));
        mInf = ng_rcp((1.0 + exp_0));
        if ((v < -50.0)) {
            const double exp_1 = ng_expc(((-v) * -0.026));
            mTau = ((1.25 + (175.03 * exp_1)) * 0.3386521313023745);
        } else {
            const double exp_2 = ng_expc(((-v) * 0.026));
            mTau = ((1.25 + (13.0 * exp_2)) * 0.3386521313023745);
        }
        const double exp_3 = ng_expc(((-(v + 54.0)) * -0.09090909090909091));
        hInf = ng_rcp((1.0 + exp_3));
        hTau = ((360.0 + ((1010.0 + (24.0 * (v + 55.0))) * ng_expc((-ng_pow(((v + 75.0) * 0.020833333333333332), 2.0))))) * 0.3386521313023745);
        v = (v - 10.0);
        m = mInf;
        h = hInf;
    }
    NG_HD_M void states()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((-(v + 1.0)) * 0.08333333333333333));
        mInf = ng_rcp((1.0 + exp_0));
        if ((v < -50.0)) {
            const double exp_1 = ng_expc(((-v) * -0.026));
            mTau = ((1.25 + (175.03 * exp_1)) * 0.3386521313023745);
        } else {
            const double exp_2 = ng_expc(((-v) * 0.026));
            mTau = ((1.25 + (13.0 * exp_2)) * 0.3386521313023745);
        }
        const double exp_3 = ng_expc(((-(v + 54.0)) * -0.09090909090909091));
        hInf = ng_rcp((1.0 + exp_3));
        hTau = ((360.0 + ((1010.0 + (24.0 * (v + 55.0))) * ng_expc((-ng_pow(((v + 75.0) * 0.020833333333333332), 2.0))))) * 0.3386521313023745);
        v = (v - 10.0);
        m = m + (1.0 - ng_expc(-(dt * ng_rcp(mTau)))) * (mInf - m);
        h = h + (1.0 - ng_expc(-(dt * ng_rcp(hTau)))) * (hInf - h);
    }
    NG_HD_M void breakpoint()
    {
        gK_Pst = (((gK_Pstbar * m) * m) * h);
        ik = (gK_Pst * (v - ek));
    }
};

/* K_Tst: generated by neurogpu_amd.nmodl from K_Tst.mod */
struct Mech_K_Tst {
    double v = 0.0, dt = 0.0;
    double gK_Tstbar = 1e-05;
    double m = 0.0;
    double h = 0.0;
    double ek = 0.0;
    double ik = 0.0;
    double gK_Tst = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double hInf = 0.0;
    double hTau = 0.0;
    NG_HD_M void rates()
    {
        double qt = 0.0;
        qt = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((-(v + 0.0)) * 0.05263157894736842));
        mInf = ng_rcp((1.0 + exp_0));
        mTau = ((0.34 + (0.92 * ng_expc((-ng_pow(((v + 71.0) * 0.01694915254237288), 2.0))))) * 0.3386521313023745);
        const double exp_1 = ng_expc(((-(v + 66.0)) * -0.1));
        hInf = ng_rcp((1.0 + exp_1));
        hTau = ((8.0 + (49.0 * ng_expc((-ng_pow(((v + 73.0) * 0.043478260869565216), 2.0))))) * 0.3386521313023745);
        v = (v - 10.0);
    }
    NG_HD_M void initial()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((-(v + 0.0)) * 0.05263157894736842));
        mInf = ng_rcp((1.0 + exp_0));
        mTau = ((0.34 + (0.92 * ng_expc((-ng_pow(((v + 71.0) * 0.01694915254237288), 2.0))))) * 0.3386521313023745);
        const double exp_1 = ng_expc(((-(v + 66.0)) * -0.1));
        hInf = ng_rcp((1.0 + exp_1));
        hTau = ((8.0 + (49.0 * ng_expc((-ng_pow(((v + 73.0) * 0.043478260869565216), 2.0))))) * 0.3386521313023745);
        v = (v - 10.0);
        m = mInf;
        h = hInf;
    }
    NG_HD_M void states()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        v = (v + 10.0);
        const double exp_0 = ng_expc(((-(v + 0.0)) * 0.05263157894736842));
        mInf = ng_rcp((1.0 + exp_0));
        mTau = ((0.34 + (0.92 * ng_expc((-ng_pow(((v + 71.0) * 0.01694915254237288), 2.0))))) * 0.3386521313023745);
        const double exp_1 = ng_expc(((-(v + 66.0)) * -0.1));
        hInf = ng_rcp((1.0 + exp_1));
        hTau = ((8.0 + (49.0 * ng_expc((-ng_pow(((v + 73.0) * 0.043478260869565216), 2.0))))) * 0.3386521313023745);
        v = (v - 10.0);
        m = m + (1.0 - ng_expc(-(dt * ng_rcp(mTau)))) * (mInf - m);
        h = h + (1.0 - ng_expc(-(dt * ng_rcp(hTau)))) * (hInf - h);
    }
    NG_HD_M void breakpoint()
    {
        gK_Tst = ((gK_Tstbar * ng_pow(m, 4.0)) * h);
        ik = (gK_Tst * (v - ek));
    }
};

/* Nap_Et2: generated by neurogpu_amd.nmodl from Nap_Et2.mod */
struct Mech_Nap_Et2 {
    double v = 0.0, dt = 0.0;
    double gNap_Et2bar = 1e-05;
    double m = 0.0;
    double h = 0.0;
    double ena = 0.0;
    double ina = 0.0;
    double gNap_Et2 = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double mAlpha = 0.0;
    double mBeta = 0.0;
    double hInf = 0.0;
    double hTau = 0.0;
    double hAlpha = 0.0;
    double hBeta = 0.0;
    NG_HD_M void rates()
    {
        double qt = 0.0;
        qt = 2.952882641412121;
        const double exp_0 = ng_expc(((v - -52.6) * -0.2173913043478261));
        mInf = ng_rcp((1.0 + exp_0));
        if ((v == -38.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((-(v - -38.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -38.0)), (1.0 - exp_1));
        mBeta = ng_div((0.124 * ((-v) - 38.0)), (1.0 - ng_rcp(exp_1)));
        mTau = ((6.0 * ng_rcp((mAlpha + mBeta))) * 0.3386521313023745);
        if ((v == -17.0)) {
            v = (v + 0.0001);
        }
        if ((v == -64.4)) {
            v = (v + 0.0001);
        }
        const double exp_2 = ng_expc(((v - -48.8) * 0.1));
        hInf = ng_rcp((1.0 + exp_2));
        const double exp_3 = ng_expc(((v + 17.0) * 0.2159827213822894));
        hAlpha = ng_div((-2.88e-06 * (v + 17.0)), (1.0 - exp_3));
        const double exp_4 = ng_expc(((-(v + 64.4)) * 0.3802281368821293));
        hBeta = ng_div((6.94e-06 * (v + 64.4)), (1.0 - exp_4));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
    }
    NG_HD_M void initial()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        const double exp_0 = ng_expc(((v - -52.6) * -0.2173913043478261));
        mInf = ng_rcp((1.0 + exp_0));
        if ((v == -38.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((-(v - -38.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -38.0)), (1.0 - exp_1));
        mBeta = ng_div((0.124 * ((-v) - 38.0)), (1.0 - ng_rcp(exp_1)));
        mTau = ((6.0 * ng_rcp((mAlpha + mBeta))) * 0.3386521313023745);
        if ((v == -17.0)) {
            v = (v + 0.0001);
        }
        if ((v == -64.4)) {
            v = (v + 0.0001);
        }
        const double exp_2 = ng_expc(((v - -48.8) * 0.1));
        hInf = ng_rcp((1.0 + exp_2));
        const double exp_3 = ng_expc(((v + 17.0) * 0.2159827213822894));
        hAlpha = ng_div((-2.88e-06 * (v + 17.0)), (1.0 - exp_3));
        const double exp_4 = ng_expc(((-(v + 64.4)) * 0.3802281368821293));
        hBeta = ng_div((6.94e-06 * (v + 64.4)), (1.0 - exp_4));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
        m = mInf;
        h = hInf;
    }
    NG_HD_M void states()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        const double exp_0 = ng_expc(((v - -52.6) * -0.2173913043478261));
        mInf = ng_rcp((1.0 + exp_0));
        if ((v == -38.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((-(v - -38.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -38.0)), (1.0 - exp_1));
        mBeta = ng_div((0.124 * ((-v) - 38.0)), (1.0 - ng_rcp(exp_1)));
        mTau = ((6.0 * ng_rcp((mAlpha + mBeta))) * 0.3386521313023745);
        if ((v == -17.0)) {
            v = (v + 0.0001);
        }
        if ((v == -64.4)) {
            v = (v + 0.0001);
        }
        const double exp_2 = ng_expc(((v - -48.8) * 0.1));
        hInf = ng_rcp((1.0 + exp_2));
        const double exp_3 = ng_expc(((v + 17.0) * 0.2159827213822894));
        hAlpha = ng_div((-2.88e-06 * (v + 17.0)), (1.0 - exp_3));
        const double exp_4 = ng_expc(((-(v + 64.4)) * 0.3802281368821293));
        hBeta = ng_div((6.94e-06 * (v + 64.4)), (1.0 - exp_4));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
        m = m + (1.0 - ng_expc(-(dt * ng_rcp(mTau)))) * (mInf - m);
        h = h + (1.0 - ng_expc(-(dt * ((hAlpha + hBeta) * 2.952882641412121)))) * (hInf - h);
    }
    NG_HD_M void breakpoint()
    {
        gNap_Et2 = ((((gNap_Et2bar * m) * m) * m) * h);
        ina = (gNap_Et2 * (v - ena));
    }
};

/* NaTa_t: generated by neurogpu_amd.nmodl from NaTa_t.mod */
struct Mech_NaTa_t {
    double v = 0.0, dt = 0.0;
    double gNaTa_tbar = 1e-05;
    double m = 0.0;
    double h = 0.0;
    double ena = 0.0;
    double ina = 0.0;
    double gNaTa_t = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double mAlpha = 0.0;
    double mBeta = 0.0;
    double hInf = 0.0;
    double hTau = 0.0;
    double hAlpha = 0.0;
    double hBeta = 0.0;
    NG_HD_M void rates()
    {
        double qt = 0.0;
        qt = 2.952882641412121;
        if ((v == -38.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-(v - -38.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -38.0)), (1.0 - exp_0));
        mBeta = ng_div((0.124 * ((-v) - 38.0)), (1.0 - ng_rcp(exp_0)));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        if ((v == -66.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((v - -66.0) * 0.16666666666666666));
        hAlpha = ng_div((-0.015 * (v - -66.0)), (1.0 - exp_1));
        hBeta = ng_div((-0.015 * ((-v) - 66.0)), (1.0 - ng_rcp(exp_1)));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
    }
    NG_HD_M void initial()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        if ((v == -38.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-(v - -38.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -38.0)), (1.0 - exp_0));
        mBeta = ng_div((0.124 * ((-v) - 38.0)), (1.0 - ng_rcp(exp_0)));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        if ((v == -66.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((v - -66.0) * 0.16666666666666666));
        hAlpha = ng_div((-0.015 * (v - -66.0)), (1.0 - exp_1));
        hBeta = ng_div((-0.015 * ((-v) - 66.0)), (1.0 - ng_rcp(exp_1)));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
        m = mInf;
        h = hInf;
    }
    NG_HD_M void states()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        if ((v == -38.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-(v - -38.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -38.0)), (1.0 - exp_0));
        mBeta = ng_div((0.124 * ((-v) - 38.0)), (1.0 - ng_rcp(exp_0)));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        if ((v == -66.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((v - -66.0) * 0.16666666666666666));
        hAlpha = ng_div((-0.015 * (v - -66.0)), (1.0 - exp_1));
        hBeta = ng_div((-0.015 * ((-v) - 66.0)), (1.0 - ng_rcp(exp_1)));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
        m = m + (1.0 - ng_expc(-(dt * ((mAlpha + mBeta) * 2.952882641412121)))) * (mInf - m);
        h = h + (1.0 - ng_expc(-(dt * ((hAlpha + hBeta) * 2.952882641412121)))) * (hInf - h);
    }
    NG_HD_M void breakpoint()
    {
        gNaTa_t = ((((gNaTa_tbar * m) * m) * m) * h);
        ina = (gNaTa_t * (v - ena));
    }
};

/* NaTs2_t: generated by neurogpu_amd.nmodl from NaTs2_t.mod */
struct Mech_NaTs2_t {
    double v = 0.0, dt = 0.0;
    double gNaTs2_tbar = 1e-05;
    double m = 0.0;
    double h = 0.0;
    double ena = 0.0;
    double ina = 0.0;
    double gNaTs2_t = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    double mAlpha = 0.0;
    double mBeta = 0.0;
    double hInf = 0.0;
    double hTau = 0.0;
    double hAlpha = 0.0;
    double hBeta = 0.0;
    NG_HD_M void rates()
    {
        double qt = 0.0;
        qt = 2.952882641412121;
        if ((v == -32.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-(v - -32.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -32.0)), (1.0 - exp_0));
        mBeta = ng_div((0.124 * ((-v) - 32.0)), (1.0 - ng_rcp(exp_0)));
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
        if ((v == -60.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((v - -60.0) * 0.16666666666666666));
        hAlpha = ng_div((-0.015 * (v - -60.0)), (1.0 - exp_1));
        hBeta = ng_div((-0.015 * ((-v) - 60.0)), (1.0 - ng_rcp(exp_1)));
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
    }
    NG_HD_M void initial()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        if ((v == -32.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-(v - -32.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -32.0)), (1.0 - exp_0));
        mBeta = ng_div((0.124 * ((-v) - 32.0)), (1.0 - ng_rcp(exp_0)));
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
        if ((v == -60.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((v - -60.0) * 0.16666666666666666));
        hAlpha = ng_div((-0.015 * (v - -60.0)), (1.0 - exp_1));
        hBeta = ng_div((-0.015 * ((-v) - 60.0)), (1.0 - ng_rcp(exp_1)));
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
        m = mInf;
        h = hInf;
    }
    NG_HD_M void states()
    {
        double qt_l0 = 0.0;
        qt_l0 = 2.952882641412121;
        if ((v == -32.0)) {
            v = (v + 0.0001);
        }
        const double exp_0 = ng_expc(((-(v - -32.0)) * 0.16666666666666666));
        mAlpha = ng_div((0.182 * (v - -32.0)), (1.0 - exp_0));
        mBeta = ng_div((0.124 * ((-v) - 32.0)), (1.0 - ng_rcp(exp_0)));
        mInf = ng_div(mAlpha, (mAlpha + mBeta));
        mTau = (ng_rcp((mAlpha + mBeta)) * 0.3386521313023745);
        if ((v == -60.0)) {
            v = (v + 0.0001);
        }
        const double exp_1 = ng_expc(((v - -60.0) * 0.16666666666666666));
        hAlpha = ng_div((-0.015 * (v - -60.0)), (1.0 - exp_1));
        hBeta = ng_div((-0.015 * ((-v) - 60.0)), (1.0 - ng_rcp(exp_1)));
        hInf = ng_div(hAlpha, (hAlpha + hBeta));
        hTau = (ng_rcp((hAlpha + hBeta)) * 0.3386521313023745);
        m = m + (1.0 - ng_expc(-(dt * ((mAlpha + mBeta) * 2.952882641412121)))) * (mInf - m);
        h = h + (1.0 - ng_expc(-(dt * ((hAlpha + hBeta) * 2.952882641412121)))) * (hInf - h);
    }
    NG_HD_M void breakpoint()
    {
        gNaTs2_t = ((((gNaTs2_tbar * m) * m) * m) * h);
        ina = (gNaTs2_t * (v - ena));
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

/* SK_E2: generated by neurogpu_amd.nmodl from SK_E2.mod */
struct Mech_SK_E2 {
    double v = 0.0, dt = 0.0;
    double gSK_E2bar = 1e-06;
    double zTau = 1.0;
    double ek = 0.0;
    double cai = 0.0;
    double z = 0.0;
    double zInf = 0.0;
    double ik = 0.0;
    double gSK_E2 = 0.0;
    NG_HD_M void rates(double ca)
    {
        if ((ca < 1e-07)) {
            ca = (ca + 1e-07);
        }
        zInf = ng_rcp((1.0 + ng_expc((4.8 * ng_logc(ng_div(0.00043, ca))))));
    }
    NG_HD_M void initial()
    {
        rates(cai);
        z = zInf;
    }
    NG_HD_M void states()
    {
        rates(cai);
        z = z + (1.0 - ng_expc(-(dt * ng_rcp(zTau)))) * (zInf - z);
    }
    NG_HD_M void breakpoint()
    {
        gSK_E2 = (gSK_E2bar * z);
        ik = (gSK_E2 * (v - ek));
    }
};

/* SKv3_1: generated by neurogpu_amd.nmodl from SKv3_1.mod */
struct Mech_SKv3_1 {
    double v = 0.0, dt = 0.0;
    double gSKv3_1bar = 1e-05;
    double m = 0.0;
    double ek = 0.0;
    double ik = 0.0;
    double gSKv3_1 = 0.0;
    double mInf = 0.0;
    double mTau = 0.0;
    NG_HD_M void rates()
    {
        const double exp_0 = ng_expc(((v - 18.7) * -0.10309278350515465));
        mInf = ng_rcp((1.0 + exp_0));
        const double exp_1 = ng_expc(((v - -46.56) * -0.022655188038060714));
        mTau = ng_div(4.0, (1.0 + exp_1));
    }
    NG_HD_M void initial()
    {
        const double exp_0 = ng_expc(((v - 18.7) * -0.10309278350515465));
        mInf = ng_rcp((1.0 + exp_0));
        const double exp_1 = ng_expc(((v - -46.56) * -0.022655188038060714));
        mTau = ng_div(4.0, (1.0 + exp_1));
        m = mInf;
    }
    NG_HD_M void states()
    {
        const double exp_0 = ng_expc(((v - 18.7) * -0.10309278350515465));
        mInf = ng_rcp((1.0 + exp_0));
        const double exp_1 = ng_expc(((v - -46.56) * -0.022655188038060714));
        mTau = ng_div(4.0, (1.0 + exp_1));
        m = m + (1.0 - ng_expc(-(dt * ((1.0 + exp_1) * 0.25)))) * (mInf - m);
    }
    NG_HD_M void breakpoint()
    {
        gSKv3_1 = (gSKv3_1bar * m);
        ik = (gSKv3_1 * (v - ek));
    }
};

/* SetGen_ttpc: the simulation kernel's mechanism set over the generated mechanisms
 * Ca_HVA, Ca_LVAst, CaDynamics_E2, Ih, Im, K_Pst, K_Tst, Nap_Et2, NaTa_t, NaTs2_t, pas, SK_E2, SKv3_1 (bit k = mechanism k); generated by neurogpu_amd.nmodl. */
struct SetGen_ttpc {
    static constexpr int NMECH = 13, NPARAM = 19, NSTATE = 21;
    static constexpr int CAI = 18, ECA = 19, ICA = 20;
    /* first parameter-matrix column of each mechanism, then the total */
    static constexpr int PBASE[14] = {0, 1, 2, 6, 8, 9, 10, 11, 12, 13, 14, 16, 18, 19};
    template <class F> NG_DEV static void dispatch(unsigned m, F &&f)
    {
        ng_dispatch(m, f, NgMasks<0x408u, 0x1618u>{});
    }
    NG_DEV static unsigned ica_bits(unsigned m) { return (m & 0x4u) ? 1u << ICA : 0u; }
    NG_DEV static double eca_of(double cai) { return 13.23238136699132 * ng_logc(ng_div(2.0, cai)); }
    NG_DEV static void eca_after_cnexp(unsigned m, double *st)
    {
        if (m & 0x807u) st[ECA] = eca_of(st[CAI]);
    }
    NG_DEV static unsigned state_bits(unsigned m)
    {
        unsigned s = 0;
        if (m & (1u << 0)) s |= 0x3u;
        if (m & (1u << 1)) s |= 0xcu;
        if (m & (1u << 3)) s |= 0x10u;
        if (m & (1u << 4)) s |= 0x20u;
        if (m & (1u << 5)) s |= 0xc0u;
        if (m & (1u << 6)) s |= 0x300u;
        if (m & (1u << 7)) s |= 0xc00u;
        if (m & (1u << 8)) s |= 0x3000u;
        if (m & (1u << 9)) s |= 0xc000u;
        if (m & (1u << 11)) s |= 0x10000u;
        if (m & (1u << 12)) s |= 0x20000u;
        if (m & 0x807u) s |= (1u << CAI) | (1u << ECA);
        return s;
    }
    /* INITIAL blocks in mechanism order.  A calcium mechanism that writes cai
     * or does not read it first sets it to its resting value; every calcium
     * mechanism sets eca from cai (CuInitModel_*, Code/OneKernel/src/AllModels.cu:202-296,
     * Figures/FigureS4_mainen/CParsed/AllModels.cu) */
    template <class PF> NG_DEV static void init(unsigned m, double v, double *st, const PF &P)
    {
        if (m & (1u << 0)) {
            st[CAI] = 5e-05;
            st[ECA] = eca_of(st[CAI]);
            Mech_Ca_HVA M;
            M.v = v;
            M.gCa_HVAbar = P(0);
            M.m = st[0];
            M.h = st[1];
            M.eca = st[ECA];
            M.initial();
            st[0] = M.m;
            st[1] = M.h;
        }
        if (m & (1u << 1)) {
            st[CAI] = 5e-05;
            st[ECA] = eca_of(st[CAI]);
            Mech_Ca_LVAst M;
            M.v = v;
            M.gCa_LVAstbar = P(1);
            M.m = st[2];
            M.h = st[3];
            M.eca = st[ECA];
            M.initial();
            st[2] = M.m;
            st[3] = M.h;
        }
        if (m & (1u << 2)) {
            st[CAI] = 5e-05;
            st[ECA] = eca_of(st[CAI]);
            Mech_CaDynamics_E2 M;
            M.v = v;
            M.gamma = P(2);
            M.decay = P(3);
            M.depth = P(4);
            M.minCai = P(5);
            M.ica = 0.0;
            M.cai = st[CAI];
            M.initial();
            st[CAI] = M.cai;
        }
        if (m & (1u << 3)) {
            Mech_Ih M;
            M.v = v;
            M.gIhbar = P(6);
            M.ehcn = P(7);
            M.m = st[4];
            M.initial();
            st[4] = M.m;
        }
        if (m & (1u << 4)) {
            Mech_Im M;
            M.v = v;
            M.gImbar = P(8);
            M.m = st[5];
            M.ek = -85.0;
            M.initial();
            st[5] = M.m;
        }
        if (m & (1u << 5)) {
            Mech_K_Pst M;
            M.v = v;
            M.gK_Pstbar = P(9);
            M.m = st[6];
            M.h = st[7];
            M.ek = -85.0;
            M.initial();
            st[6] = M.m;
            st[7] = M.h;
        }
        if (m & (1u << 6)) {
            Mech_K_Tst M;
            M.v = v;
            M.gK_Tstbar = P(10);
            M.m = st[8];
            M.h = st[9];
            M.ek = -85.0;
            M.initial();
            st[8] = M.m;
            st[9] = M.h;
        }
        if (m & (1u << 7)) {
            Mech_Nap_Et2 M;
            M.v = v;
            M.gNap_Et2bar = P(11);
            M.m = st[10];
            M.h = st[11];
            M.ena = 50.0;
            M.initial();
            st[10] = M.m;
            st[11] = M.h;
        }
        if (m & (1u << 8)) {
            Mech_NaTa_t M;
            M.v = v;
            M.gNaTa_tbar = P(12);
            M.m = st[12];
            M.h = st[13];
            M.ena = 50.0;
            M.initial();
            st[12] = M.m;
            st[13] = M.h;
        }
        if (m & (1u << 9)) {
            Mech_NaTs2_t M;
            M.v = v;
            M.gNaTs2_tbar = P(13);
            M.m = st[14];
            M.h = st[15];
            M.ena = 50.0;
            M.initial();
            st[14] = M.m;
            st[15] = M.h;
        }
        if (m & (1u << 10)) {
            Mech_pas M;
            M.v = v;
            M.g = P(14);
            M.e = P(15);
            M.initial();
        }
        if (m & (1u << 11)) {
            st[ECA] = eca_of(st[CAI]);
            Mech_SK_E2 M;
            M.v = v;
            M.gSK_E2bar = P(16);
            M.zTau = P(17);
            M.z = st[16];
            M.ek = -85.0;
            M.cai = st[CAI];
            M.initial();
            st[16] = M.z;
        }
        if (m & (1u << 12)) {
            Mech_SKv3_1 M;
            M.v = v;
            M.gSKv3_1bar = P(18);
            M.m = st[17];
            M.ek = -85.0;
            M.initial();
            st[17] = M.m;
        }
    }
    /* DERIVATIVE (cnexp) blocks at the new v, in mechanism order; ica is the
     * calcium current of the last BREAKPOINT */
    template <class PF> NG_DEV static void cnexp(unsigned m, double v, double dt, double *st, double ica, const PF &P)
    {
        (void)ica;
        if (m & (1u << 0)) {
            Mech_Ca_HVA M;
            M.v = v;
            M.gCa_HVAbar = P(0);
            M.m = st[0];
            M.h = st[1];
            M.eca = st[ECA];
            M.dt = dt;
            M.states();
            st[0] = M.m;
            st[1] = M.h;
        }
        if (m & (1u << 1)) {
            Mech_Ca_LVAst M;
            M.v = v;
            M.gCa_LVAstbar = P(1);
            M.m = st[2];
            M.h = st[3];
            M.eca = st[ECA];
            M.dt = dt;
            M.states();
            st[2] = M.m;
            st[3] = M.h;
        }
        if (m & (1u << 2)) {
            Mech_CaDynamics_E2 M;
            M.v = v;
            M.gamma = P(2);
            M.decay = P(3);
            M.depth = P(4);
            M.minCai = P(5);
            M.ica = ica;
            M.cai = st[CAI];
            M.dt = dt;
            M.states();
            st[CAI] = M.cai;
        }
        if (m & (1u << 3)) {
            Mech_Ih M;
            M.v = v;
            M.gIhbar = P(6);
            M.ehcn = P(7);
            M.m = st[4];
            M.dt = dt;
            M.states();
            st[4] = M.m;
        }
        if (m & (1u << 4)) {
            Mech_Im M;
            M.v = v;
            M.gImbar = P(8);
            M.m = st[5];
            M.ek = -85.0;
            M.dt = dt;
            M.states();
            st[5] = M.m;
        }
        if (m & (1u << 5)) {
            Mech_K_Pst M;
            M.v = v;
            M.gK_Pstbar = P(9);
            M.m = st[6];
            M.h = st[7];
            M.ek = -85.0;
            M.dt = dt;
            M.states();
            st[6] = M.m;
            st[7] = M.h;
        }
        if (m & (1u << 6)) {
            Mech_K_Tst M;
            M.v = v;
            M.gK_Tstbar = P(10);
            M.m = st[8];
            M.h = st[9];
            M.ek = -85.0;
            M.dt = dt;
            M.states();
            st[8] = M.m;
            st[9] = M.h;
        }
        if (m & (1u << 7)) {
            Mech_Nap_Et2 M;
            M.v = v;
            M.gNap_Et2bar = P(11);
            M.m = st[10];
            M.h = st[11];
            M.ena = 50.0;
            M.dt = dt;
            M.states();
            st[10] = M.m;
            st[11] = M.h;
        }
        if (m & (1u << 8)) {
            Mech_NaTa_t M;
            M.v = v;
            M.gNaTa_tbar = P(12);
            M.m = st[12];
            M.h = st[13];
            M.ena = 50.0;
            M.dt = dt;
            M.states();
            st[12] = M.m;
            st[13] = M.h;
        }
        if (m & (1u << 9)) {
            Mech_NaTs2_t M;
            M.v = v;
            M.gNaTs2_tbar = P(13);
            M.m = st[14];
            M.h = st[15];
            M.ena = 50.0;
            M.dt = dt;
            M.states();
            st[14] = M.m;
            st[15] = M.h;
        }
        if (m & (1u << 11)) {
            Mech_SK_E2 M;
            M.v = v;
            M.gSK_E2bar = P(16);
            M.zTau = P(17);
            M.z = st[16];
            M.ek = -85.0;
            M.cai = st[CAI];
            M.dt = dt;
            M.states();
            st[16] = M.z;
        }
        if (m & (1u << 12)) {
            Mech_SKv3_1 M;
            M.v = v;
            M.gSKv3_1bar = P(18);
            M.m = st[17];
            M.ek = -85.0;
            M.dt = dt;
            M.states();
            st[17] = M.m;
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
            Mech_Ca_HVA M;
            M.v = v;
            M.gCa_HVAbar = P(0);
            M.m = st[0];
            M.h = st[1];
            M.eca = st[ECA];
            M.breakpoint();
            s += M.ica;
            if (add_ica) ica += M.ica;
        }
        if (m & (1u << 1)) {
            Mech_Ca_LVAst M;
            M.v = v;
            M.gCa_LVAstbar = P(1);
            M.m = st[2];
            M.h = st[3];
            M.eca = st[ECA];
            M.breakpoint();
            s += M.ica;
            if (add_ica) ica += M.ica;
        }
        if (m & (1u << 3)) {
            Mech_Ih M;
            M.v = v;
            M.gIhbar = P(6);
            M.ehcn = P(7);
            M.m = st[4];
            M.breakpoint();
            s += M.ihcn;
        }
        if (m & (1u << 4)) {
            Mech_Im M;
            M.v = v;
            M.gImbar = P(8);
            M.m = st[5];
            M.ek = -85.0;
            M.breakpoint();
            s += M.ik;
        }
        if (m & (1u << 5)) {
            Mech_K_Pst M;
            M.v = v;
            M.gK_Pstbar = P(9);
            M.m = st[6];
            M.h = st[7];
            M.ek = -85.0;
            M.breakpoint();
            s += M.ik;
        }
        if (m & (1u << 6)) {
            Mech_K_Tst M;
            M.v = v;
            M.gK_Tstbar = P(10);
            M.m = st[8];
            M.h = st[9];
            M.ek = -85.0;
            M.breakpoint();
            s += M.ik;
        }
        if (m & (1u << 7)) {
            Mech_Nap_Et2 M;
            M.v = v;
            M.gNap_Et2bar = P(11);
            M.m = st[10];
            M.h = st[11];
            M.ena = 50.0;
            M.breakpoint();
            s += M.ina;
        }
        if (m & (1u << 8)) {
            Mech_NaTa_t M;
            M.v = v;
            M.gNaTa_tbar = P(12);
            M.m = st[12];
            M.h = st[13];
            M.ena = 50.0;
            M.breakpoint();
            s += M.ina;
        }
        if (m & (1u << 9)) {
            Mech_NaTs2_t M;
            M.v = v;
            M.gNaTs2_tbar = P(13);
            M.m = st[14];
            M.h = st[15];
            M.ena = 50.0;
            M.breakpoint();
            s += M.ina;
        }
        if (m & (1u << 10)) {
            Mech_pas M;
            M.v = v;
            M.g = P(14);
            M.e = P(15);
            M.breakpoint();
            s += M.i;
        }
        if (m & (1u << 11)) {
            Mech_SK_E2 M;
            M.v = v;
            M.gSK_E2bar = P(16);
            M.zTau = P(17);
            M.z = st[16];
            M.ek = -85.0;
            M.cai = st[CAI];
            M.breakpoint();
            s += M.ik;
        }
        if (m & (1u << 12)) {
            Mech_SKv3_1 M;
            M.v = v;
            M.gSKv3_1bar = P(18);
            M.m = st[17];
            M.ek = -85.0;
            M.breakpoint();
            s += M.ik;
        }
        sum = s;
    }
};

} // namespace ng_gen_ttpc

using ng_gen_ttpc::SetGen_ttpc;
