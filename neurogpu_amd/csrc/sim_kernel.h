/*
 * sim_kernel.h -- the persistent simulation kernel for gfx950.
 *
 * One run (= parameter set x stimulus) is advanced by a group of W wavefronts
 * (64*W lanes).  Its segments are spread over the lanes: position
 * q = slot*LW + lane, NSLOT slots per lane, with segments sorted by their
 * mechanism set so that the 64 lanes of a wave mostly execute the same
 * mechanism code.  Per step:
 *
 *   1. fused state update + currents: load the mechanism states of the
 *      compartment once, apply the previous step's cnexp update at the current
 *      v, evaluate BREAKPOINT at v+0.001 and v, store the states once;
 *   2. assembly of the Hines rows into LDS (rhs, diagonal, child scatter);
 *   3. triangularisation by tree levels: each lane eliminates one run of
 *      consecutive rows, then each father folds in its related run ends;
 *   4. back-substitution by pointer jumping over the ancestor tables;
 *   5. v += x.
 *
 * Levels are synchronised at wavefront scope when W == 1 (no barrier
 * instruction: a wave's LDS operations complete in issue order, so only the
 * compiler is fenced) and at workgroup scope when W > 1.
 *
 * The algorithm and arithmetic follow the reference kernel:
 *   assembly        Code/OneKernel/src/CudaStuff.cu:263-305
 *   BeforeLU        Code/OneKernel/src/CudaStuff.cu:31-69
 *   BkSub (BKSUB1)  Code/OneKernel/src/CudaStuff.cu:71-103
 *   v update        Code/OneKernel/src/CudaStuff.cu:312-315
 *   record/stimulus Code/OneKernel/src/CudaStuff.cu:241-275
 * Deliberate differences: states are per run (the reference shares them
 * between the stimuli of a block), sibling scatter is ordered (the reference
 * races), duplicate padded level entries are not replayed, and the cnexp of
 * step t-1 is executed at the start of step t on the same loaded state.
 */
#ifndef NGPU_SIM_KERNEL_H
#define NGPU_SIM_KERNEL_H

#include <hip/hip_runtime.h>
#include "mechs.h"

#define NG_EPS_V 0.001

struct SimArgs {
    int N, NQ, LW, runs_per_block;
    int nvar, nrec, nstim, nt, nt_amps, t0, t1, nruns, depth, logn;
    int stim_q;
    double stim_area, v_init;
    const int *q_h, *q_ph, *q_meta;      /* [NQ]: row, parent row, mask | son<<24 (-1 = padding) */
    const double *q_cm;                  /* [NQ] */
    const double *row_e, *row_f;         /* [N] by Hines row */
    const int *pvar;                     /* [NPARAM*NQ] varying-column index or -1 */
    const double *pval;                  /* [NPARAM*NQ] value when uniform */
    const double *pv;                    /* [nsets*nvar] varying parameter values */
    const int *chains, *lvl_ptr;         /* chains: (i0, i1) pairs; lvl_ptr[depth+2] */
    const int *fath, *flvl_ptr, *fchild; /* fath: (J, c0, c1) triples; flvl_ptr[depth+1] */
    const int *fidx;                     /* [logn*N] 0-based ancestor row, N = beyond root */
    const double *dts, *amps;
    const int *rec_q;
    double *states;                      /* [nruns][NSTATE][NQ] */
    double *carry;                       /* [nruns][3*NQ + N + 1] */
    double *vhot;                        /* [nruns][nrec][nt] */
};

template <bool MULTI>
__device__ __forceinline__ void group_sync()
{
    if constexpr (MULTI) {
        __syncthreads();
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

struct ParamAccess {
    const int *pvar;
    const double *pval, *pv;
    int q, NQ;
    __device__ __forceinline__ double operator()(int j) const
    {
        int k = pvar[j * NQ + q];
        return k < 0 ? pval[j * NQ + q] : pv[k];
    }
};

template <class SET, int NSLOT, bool MULTI>
__global__ void __launch_bounds__(MULTI ? 512 : 256) ngpu_sim_kernel(SimArgs a)
{
    extern __shared__ double lds[];
    const int LW = a.LW;
    const int group = threadIdx.x / LW;
    const int g = threadIdx.x - group * LW;
    const int run = blockIdx.x * a.runs_per_block + group;
    if (run >= a.nruns) return;
    const int N = a.N, NQ = a.NQ;
    const int per_run = 2 * (N + 1) + 3 * NQ + a.nrec * 64;
    double *X = lds + (size_t)group * per_run;   /* rhs, then x (PX)          */
    double *U = X + (N + 1);                      /* diagonal, then PF          */
    double *V = U + (N + 1);                      /* v per position             */
    double *DV = V + NQ;                          /* axial dv per position      */
    double *ICA = DV + NQ;                        /* ica of the last BREAKPOINT */
    double *REC = ICA + NQ;
    const int set = run / a.nstim, stim = run - set * a.nstim;
    double *S = a.states + (size_t)run * SET::NSTATE * NQ;
    double *CY = a.carry + (size_t)run * (3 * NQ + N + 1);
    const double *pv = a.pv + (size_t)set * a.nvar;

    if (a.t0 == 0) {
        for (int i = g; i <= N; i += LW) { X[i] = 0.0; U[i] = 0.0; }
#pragma unroll 1
        for (int q = g; q < N; q += LW) {
            const unsigned m = (unsigned)a.q_meta[q] & 0xffffffu;
            ParamAccess P{a.pvar, a.pval, pv, q, NQ};
            V[q] = a.v_init; DV[q] = 0.0; ICA[q] = 0.0;
            double st[SET::NSTATE];
#pragma unroll
            for (int k = 0; k < SET::NSTATE; k++) st[k] = 0.0;
            SET::init(m, a.v_init, st, P);
            const unsigned sb = SET::state_bits(m);
#pragma unroll
            for (int k = 0; k < SET::NSTATE; k++)
                if (sb & (1u << k)) S[k * NQ + q] = st[k];
        }
    } else {
        for (int i = g; i < N; i += LW) X[i] = CY[3 * NQ + i];
        for (int q = g; q < N; q += LW) { V[q] = CY[q]; DV[q] = CY[NQ + q]; ICA[q] = CY[2 * NQ + q]; }
    }
    if (g == 0) { X[N] = 0.0; U[N] = 1.0; }
    group_sync<MULTI>();

    double dtp = a.t0 > 0 ? a.dts[a.t0 - 1] : 0.0;
    const double *amp_row = a.amps + (size_t)stim * a.nt_amps;
    double *out = a.vhot + (size_t)run * a.nrec * a.nt;

    for (int t = a.t0; t < a.t1; t++) {
        const double dt = a.dts[t];
        const double amp = amp_row[t];
        const int slot_t = (t - a.t0) & 63;
        /* record v before the step (CudaStuff.cu:257-261) */
        for (int r = g; r < a.nrec; r += LW) REC[r * 64 + slot_t] = V[a.rec_q[r]];

        /* dv += x[parent] - x[self] with the previous solution (CudaStuff.cu:287) */
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g;
            if (q < N) DV[q] += X[a.q_ph[q]] - X[a.q_h[q]];
        }
        group_sync<MULTI>();

        /* fused cnexp(t-1) + BREAKPOINT(t) + Hines row (CudaStuff.cu:263-294) */
#pragma unroll 1
        for (int q = g; q < N; q += LW) {
            const int h = a.q_h[q];
            const unsigned m = (unsigned)a.q_meta[q] & 0xffffffu;
            const double vq = V[q];
            ParamAccess P{a.pvar, a.pval, pv, q, NQ};
            const unsigned sb = SET::state_bits(m);
            double st[SET::NSTATE];
#pragma unroll
            for (int k = 0; k < SET::NSTATE; k++) st[k] = (sb & (1u << k)) ? S[k * NQ + q] : 0.0;
            if (t > 0) {
                SET::cnexp(m, vq, dtp, st, ICA[q], P);
#pragma unroll
                for (int k = 0; k < SET::NSTATE; k++)
                    if (sb & (1u << k)) S[k * NQ + q] = st[k];
            }
            double sumDv = 0.0, sum = 0.0, icaDv = 0.0, icaNow = 0.0;
            SET::current(m, vq + 0.001, st, P, sumDv, icaDv, false);
            SET::current(m, vq, st, P, sum, icaNow, true);
            ICA[q] = icaNow;
            const double istim = (q == a.stim_q) ? 100 * amp / a.stim_area : 0.0;
            const double gModel = (sumDv - sum) / NG_EPS_V;
            double rhs = istim - sum;
            double D = gModel + a.q_cm[q] / (dt * 1000);
            const double fh = a.row_f[h];
            D -= fh;
            rhs -= fh * DV[q];
            X[h] = rhs;
            U[h] = D;
        }
        group_sync<MULTI>();
        /* children scatter into parents, first children then second children
         * (CudaStuff.cu:296-305; ordered, not racing) */
#pragma unroll
        for (int son = 1; son <= 2; son++) {
#pragma unroll
            for (int c = 0; c < NSLOT; c++) {
                const int q = c * LW + g;
                if (q >= N || (a.q_meta[q] >> 24) != son) continue;
                const int h = a.q_h[q], ph = a.q_ph[q];
                const double e = a.row_e[h];
                X[ph] += e * DV[q];
                U[ph] -= e;
            }
            group_sync<MULTI>();
        }

        /* triangularisation by levels (BeforeLU, CudaStuff.cu:31-69) */
        for (int L = 0; L <= a.depth; L++) {
            const int k1 = a.lvl_ptr[L + 1];
            for (int k = a.lvl_ptr[L] + g; k < k1; k += LW) {
                const int i0 = a.chains[2 * k], i1 = a.chains[2 * k + 1];
                if (i0 >= i1) continue;
                double um1 = U[i0 - 1], bm1 = X[i0 - 1];
                for (int i = i0; i < i1; i++) {
                    const double e1 = a.row_e[i - 1], f1 = a.row_f[i - 1];
                    const double ui = U[i] - f1 * (e1 / um1);
                    const double bi = X[i] - bm1 * e1 / um1;
                    U[i] = ui; X[i] = bi;
                    um1 = ui; bm1 = bi;
                }
            }
            group_sync<MULTI>();
            if (L < a.depth) {
                const int f1e = a.flvl_ptr[L + 1];
                for (int k = a.flvl_ptr[L] + g; k < f1e; k += LW) {
                    const int J = a.fath[3 * k], c0 = a.fath[3 * k + 1], c1 = a.fath[3 * k + 2];
                    double uJ = U[J], bJ = X[J];
                    for (int cc = c0; cc < c1; cc++) {
                        const int tr = a.fchild[cc];
                        const double ut = U[tr], bt = X[tr], e1 = a.row_e[tr], f1 = a.row_f[tr];
                        uJ = uJ - f1 * (e1 / ut);
                        bJ = bJ - bt * e1 / ut;
                    }
                    U[J] = uJ; X[J] = bJ;
                }
                group_sync<MULTI>();
            }
        }

        /* back-substitution by pointer jumping (BKSUB1, CudaStuff.cu:71-103);
         * lanes own rows h = c*LW + g */
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int h = c * LW + g;
            if (h < N) {
                const double u = U[h];
                X[h] = X[h] / u;
                U[h] = -a.row_f[h] / u;
            }
        }
        group_sync<MULTI>();
        for (int r = 0; r < a.logn; r++) {
            double nx[NSLOT], nf[NSLOT];
            const int *fr = a.fidx + (size_t)r * N;
#pragma unroll
            for (int c = 0; c < NSLOT; c++) {
                const int h = c * LW + g;
                nx[c] = 0.0; nf[c] = 0.0;
                if (h < N) {
                    const int an = fr[h];
                    const double pxj = X[h], pfj = U[h];
                    nx[c] = pxj + X[an] * pfj;
                    nf[c] = pfj * U[an];
                }
            }
            group_sync<MULTI>();
#pragma unroll
            for (int c = 0; c < NSLOT; c++) {
                const int h = c * LW + g;
                if (h < N) { X[h] = nx[c]; U[h] = nf[c]; }
            }
            group_sync<MULTI>();
        }

        /* v update (CudaStuff.cu:312-315) */
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g;
            if (q < N) V[q] += X[a.q_h[q]];
        }
        dtp = dt;

        /* flush the recorded chunk */
        if (slot_t == 63 || t == a.t1 - 1) {
            group_sync<MULTI>();
            const int base = t - slot_t;
            for (int i = g; i < a.nrec * 64; i += LW) {
                const int r = i >> 6, j = i & 63;
                if (j <= slot_t) out[(size_t)r * a.nt + base + j] = REC[i];
            }
        }
        group_sync<MULTI>();
    }

    /* carry for the next launch */
    if (a.t1 < a.nt) {
        for (int i = g; i < N; i += LW) CY[3 * NQ + i] = X[i];
        for (int q = g; q < N; q += LW) { CY[q] = V[q]; CY[NQ + q] = DV[q]; CY[2 * NQ + q] = ICA[q]; }
    }
}

#endif
