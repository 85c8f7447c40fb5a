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
 * Per-lane values of the lane's positions (v, dv, ica) live in registers and
 * are only ever indexed with compile-time constants: the slot loop of phase 1
 * is not unrolled (the mechanism code is large), so it works on index 0 and
 * rotates the three arrays by one position per slot.  A run-time index would
 * put the arrays in scratch memory.  Position tables (row, parent row,
 * mechanism mask) are block-shared in LDS.
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
    const double *pval;                  /* [NPARAM*NQ] value, or tagged NaN = varying column */
    const double *pv;                    /* [nsets*nvar] varying parameter values */
    unsigned pconst_mask;                /* bit j: parameter j has one value wherever it is used */
    int pv_in_lds;                       /* stage the run's varying values in LDS */
    const double *pconst;                /* [NPARAM] those values */
    const int *chains, *lvl_ptr;         /* chains: (i0, i1, c0, c1); lvl_ptr[depth+2] */
    const int *fath, *flvl_ptr, *fchild; /* fathers with their own phase: (J, c0, c1); flvl_ptr[depth+1] */
    const int *fidx;                     /* [logn*N] 0-based ancestor row, N = beyond root */
    int nchains, nfath, nfchild;         /* schedule sizes (staged into LDS as int16) */
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

/* Parameter tables hold either the value or, for a run-varying column k, a
 * NaN whose high word is NG_PTAG and low word k (decoded from the run's
 * column values).  Genuine NaN parameters (quiet NaN payloads) pass through. */
#define NG_PTAG 0x7FF40000u

/* Parameter j at position q, read where a mechanism uses it: a value shared
 * by every segment comes from the constant table (a scalar load), otherwise
 * the per-position table gives the value or the tag of a run-varying column. */
struct ParamFetch {
    const double *pconst, *pval, *pv;
    unsigned mask;
    int q, NQ;
    __device__ __forceinline__ double operator()(int j) const
    {
        if (mask & (1u << j)) return pconst[j];
        double x = pval[j * NQ + q];
        const unsigned long long bits = (unsigned long long)__double_as_longlong(x);
        if ((unsigned)(bits >> 32) == NG_PTAG) x = pv[(unsigned)bits];
        return x;
    }
};

template <class SET, int NSLOT, bool MULTI>
#ifndef NG_WAVES_PER_EU
#define NG_WAVES_PER_EU 2
#endif
__global__ void __launch_bounds__(MULTI ? 512 : 256) __attribute__((amdgpu_waves_per_eu(NG_WAVES_PER_EU)))
ngpu_sim_kernel(SimArgs a)
{
    extern __shared__ double lds[];
    const int LW = a.LW;
    const int group = threadIdx.x / LW;
    const int g = threadIdx.x - group * LW;
    const int N = a.N, NQ = a.NQ;
    /* block-shared, staged once: coupling coefficients by Hines row, the
     * position tables, the level schedule and the ancestor tables (int16 row
     * indices; -1 marks a padding position) */
    const int npv = a.pv_in_lds ? a.nvar : 0;
    const int per_run = 2 * (N + 1) + npv + a.nrec * 64;
    double *RE = lds, *RF = lds + N;
    int *QM = (int *)(lds + 2 * N + (size_t)a.runs_per_block * per_run);
    short *QH = (short *)(QM + NQ);
    short *QP = QH + NQ;
    short *FIDX = QP + NQ;
    short *CH = FIDX + a.logn * N;
    short *FA = CH + 4 * a.nchains;
    short *FC = FA + 3 * a.nfath;
    short *LP = FC + a.nfchild;
    short *FLP = LP + (a.depth + 2);
    for (int i = threadIdx.x; i < N; i += blockDim.x) { RE[i] = a.row_e[i]; RF[i] = a.row_f[i]; }
    for (int i = threadIdx.x; i < NQ; i += blockDim.x) {
        const bool in = i < N;
        QM[i] = in ? a.q_meta[i] : 0;
        QH[i] = (short)(in ? a.q_h[i] : -1);
        QP[i] = (short)(in ? a.q_ph[i] : 0);
    }
    for (int i = threadIdx.x; i < a.logn * N; i += blockDim.x) FIDX[i] = (short)a.fidx[i];
    for (int i = threadIdx.x; i < 4 * a.nchains; i += blockDim.x) CH[i] = (short)a.chains[i];
    for (int i = threadIdx.x; i < 3 * a.nfath; i += blockDim.x) FA[i] = (short)a.fath[i];
    for (int i = threadIdx.x; i < a.nfchild; i += blockDim.x) FC[i] = (short)a.fchild[i];
    for (int i = threadIdx.x; i < a.depth + 2; i += blockDim.x) LP[i] = (short)a.lvl_ptr[i];
    for (int i = threadIdx.x; i < a.depth + 1; i += blockDim.x) FLP[i] = (short)a.flvl_ptr[i];
    __syncthreads();
    const int run = blockIdx.x * a.runs_per_block + group;
    if (run >= a.nruns) return;
    double *X = lds + 2 * N + (size_t)group * per_run;   /* rhs, then x (PX)          */
    double *U = X + (N + 1);                              /* diagonal, then PF          */
    double *PV = U + (N + 1);                             /* run-varying parameters     */
    double *REC = PV + npv;
    const int set = run / a.nstim, stim = run - set * a.nstim;
    double *S = a.states + (size_t)run * SET::NSTATE * NQ;
    double *CY = a.carry + (size_t)run * (3 * NQ + N + 1);
    const double *pvg = a.pv + (size_t)set * a.nvar;
    const double *pvsrc = pvg;
    if (a.pv_in_lds) {
        for (int k = g; k < a.nvar; k += LW) PV[k] = pvg[k];
        pvsrc = PV;
    }

    double v[NSLOT], dv[NSLOT], ica[NSLOT];
#pragma unroll
    for (int c = 0; c < NSLOT; c++) { v[c] = (c * LW + g < N) ? a.v_init : 0.0; dv[c] = 0.0; ica[c] = 0.0; }
    if (a.t0 == 0) {
        for (int i = g; i <= N; i += LW) { X[i] = 0.0; U[i] = 0.0; }
#pragma unroll 1
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g;
            if (q >= N) continue;
            const unsigned m = (unsigned)QM[q] & 0xffffffu;
            const ParamFetch P{a.pconst, a.pval, pvsrc, a.pconst_mask, q, NQ};
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
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g;
            if (q < N) { v[c] = CY[q]; dv[c] = CY[NQ + q]; ica[c] = CY[2 * NQ + q]; }
        }
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
        for (int r = 0; r < a.nrec; r++) {
            const int rq = a.rec_q[r];
#pragma unroll
            for (int c = 0; c < NSLOT; c++)
                if (c * LW + g == rq) REC[r * 64 + slot_t] = v[c];
        }

        /* dv += x[parent] - x[self] with the previous solution (CudaStuff.cu:287) */
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g, h = QH[q];
            if (h >= 0) dv[c] += X[QP[q]] - X[h];
        }
        group_sync<MULTI>();

        /* fused cnexp(t-1) + BREAKPOINT(t) + Hines row (CudaStuff.cu:263-294).
         * eca is stored by INITIAL only: at t = 0 it is loaded, afterwards it is
         * recomputed from cai after the cnexp pass. */
        const unsigned keep = (t > 0) ? ~(1u << SET::ECA) : ~0u;
#pragma unroll 1
        for (int c = 0; c < NSLOT; c++) {
            /* slot c's values are at index 0: the arrays rotate by one per slot */
            const int q = c * LW + g;
            const double vq = v[0], dvq = dv[0];
            double icaq = ica[0];
            double st[SET::NSTATE];
            {
                const unsigned sbn = SET::state_bits((unsigned)QM[q] & 0xffffffu) & keep;
#pragma unroll
                for (int k = 0; k < SET::NSTATE; k++) st[k] = (q < N && (sbn & (1u << k))) ? S[k * NQ + q] : 0.0;
            }
            if (q < N) {
                const int h = QH[q];
                const unsigned m = (unsigned)QM[q] & 0xffffffu;
                const double cmq = a.q_cm[q];
                const double fh = RF[h];
                const unsigned sb = SET::state_bits(m) & keep;
                const ParamFetch P{a.pconst, a.pval, pvsrc, a.pconst_mask, q, NQ};
                if (t > 0) {
                    SET::cnexp(m, vq, dtp, st, icaq, P);
                    SET::eca_after_cnexp(m, st);
#pragma unroll
                    for (int k = 0; k < SET::NSTATE; k++)
                        if (sb & (1u << k)) S[k * NQ + q] = st[k];
                }
                double sumDv = 0.0, sum = 0.0, icaDv = 0.0, icaNow = 0.0;
                SET::current(m, vq + 0.001, st, P, sumDv, icaDv, false);
                SET::current(m, vq, st, P, sum, icaNow, true);
                icaq = icaNow;
                const double istim = (q == a.stim_q) ? 100 * amp / a.stim_area : 0.0;
                const double gModel = (sumDv - sum) / NG_EPS_V;
                double rhs = istim - sum;
                double D = gModel + cmq / (dt * 1000);
                D -= fh;
                rhs -= fh * dvq;
                X[h] = rhs;
                U[h] = D;
            }
#pragma unroll
            for (int i = 0; i + 1 < NSLOT; i++) { v[i] = v[i + 1]; dv[i] = dv[i + 1]; ica[i] = ica[i + 1]; }
            v[NSLOT - 1] = vq; dv[NSLOT - 1] = dvq; ica[NSLOT - 1] = icaq;
        }
        group_sync<MULTI>();
        /* children scatter into parents, first children then second children
         * (CudaStuff.cu:296-305; ordered, not racing) */
#pragma unroll
        for (int son = 1; son <= 2; son++) {
#pragma unroll
            for (int c = 0; c < NSLOT; c++) {
                const int q = c * LW + g, h = QH[q];
                if (h < 0 || (QM[q] >> 24) != son) continue;
                const double e = RE[h];
                X[QP[q]] += e * dv[c];
                U[QP[q]] -= e;
            }
            group_sync<MULTI>();
        }

        /* triangularisation by levels (BeforeLU, CudaStuff.cu:31-69).  Level L
         * first folds the fathers of level L-1 that have their own phase, then
         * eliminates its chains; a chain (i0, i1, c0, c1) with c0 < c1 first
         * folds the father row i0-1 over the related rows FC[c0..c1). */
        for (int L = 0; L <= a.depth; L++) {
            if (L > 0 && FLP[L] > FLP[L - 1]) {
                const int f1e = FLP[L];
                for (int k = FLP[L - 1] + g; k < f1e; k += LW) {
                    const int J = FA[3 * k], c0 = FA[3 * k + 1], c1 = FA[3 * k + 2];
                    double uJ = U[J], bJ = X[J];
                    for (int cc = c0; cc < c1; cc++) {
                        const int tr = FC[cc];
                        const double ut = U[tr], bt = X[tr], e1 = RE[tr], f1 = RF[tr];
                        uJ = uJ - f1 * (e1 / ut);
                        bJ = bJ - bt * e1 / ut;
                    }
                    U[J] = uJ; X[J] = bJ;
                }
                group_sync<MULTI>();
            }
            const int k1 = LP[L + 1];
            for (int k = LP[L] + g; k < k1; k += LW) {
                const int i0 = CH[4 * k], i1 = CH[4 * k + 1], c0 = CH[4 * k + 2], c1 = CH[4 * k + 3];
                double um1 = U[i0 - 1], bm1 = X[i0 - 1];
                if (c0 < c1) {
                    for (int cc = c0; cc < c1; cc++) {
                        const int tr = FC[cc];
                        const double ut = U[tr], bt = X[tr], e1 = RE[tr], f1 = RF[tr];
                        um1 = um1 - f1 * (e1 / ut);
                        bm1 = bm1 - bt * e1 / ut;
                    }
                    U[i0 - 1] = um1; X[i0 - 1] = bm1;
                }
                if (i0 >= i1) continue;
                double e1 = RE[i0 - 1], f1 = RF[i0 - 1], ui = U[i0], xi = X[i0];
                for (int i = i0; i < i1; i++) {
                    double e1n = 0.0, f1n = 0.0, un = 0.0, xn = 0.0;
                    if (i + 1 < i1) { e1n = RE[i]; f1n = RF[i]; un = U[i + 1]; xn = X[i + 1]; }
                    const double unew = ui - f1 * (e1 / um1);
                    const double bnew = xi - bm1 * e1 / um1;
                    U[i] = unew; X[i] = bnew;
                    um1 = unew; bm1 = bnew;
                    e1 = e1n; f1 = f1n; ui = un; xi = xn;
                }
            }
            group_sync<MULTI>();
        }

        /* back-substitution by pointer jumping (BKSUB1, CudaStuff.cu:71-103);
         * lanes own rows h = c*LW + g */
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int h = c * LW + g;
            if (h < N) {
                const double u = U[h];
                X[h] = X[h] / u;
                U[h] = -RF[h] / u;
            }
        }
        group_sync<MULTI>();
        for (int r = 0; r < a.logn; r++) {
            double nx[NSLOT], nf[NSLOT];
            int an[NSLOT];
            const short *fr = FIDX + r * N;
#pragma unroll
            for (int c = 0; c < NSLOT; c++) {
                const int h = c * LW + g;
                an[c] = h < N ? fr[h] : N;
            }
#pragma unroll
            for (int c = 0; c < NSLOT; c++) {
                const int h = c * LW + g;
                nx[c] = 0.0; nf[c] = 0.0;
                if (h < N) {
                    const double pxj = X[h], pfj = U[h];
                    nx[c] = pxj + X[an[c]] * pfj;
                    nf[c] = pfj * U[an[c]];
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
        for (int c = 0; c < NSLOT; c++)
            { const int h = QH[c * LW + g]; if (h >= 0) v[c] += X[h]; }
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
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g;
            if (q < N) { CY[q] = v[c]; CY[NQ + q] = dv[c]; CY[2 * NQ + q] = ica[c]; }
        }
    }
}

#endif
