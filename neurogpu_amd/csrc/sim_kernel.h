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
 * Per-lane values of the lane's positions (v, dv) live in registers and are
 * only ever indexed with compile-time constants: the slot loop of phase 1 is
 * not unrolled (the mechanism code is large), so it works on index 0 and
 * rotates the two arrays by one position per slot.  The calcium current that
 * one step's BREAKPOINT hands to the next step's DERIVATIVE pass is a state
 * slot (SET::ICA) of the compartments whose mechanisms read it.  A run-time index would
 * put the arrays in scratch memory.  Position tables (row, parent row,
 * mechanism mask) are block-shared in LDS.
 *
 * Levels are synchronised at wavefront scope when W == 1 (no barrier
 * instruction: a wave's LDS operations complete in issue order, so only the
 * compiler is fenced) and at workgroup scope when W > 1.
 *
 * Two kernels share the per-run code (sim_run):
 *   ngpu_sim_kernel   one morphology; workgroup b runs R = runs_per_block runs;
 *   ngpu_batch_kernel persistent: each workgroup of 8 waves takes items from a
 *                     device work queue (an atomic counter over an item list);
 *                     an item is up to 8/W runs of one morphology, so runs of
 *                     uneven size and step count share one launch.
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
    const short *fidx16;                 /* [logn*N] 0-based ancestor row, N = beyond root */
    int fidx_in_lds;                     /* stage the ancestor table in LDS */
    int block_sync;                      /* batch kernel: synchronise levels at workgroup scope */
    int nslot;                           /* batch kernel: positions per lane of this layout (NQ / LW) */
    int nchains, nfath, nfchild;         /* schedule sizes (staged into LDS as int16) */
    const double *dts, *amps;
    const int *rec_q;
    double *states;                      /* [nruns][NSTATE][NQ] */
    double *carry;                       /* [nruns][2*NQ + N + 1]: v, dv, x between launches */
    double *vhot;                        /* [nruns][nrec][nt] */
};

/* Level synchronisation of one run's waves. */
__device__ __forceinline__ void ng_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
struct SyncWave  { __device__ __forceinline__ void operator()() const { ng_wave_sync(); } };
struct SyncBlock { __device__ __forceinline__ void operator()() const { __syncthreads(); } };
struct SyncDyn {
    bool block;
    __device__ __forceinline__ void operator()() const { if (block) __syncthreads(); else ng_wave_sync(); }
};

/* LDS layout of a workgroup, in this order:
 *   RE[N], RF[N]                                   doubles (by Hines row)
 *   R per-run regions: X[N+1], U[N+1], PV[npv], REC[64*nrec]   doubles
 *   QM[NQS] int, QH[NQS] short, QP[NQS] short      NQS = NSLOT*LW positions
 *   FIDX[logn*N] short                             only when staged
 *   CH[4*nchains], FA[3*nfath], FC[nfchild], LP[depth+2], FLP[depth+1] short */
__host__ __device__ inline size_t ng_lds_per_run(int N, int npv, int nrec)
{
    return (size_t)(2 * (N + 1) + npv + 64 * nrec);
}
__host__ __device__ inline size_t ng_lds_bytes(int N, int NQS, int R, int npv, int nrec, int logn, int fidx_in_lds,
                                               int nchains, int nfath, int nfchild, int depth)
{
    const size_t nshort = 2 * (size_t)NQS + (fidx_in_lds ? (size_t)logn * N : 0) + 4 * (size_t)nchains +
                          3 * (size_t)nfath + (size_t)nfchild + (size_t)(depth + 2) + (size_t)(depth + 1);
    const size_t dbl = 2 * (size_t)N + (size_t)R * ng_lds_per_run(N, npv, nrec);
    return dbl * sizeof(double) + (((size_t)NQS * sizeof(int) + nshort * sizeof(short) + 7) / 8) * 8;
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

/* Pointers into the block-shared part of the LDS layout. */
struct Tables {
    double *RE, *RF;
    int *QM;
    short *QH, *QP;
    short *FIDX;                         /* NULL: read a.fidx16 from global memory */
    short *CH, *FA, *FC, *LP, *FLP;
};

/* Stage the block-shared tables of morphology a for R runs and NQS = NSLOT*LW
 * positions; positions beyond the morphology's own table are padding. */
__device__ __forceinline__ Tables stage_tables(const SimArgs &a, double *lds, int R, int NQS, int npv)
{
    const int N = a.N;
    Tables T;
    T.RE = lds;
    T.RF = lds + N;
    T.QM = (int *)(lds + 2 * N + (size_t)R * ng_lds_per_run(N, npv, a.nrec));
    T.QH = (short *)(T.QM + NQS);
    T.QP = T.QH + NQS;
    short *next = T.QP + NQS;
    T.FIDX = a.fidx_in_lds ? next : nullptr;
    T.CH = next + (a.fidx_in_lds ? a.logn * N : 0);
    T.FA = T.CH + 4 * a.nchains;
    T.FC = T.FA + 3 * a.nfath;
    T.LP = T.FC + a.nfchild;
    T.FLP = T.LP + (a.depth + 2);
    for (int i = threadIdx.x; i < N; i += blockDim.x) { T.RE[i] = a.row_e[i]; T.RF[i] = a.row_f[i]; }
    for (int i = threadIdx.x; i < NQS; i += blockDim.x) {
        const bool in = i < N;
        T.QM[i] = in ? a.q_meta[i] : 0;
        T.QH[i] = (short)(in ? a.q_h[i] : -1);
        T.QP[i] = (short)(in ? a.q_ph[i] : 0);
    }
    if (a.fidx_in_lds)
        for (int i = threadIdx.x; i < a.logn * N; i += blockDim.x) T.FIDX[i] = a.fidx16[i];
    for (int i = threadIdx.x; i < 4 * a.nchains; i += blockDim.x) T.CH[i] = (short)a.chains[i];
    for (int i = threadIdx.x; i < 3 * a.nfath; i += blockDim.x) T.FA[i] = (short)a.fath[i];
    for (int i = threadIdx.x; i < a.nfchild; i += blockDim.x) T.FC[i] = (short)a.fchild[i];
    for (int i = threadIdx.x; i < a.depth + 2; i += blockDim.x) T.LP[i] = (short)a.lvl_ptr[i];
    for (int i = threadIdx.x; i < a.depth + 1; i += blockDim.x) T.FLP[i] = (short)a.flvl_ptr[i];
    return T;
}

/* Steps t0..t1 of one run by the LW lanes of its group (lane g).  X is the
 * run's LDS region.  Runs from nruns up to the next multiple of 8 are padding
 * (buffers are allocated for them, their parameters are the last set's): the
 * batch kernel gives them to the unused groups of a partial item. */
template <class SET, int NSLOT, class SYNC>
__device__ __forceinline__ void sim_run(const SimArgs &a, const Tables &T, double *X, int run, int g,
                                        const SYNC &sync)
{
    const int LW = a.LW, N = a.N, NQ = a.NQ;
    const int npv = a.pv_in_lds ? a.nvar : 0;
    double *RE = T.RE, *RF = T.RF;
    int *QM = T.QM;
    short *QH = T.QH, *QP = T.QP, *CH = T.CH, *FA = T.FA, *FC = T.FC, *LP = T.LP, *FLP = T.FLP;
    double *U = X + (N + 1);                              /* diagonal, then PF          */
    double *PV = U + (N + 1);                             /* run-varying parameters     */
    double *REC = PV + npv;
    const int set0 = run / a.nstim, stim = run - set0 * a.nstim;
    const int set = min(set0, a.nruns / a.nstim - 1);
    double *S = a.states + (size_t)run * SET::NSTATE * NQ;
    double *CY = a.carry + (size_t)run * (2 * NQ + N + 1);
    const double *pvg = a.pv + (size_t)set * a.nvar;
    const double *pvsrc = pvg;
    if (a.pv_in_lds) {
        for (int k = g; k < a.nvar; k += LW) PV[k] = pvg[k];
        pvsrc = PV;
    }

    double v[NSLOT], dv[NSLOT];
#pragma unroll
    for (int c = 0; c < NSLOT; c++) { v[c] = (c * LW + g < N) ? a.v_init : 0.0; dv[c] = 0.0; }
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
        for (int i = g; i < N; i += LW) X[i] = CY[2 * NQ + i];
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g;
            if (q < N) { v[c] = CY[q]; dv[c] = CY[NQ + q]; }
        }
    }
    if (g == 0) { X[N] = 0.0; U[N] = 1.0; }
    sync();

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
        sync();

        /* fused cnexp(t-1) + BREAKPOINT(t) + Hines row (CudaStuff.cu:263-294).
         * eca is stored by INITIAL only: at t = 0 it is loaded, afterwards it is
         * recomputed from cai after the cnexp pass. */
        const unsigned keep = (t > 0) ? ~(1u << SET::ECA) : ~0u;
#pragma unroll 1
        for (int c = 0; c < NSLOT; c++) {
            /* slot c's values are at index 0: the arrays rotate by one per slot */
            const int q = c * LW + g;
            const double vq = v[0], dvq = dv[0];
            double st[SET::NSTATE];
            {
                const unsigned mq = (unsigned)QM[q] & 0xffffffu;
                const unsigned sbn = (SET::state_bits(mq) & keep) | (t > 0 ? SET::ica_bits(mq) : 0u);
#pragma unroll
                for (int k = 0; k < SET::NSTATE; k++) st[k] = (q < N && (sbn & (1u << k))) ? S[k * NQ + q] : 0.0;
            }
            if (q < N) {
                const int h = QH[q];
                const unsigned m = (unsigned)QM[q] & 0xffffffu;
                const double cmq = a.q_cm[q];
                const double fh = RF[h];
                const unsigned sb = SET::state_bits(m) & keep;
                const unsigned ib = SET::ica_bits(m);
                const ParamFetch P{a.pconst, a.pval, pvsrc, a.pconst_mask, q, NQ};
                if (t > 0) {
                    SET::cnexp(m, vq, dtp, st, ib ? st[SET::ICA] : 0.0, P);
                    SET::eca_after_cnexp(m, st);
#pragma unroll
                    for (int k = 0; k < SET::NSTATE; k++)
                        if (sb & (1u << k)) S[k * NQ + q] = st[k];
                }
                double sumDv = 0.0, sum = 0.0, icaDv = 0.0, icaNow = 0.0;
                SET::current(m, vq + 0.001, st, P, sumDv, icaDv, false);
                SET::current(m, vq, st, P, sum, icaNow, true);
                if (ib) S[SET::ICA * NQ + q] = icaNow;
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
            for (int i = 0; i + 1 < NSLOT; i++) { v[i] = v[i + 1]; dv[i] = dv[i + 1]; }
            v[NSLOT - 1] = vq; dv[NSLOT - 1] = dvq;
        }
        sync();
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
            sync();
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
                sync();
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
            sync();
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
        sync();
        for (int r = 0; r < a.logn; r++) {
            double nx[NSLOT], nf[NSLOT];
            int an[NSLOT];
            if (T.FIDX) {
                const short *fr = T.FIDX + r * N;
#pragma unroll
                for (int c = 0; c < NSLOT; c++) {
                    const int h = c * LW + g;
                    an[c] = h < N ? fr[h] : N;
                }
            } else {
                const short *fr = a.fidx16 + r * N;
#pragma unroll
                for (int c = 0; c < NSLOT; c++) {
                    const int h = c * LW + g;
                    an[c] = h < N ? fr[h] : N;
                }
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
            sync();
#pragma unroll
            for (int c = 0; c < NSLOT; c++) {
                const int h = c * LW + g;
                if (h < N) { X[h] = nx[c]; U[h] = nf[c]; }
            }
            sync();
        }

        /* v update (CudaStuff.cu:312-315) */
#pragma unroll
        for (int c = 0; c < NSLOT; c++)
            { const int h = QH[c * LW + g]; if (h >= 0) v[c] += X[h]; }
        dtp = dt;

        /* flush the recorded chunk */
        if (slot_t == 63 || t == a.t1 - 1) {
            sync();
            const int base = t - slot_t;
            for (int i = g; i < a.nrec * 64; i += LW) {
                const int r = i >> 6, j = i & 63;
                if (j <= slot_t) out[(size_t)r * a.nt + base + j] = REC[i];
            }
        }
        sync();
    }

    /* carry for the next launch */
    if (a.t1 < a.nt) {
        for (int i = g; i < N; i += LW) CY[2 * NQ + i] = X[i];
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g;
            if (q < N) { CY[q] = v[c]; CY[NQ + q] = dv[c]; }
        }
    }
}

#ifndef NG_WAVES_PER_EU
#define NG_WAVES_PER_EU 2
#endif

template <class SET, int NSLOT, bool MULTI>
__global__ void __launch_bounds__(MULTI ? 512 : 256) __attribute__((amdgpu_waves_per_eu(NG_WAVES_PER_EU)))
ngpu_sim_kernel(SimArgs a)
{
    extern __shared__ double lds[];
    const int LW = a.LW;
    const int group = threadIdx.x / LW;
    const int g = threadIdx.x - group * LW;
    const int npv = a.pv_in_lds ? a.nvar : 0;
    const Tables T = stage_tables(a, lds, a.runs_per_block, NSLOT * LW, npv);
    __syncthreads();
    const int run = blockIdx.x * a.runs_per_block + group;
    if (run >= a.nruns) return;          /* W > 1 implies one run per workgroup */
    double *X = lds + 2 * a.N + (size_t)group * ng_lds_per_run(a.N, npv, a.nrec);
    if constexpr (MULTI) sim_run<SET, NSLOT>(a, T, X, run, g, SyncBlock{});
    else sim_run<SET, NSLOT>(a, T, X, run, g, SyncWave{});
}

/* Work queue of the batch kernel: item i runs runs [y, y + z) of morphology x.
 * The morphologies' arguments travel in the kernel arguments, so that their
 * pointers are known to address global memory. */
#define NG_BATCH_MAX_MORPH 8
struct BatchArgs {
    SimArgs morph[NG_BATCH_MAX_MORPH];
    const int4 *items;
    int nitems;
    int *head;                           /* next item; zero before the launch */
    unsigned long long *stamps;          /* [nitems][2] start and end, device clock */
};

#define NG_BATCH_THREADS 512
#define NG_BATCH_NSLOT 6                 /* positions per lane: up to 64*8*6 = 3072 segments */

/* NSLOT positions per lane at most; layouts of up to NSLOT_SMALL positions
 * per lane run a copy of the step code specialised for that count. */
template <class SET, int NSLOT, int NSLOT_SMALL>
__global__ void __launch_bounds__(NG_BATCH_THREADS) __attribute__((amdgpu_waves_per_eu(NG_WAVES_PER_EU)))
ngpu_batch_kernel(BatchArgs b)
{
    extern __shared__ double lds[];
    __shared__ int s_item;
    for (;;) {
        __syncthreads();                 /* the previous item is done with LDS */
        if (threadIdx.x == 0) s_item = atomicAdd(b.head, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= b.nitems) break;
        const int4 item = b.items[it];
        const int mi = __builtin_amdgcn_readfirstlane(item.x);
        const int run0 = __builtin_amdgcn_readfirstlane(item.y);
        const int nrun = __builtin_amdgcn_readfirstlane(item.z);
        const SimArgs &a = b.morph[mi];
        if (threadIdx.x == 0) b.stamps[2 * it] = wall_clock64();
        const int LW = a.LW, R = NG_BATCH_THREADS / LW;
        const int group = threadIdx.x / LW;
        const int g = threadIdx.x - group * LW;
        const int npv = a.pv_in_lds ? a.nvar : 0;
        const Tables T = stage_tables(a, lds, R, NSLOT * LW, npv);
        __syncthreads();
        (void)nrun;                      /* groups beyond nrun simulate padding runs */
        double *X = lds + 2 * a.N + (size_t)group * ng_lds_per_run(a.N, npv, a.nrec);
        if (a.nslot <= NSLOT_SMALL) sim_run<SET, NSLOT_SMALL>(a, T, X, run0 + group, g, SyncDyn{a.block_sync != 0});
        else sim_run<SET, NSLOT>(a, T, X, run0 + group, g, SyncDyn{a.block_sync != 0});
        __syncthreads();
        if (threadIdx.x == 0) b.stamps[2 * it + 1] = wall_clock64();
    }
}

#endif
