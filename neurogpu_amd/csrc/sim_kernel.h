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
 *   3. triangularisation by tree levels: each lane eliminates one chain of
 *      consecutive rows (a section), folding the ends of finished child
 *      chains into the chain's first row, with the reference's arithmetic;
 *      then every row's pivot quotients PX = b/u and PF = -f/u, in parallel;
 *   4. back-substitution by the same chains in reverse level order,
 *      x_i = PX_i + PF_i * x_parent(i), the parent of a chain's end row
 *      coming from a chain of a higher level;
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
 * The algorithm follows the reference kernel:
 *   assembly        Code/OneKernel/src/CudaStuff.cu:263-305
 *   BeforeLU        Code/OneKernel/src/CudaStuff.cu:31-69
 *   v update        Code/OneKernel/src/CudaStuff.cu:312-315
 *   record/stimulus Code/OneKernel/src/CudaStuff.cu:241-275
 * The elimination visits the rows in the reference's level order; the
 * back-substitution solves the same triangular system as its pointer-jumping
 * BKSUB1 (CudaStuff.cu:71-103) row by row instead of in log2(depth) rounds
 * over the whole tree.  Deliberate differences: states are per run (the
 * reference shares them between the stimuli of a block), sibling scatter is
 * ordered (the reference races), duplicate padded level entries are not
 * replayed, the cnexp of step t-1 is executed at the start of step t on the
 * same loaded state.  The elimination keeps the reference's operation order
 * (u_i -= f (e / u_{i-1}), b_i -= b_{i-1} e / u_{i-1}): the reference trees
 * couple zero-area nodes with coefficients near 1e22, and there the order of
 * these operations, not the rounding of the membrane terms, sets the result.
 */
#ifndef NGPU_SIM_KERNEL_H
#define NGPU_SIM_KERNEL_H

#include <hip/hip_runtime.h>
#include "mechs.h"

#define NG_EPS_V 0.001

struct SimArgs {
    int N, NQ, LW, runs_per_block;
    int nvar, nrec, nstim, nt, nt_amps, t0, t1, nruns, depth;
    int stim_q;
    double stim_area, v_init;
    const int *q_h, *q_ph, *q_meta;      /* [NQ]: row, parent row, mask | son<<24 (-1 = padding) */
    const double *q_cm;                  /* [NQ] */
    const double *row_ef;                /* [2N]: e, f interleaved by Hines row */
    const double *pval;                  /* [NPARAM*NQ] value, or tagged NaN = varying column */
    const double *pv;                    /* [nsets*nvar] varying parameter values */
    unsigned pconst_mask;                /* bit j: parameter j has one value wherever it is used */
    const double *pconst;                /* [NPARAM] those values */
    const int *chains, *lvl_ptr;         /* chains: (i0, i1, tA, tB, pe); lvl_ptr[depth+2] */
    const int *fath, *flvl_ptr, *fchild; /* fathers with their own phase: (J, c0, c1); flvl_ptr[depth+1] */
    int lds_staging;                     /* row coefficients e, f and the run's varying
                                            parameter values staged in LDS (else read
                                            from global memory); selects the kernel's
                                            STAGE instantiation */
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
 *   EF[2N]                                         e, f interleaved by Hines row, when staged
 *   R per-run regions: XU[2(N+1)], PV[npv], REC[64*nrec]   doubles; XU holds
 *                      (b, u), then (PX, PF), then (x, PF) interleaved by row, so
 *                      that a row is one 16-byte access; npv = nvar rounded up to
 *                      even when staged, else 0 (regions stay 16-byte aligned)
 *   QM[NQS] int, QH[NQS] short, QP[NQS] short      NQS = NSLOT*LW positions
 *   CH[5*nchains], FA[3*nfath], FC[nfchild], LP[depth+2], FLP[depth+1] short */
__host__ __device__ inline size_t ng_lds_per_run(int N, int npv, int nrec)
{
    return (size_t)(2 * (N + 1) + ((npv + 1) & ~1) + 64 * nrec);
}
__host__ __device__ inline size_t ng_lds_bytes(int N, int NQS, int R, int npv, int nrec, int staged,
                                               int nchains, int nfath, int nfchild, int depth)
{
    const size_t nshort = 2 * (size_t)NQS + 5 * (size_t)nchains + 3 * (size_t)nfath + (size_t)nfchild +
                          (size_t)(depth + 2) + (size_t)(depth + 1);
    const size_t dbl = (staged ? 2 * (size_t)N : 0) + (size_t)R * ng_lds_per_run(N, staged ? npv : 0, nrec);
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

/* Pointers into the block-shared part of the LDS layout; EF points to LDS or
 * to global memory depending on the STAGE instantiation, so that every access
 * compiles to an LDS or a global load (never a flat one). */
struct Tables {
    const double2 *EF;                   /* (e, f) by Hines row */
    double *runs;                        /* first per-run region */
    int *QM;
    short *QH, *QP;
    short *CH, *FA, *FC, *LP, *FLP;
};

/* Stage the block-shared tables of morphology a for R runs and NQS = NSLOT*LW
 * positions; positions beyond the morphology's own table are padding. */
template <bool STAGE>
__device__ __forceinline__ Tables stage_tables(const SimArgs &a, double *lds, int R, int NQS)
{
    const int N = a.N;
    const int npv = STAGE ? a.nvar : 0;
    Tables T;
    if constexpr (STAGE) {
        for (int i = threadIdx.x; i < 2 * N; i += blockDim.x) lds[i] = a.row_ef[i];
        T.EF = (const double2 *)lds;
        T.runs = lds + 2 * N;
    } else {
        T.EF = (const double2 *)a.row_ef;
        T.runs = lds;
    }
    T.QM = (int *)(T.runs + (size_t)R * ng_lds_per_run(N, npv, a.nrec));
    T.QH = (short *)(T.QM + NQS);
    T.QP = T.QH + NQS;
    T.CH = T.QP + NQS;
    T.FA = T.CH + 5 * a.nchains;
    T.FC = T.FA + 3 * a.nfath;
    T.LP = T.FC + a.nfchild;
    T.FLP = T.LP + (a.depth + 2);
    for (int i = threadIdx.x; i < NQS; i += blockDim.x) {
        const bool in = i < N;
        T.QM[i] = in ? a.q_meta[i] : 0;
        T.QH[i] = (short)(in ? a.q_h[i] : -1);
        T.QP[i] = (short)(in ? a.q_ph[i] : 0);
    }
    for (int i = threadIdx.x; i < 5 * a.nchains; i += blockDim.x) T.CH[i] = (short)a.chains[i];
    for (int i = threadIdx.x; i < 3 * a.nfath; i += blockDim.x) T.FA[i] = (short)a.fath[i];
    for (int i = threadIdx.x; i < a.nfchild; i += blockDim.x) T.FC[i] = (short)a.fchild[i];
    for (int i = threadIdx.x; i < a.depth + 2; i += blockDim.x) T.LP[i] = (short)a.lvl_ptr[i];
    for (int i = threadIdx.x; i < a.depth + 1; i += blockDim.x) T.FLP[i] = (short)a.flvl_ptr[i];
    return T;
}

/* n1/d and n2/d, correctly rounded, sharing one reciprocal of d.  When d has
 * a binary exponent in [-300, 300) and n1, n2 have exponents in [-400, 400),
 * the IEEE division on gfx9 (v_div_scale, rcp, two Newton steps, product,
 * residual correction, v_div_fmas, v_div_fixup) neither scales nor fixes up,
 * and its remaining steps are exactly the ones below, so the quotients are
 * bit for bit those of n1/d and n2/d.  The test is a vote of the wave, so the
 * common case is straight-line code; a wave with any other operand (zero,
 * huge, tiny, non-finite) takes the divisions themselves.  The reciprocal
 * refinement is done once. */
__device__ __forceinline__ int ng_exp_in(double x, unsigned lo, unsigned span)
{
    const unsigned ex = ((unsigned)__double2hiint(x) >> 20) & 0x7ffu;
    return ex - lo < span;
}
__device__ __forceinline__ void ng_div2(double n1, double n2, double d, double &q1, double &q2)
{
    const int plain = ng_exp_in(d, 1023u - 300u, 600u) & ng_exp_in(n1, 1023u - 400u, 800u) &
                      ng_exp_in(n2, 1023u - 400u, 800u);
    if (__all(plain)) {
        double r = __builtin_amdgcn_rcp(d);
        double e = fma(-d, r, 1.0);
        r = fma(r, e, r);
        e = fma(-d, r, 1.0);
        r = fma(r, e, r);
        double q = n1 * r;
        q1 = fma(fma(-d, q, n1), r, q);
        q = n2 * r;
        q2 = fma(fma(-d, q, n2), r, q);
    } else {
        q1 = n1 / d;
        q2 = n2 / d;
    }
}

/* Eliminate the final row (ut, bt) with coefficients e, f into the running
 * father values (uJ, bJ) with the reference's arithmetic
 * (CudaStuff.cu:53-64): uJ -= f (e / ut), bJ -= (bt e) / ut. */
__device__ __forceinline__ void ng_fold(double ut, double bt, double e, double f, double &uJ, double &bJ)
{
    double q1, q2;
    ng_div2(e, bt * e, ut, q1, q2);
    uJ = uJ - f * q1;
    bJ = bJ - q2;
}

/* Steps t0..t1 of one run by the LW lanes of its group (lane g).  XU is the
 * run's LDS region.  Runs from nruns up to the next multiple of 8 are padding
 * (buffers are allocated for them, their parameters are the last set's): the
 * batch kernel gives them to the unused groups of a partial item. */
template <class SET, int NSLOT, bool STAGE, class SYNC>
__device__ __forceinline__ void sim_run(const SimArgs &a, const Tables &T, double *XU, int run, int g,
                                        const SYNC &sync)
{
    const int LW = a.LW, N = a.N, NQ = a.NQ;
    const int npv = STAGE ? ((a.nvar + 1) & ~1) : 0;
    const double2 *EF = T.EF;
    int *QM = T.QM;
    short *QH = T.QH, *QP = T.QP, *CH = T.CH, *FA = T.FA, *FC = T.FC, *LP = T.LP, *FLP = T.FLP;
    double2 *XU2 = (double2 *)XU;                         /* (b, u) by Hines row        */
    double *PV = XU + 2 * (N + 1);                        /* run-varying parameters     */
    double *REC = PV + npv;
    const int set0 = run / a.nstim, stim = run - set0 * a.nstim;
    const int set = min(set0, a.nruns / a.nstim - 1);
    double *S = a.states + (size_t)run * SET::NSTATE * NQ;
    double *CY = a.carry + (size_t)run * (2 * NQ + N + 1);
    const double *pvg = a.pv + (size_t)set * a.nvar;
    const double *pvsrc = pvg;
    if constexpr (STAGE) {
        for (int k = g; k < a.nvar; k += LW) PV[k] = pvg[k];
        pvsrc = PV;
    }

    double v[NSLOT], dv[NSLOT];
#pragma unroll
    for (int c = 0; c < NSLOT; c++) { v[c] = (c * LW + g < N) ? a.v_init : 0.0; dv[c] = 0.0; }
    if (a.t0 == 0) {
        for (int i = g; i < 2 * (N + 1); i += LW) XU[i] = 0.0;
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
        for (int i = g; i < N; i += LW) XU[2 * i] = CY[2 * NQ + i];
#pragma unroll
        for (int c = 0; c < NSLOT; c++) {
            const int q = c * LW + g;
            if (q < N) { v[c] = CY[q]; dv[c] = CY[NQ + q]; }
        }
    }
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
            if (h >= 0) dv[c] += XU[2 * QP[q]] - XU[2 * h];
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
            /* the mechanism code is specialised for the set's common mechanism
             * masks (straight-line code, independent gates interleaved) */
            SET::dispatch((unsigned)QM[q] & 0xffffffu, [&](auto mask) {
                const unsigned m = mask;
                double st[SET::NSTATE];
                const unsigned sbn = (SET::state_bits(m) & keep) | (t > 0 ? SET::ica_bits(m) : 0u);
#pragma unroll
                for (int k = 0; k < SET::NSTATE; k++) st[k] = (q < N && (sbn & (1u << k))) ? S[k * NQ + q] : 0.0;
                if (q < N) {
                    const int h = QH[q];
                    const double cmq = a.q_cm[q];
                    const double fh = EF[h].y;
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
                    XU2[h] = make_double2(rhs, D);
                }
            });
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
                const double e = EF[h].x;
                double2 p = XU2[QP[q]];
                p.x += e * dv[c];
                p.y -= e;
                XU2[QP[q]] = p;
            }
            sync();
        }

        /* triangularisation by levels (BeforeLU, CudaStuff.cu:31-69), with the
         * reference's arithmetic.  Level L first folds the fathers of level
         * L-1 that have their own phase, then eliminates its chains; a chain
         * (i0, i1, tA, tB, pe) folds up to two related rows tA, tB (-1 = none)
         * into its first row i0-1, then eliminates rows i0-1 .. i1-2 each into
         * the next.  Every load that does not depend on the chain's own
         * results is issued up front, and the rows' values are loaded two
         * rows ahead (indices clamped to the chain, so unconditionally). */
        for (int L = 0; L <= a.depth; L++) {
            if (L > 0 && FLP[L] > FLP[L - 1]) {
                const int f1e = FLP[L];
                for (int k = FLP[L - 1] + g; k < f1e; k += LW) {
                    const int J = FA[3 * k], c0 = FA[3 * k + 1], c1 = FA[3 * k + 2];
                    double2 pJ = XU2[J];
                    for (int cc = c0; cc < c1; cc++) {
                        const int tr = FC[cc];
                        const double2 pt = XU2[tr], et = EF[tr];
                        ng_fold(pt.y, pt.x, et.x, et.y, pJ.y, pJ.x);
                    }
                    XU2[J] = pJ;
                }
                sync();
            }
            const int k1 = LP[L + 1];
            for (int k = LP[L] + g; k < k1; k += LW) {
                const short *ch = CH + 5 * k;
                const int i0 = ch[0], i1 = ch[1], tA = ch[2], tB = ch[3];
                const int ta = max(tA, 0), tb = max(tB, 0), j0 = min(i0, i1 - 1);
                const double2 p0 = XU2[i0 - 1], pa = XU2[ta], ea = EF[ta], pb = XU2[tb], eb = EF[tb];
                double2 ef = EF[i0 - 1], pn = XU2[j0];
                double u = p0.y, b = p0.x;
                if (tA >= 0) {
                    ng_fold(pa.y, pa.x, ea.x, ea.y, u, b);
                    if (tB >= 0) ng_fold(pb.y, pb.x, eb.x, eb.y, u, b);
                    XU2[i0 - 1] = make_double2(b, u);
                }
                for (int i = i0; i < i1; i++) {
                    const int j = min(i + 1, i1 - 1);
                    const double2 ef2 = EF[j - 1], pn2 = XU2[j];
                    double q1, q2;
                    ng_div2(ef.x, b * ef.x, u, q1, q2);
                    const double unew = pn.y - ef.y * q1;
                    const double bnew = pn.x - q2;
                    XU2[i] = make_double2(bnew, unew);
                    u = unew; b = bnew;
                    ef = ef2; pn = pn2;
                }
            }
            sync();
        }

        /* pivot quotients of every row, in parallel: PX = b/u and PF = -f/u
         * (BKSUB1's first step, CudaStuff.cu:75-78) replace b and u */
        for (int h = g; h < N; h += LW) {
            const double2 p = XU2[h];
            const double r = ng_rcp(p.y);
            XU2[h] = make_double2(p.x * r, -(EF[h].y * r));
        }
        sync();

        /* back-substitution by chains, highest level first: a chain's last
         * row takes x of its parent row pe (a row of a higher level; pe = N
         * marks the root, x = PX), then x_i = PX_i + PF_i x_{i+1} down to the
         * chain's first row, with the quotients loaded two rows ahead; x
         * replaces PX */
        for (int L = a.depth; L >= 0; L--) {
            const int k1 = LP[L + 1];
            for (int k = LP[L] + g; k < k1; k += LW) {
                const short *ch = CH + 5 * k;
                const int i0 = ch[0], i1 = ch[1], pe = ch[4];
                const int h = i1 - 1, lo = i0 - 1;
                const int j1 = max(h - 1, lo), j2 = max(h - 2, lo);
                const double2 ph = XU2[h];
                const double xp = XU[2 * min(pe, N - 1)];
                double2 p1 = XU2[j1], p2 = XU2[j2];
                double x = ph.x;
                if (pe < N) x = x + ph.y * xp;
                XU[2 * h] = x;
                for (int i = h - 1; i >= lo; i--) {
                    const double2 p3 = XU2[max(i - 2, lo)];
                    x = p1.x + p1.y * x;
                    XU[2 * i] = x;
                    p1 = p2; p2 = p3;
                }
            }
            sync();
        }

        /* v update (CudaStuff.cu:312-315) */
#pragma unroll
        for (int c = 0; c < NSLOT; c++)
            { const int h = QH[c * LW + g]; if (h >= 0) v[c] += XU[2 * h]; }
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
        for (int i = g; i < N; i += LW) CY[2 * NQ + i] = XU[2 * i];
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

template <class SET, int NSLOT, bool MULTI, bool STAGE>
__global__ void __launch_bounds__(MULTI ? 512 : 256) __attribute__((amdgpu_waves_per_eu(NG_WAVES_PER_EU)))
ngpu_sim_kernel(SimArgs a)
{
    extern __shared__ double lds[];
    const int LW = a.LW;
    const int group = threadIdx.x / LW;
    const int g = threadIdx.x - group * LW;
    const Tables T = stage_tables<STAGE>(a, lds, a.runs_per_block, NSLOT * LW);
    __syncthreads();
    const int run = blockIdx.x * a.runs_per_block + group;
    if (run >= a.nruns) return;          /* W > 1 implies one run per workgroup */
    double *XU = T.runs + (size_t)group * ng_lds_per_run(a.N, STAGE ? a.nvar : 0, a.nrec);
    if constexpr (MULTI) sim_run<SET, NSLOT, STAGE>(a, T, XU, run, g, SyncBlock{});
    else sim_run<SET, NSLOT, STAGE>(a, T, XU, run, g, SyncWave{});
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
        const SyncDyn sync{a.block_sync != 0};
        (void)nrun;                      /* groups beyond nrun simulate padding runs */
        if (a.lds_staging) {
            const Tables T = stage_tables<true>(a, lds, R, NSLOT * LW);
            __syncthreads();
            double *XU = T.runs + (size_t)group * ng_lds_per_run(a.N, a.nvar, a.nrec);
            if (a.nslot <= NSLOT_SMALL) sim_run<SET, NSLOT_SMALL, true>(a, T, XU, run0 + group, g, sync);
            else sim_run<SET, NSLOT, true>(a, T, XU, run0 + group, g, sync);
        } else {
            const Tables T = stage_tables<false>(a, lds, R, NSLOT * LW);
            __syncthreads();
            double *XU = T.runs + (size_t)group * ng_lds_per_run(a.N, 0, a.nrec);
            if (a.nslot <= NSLOT_SMALL) sim_run<SET, NSLOT_SMALL, false>(a, T, XU, run0 + group, g, sync);
            else sim_run<SET, NSLOT, false>(a, T, XU, run0 + group, g, sync);
        }
        __syncthreads();
        if (threadIdx.x == 0) b.stamps[2 * it + 1] = wall_clock64();
    }
}

/* Host-side launchers of one mechanism set's kernels (instantiated once per
 * set, in sim_<set>.hip). */
template <class SET, int NS, bool MULTI, bool STAGE>
static hipError_t ng_launch_sim(dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    auto k = ngpu_sim_kernel<SET, NS, MULTI, STAGE>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, grid, block, lds, s, a);
    return hipGetLastError();
}

template <class SET, int NS>
static hipError_t ng_launch_sim_ns(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    if (multi) return a.lds_staging ? ng_launch_sim<SET, NS, true, true>(grid, block, lds, s, a)
                                    : ng_launch_sim<SET, NS, true, false>(grid, block, lds, s, a);
    return a.lds_staging ? ng_launch_sim<SET, NS, false, true>(grid, block, lds, s, a)
                         : ng_launch_sim<SET, NS, false, false>(grid, block, lds, s, a);
}

/* nslot positions per lane (1, 2, 4, 6 or 8); multi: more than one wave per run */
template <class SET>
static hipError_t ng_launch_set(int nslot, bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                                const SimArgs &a)
{
    switch (nslot) {
    case 1: return ng_launch_sim_ns<SET, 1>(multi, grid, block, lds, s, a);
    case 2: return ng_launch_sim_ns<SET, 2>(multi, grid, block, lds, s, a);
    case 4: return ng_launch_sim_ns<SET, 4>(multi, grid, block, lds, s, a);
    case 6: return ng_launch_sim_ns<SET, 6>(multi, grid, block, lds, s, a);
    case 8: return ng_launch_sim_ns<SET, 8>(multi, grid, block, lds, s, a);
    default: return hipErrorInvalidValue;
    }
}

template <class SET>
static hipError_t ng_launch_batch_set(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b)
{
    auto k = ngpu_batch_kernel<SET, NG_BATCH_NSLOT, 4>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, grid, dim3(NG_BATCH_THREADS), lds, s, b);
    return hipGetLastError();
}

#endif
