/*
 * sim_kernel.h -- the persistent simulation kernel for gfx950.
 *
 * A group of W wavefronts (LW = 64*W lanes) advances R runs (run = parameter
 * set x stimulus) together through steps t0..t1; a workgroup holds one or
 * more groups.  The positions of the R runs are spread over the group's
 * lanes: position p = c*LW + lane is position q = p mod NQ of run p / NQ,
 * where the NQ >= N positions of a run are its segments sorted by mechanism
 * mask (so the 64 lanes of a wave share a run and mostly one mask).  In the
 * solve phases the lanes take (run, chain) or (run, row) tasks, so the serial
 * chain latency of the tree is paid once for R runs.  Per step:
 *
 *   1. mechanism pass: per position, load the mechanism states, v and dv
 *      (state slots), apply the previous step's cnexp at v, evaluate
 *      BREAKPOINT at v+0.001 and v, store the states, and assemble the Hines
 *      row (rhs, diagonal) in LDS, folding in the children's terms in son
 *      order (the reference's separate scatter pass);
 *   2. triangularisation by tree levels: each task eliminates one chain of
 *      consecutive rows (a section), folding the ends of finished child
 *      chains into the chain's first row, with the reference's arithmetic;
 *      then every row's pivot quotients PX = b/u and PF = -f/u, in parallel;
 *   3. back-substitution by the same chains in reverse level order,
 *      x_i = PX_i + PF_i * x_parent(i), the parent of a chain's last row
 *      coming from a chain of a higher level;
 *   4. per position: v += x, dv += x[parent] - x[self], stored.
 *
 * Registers hold one position's values at a time, so the number of positions
 * per lane is a run-time loop count.  Groups of one wave synchronise at
 * wavefront scope (no barrier instruction), larger groups at workgroup scope.
 *
 * Two kernels share the per-group code (sim_runs):
 *   ngpu_sim_kernel   one morphology; workgroup b runs runs_per_block runs;
 *   ngpu_batch_kernel persistent: each workgroup of 8 waves takes items from a
 *                     device work queue (an atomic counter over an item list);
 *                     an item is the runs of one workgroup of one morphology,
 *                     so runs of uneven size and step count share one launch.
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
 * same loaded state, and the dv update of step t is done at the end of step
 * t-1 (the same values).  The elimination keeps the reference's operation
 * order (u_i -= f (e / u_{i-1}), b_i -= b_{i-1} e / u_{i-1}): the reference
 * trees couple zero-area nodes with coefficients near 1e22, and there the
 * order of these operations, not the rounding of the membrane terms, sets
 * the result.
 */
#ifndef NGPU_SIM_KERNEL_H
#define NGPU_SIM_KERNEL_H

#include <hip/hip_runtime.h>
#include "mechs.h"

#define NG_EPS_V 0.001

/* Per-phase cycle counters, compiled in only with -DNG_PHASE_CYCLES=1 (the
 * profiling build, `make phases`): thread 0 of every workgroup adds the
 * cycles (s_memtime) of each phase of each step -- mechanism pass,
 * elimination, pivot quotients, back-substitution, v/dv update and trace --
 * and the number of steps to ng_phase_cycles[0..5]. */
#ifndef NG_PHASE_CYCLES
#define NG_PHASE_CYCLES 0
#endif
#if NG_PHASE_CYCLES
static __device__ unsigned long long ng_phase_cycles[8];
#define NG_STAMP(var) const unsigned long long var = clock64()
#else
#define NG_STAMP(var)
#endif

struct SimArgs {
    int N;                               /* segments */
    int NQ;                              /* positions per run: N rounded up to a multiple of 64 */
    int LW;                              /* lanes per group (64 * waves) */
    int runs_per_group;                  /* R: runs sharing a group's lanes */
    int nvar, nrec, nstim, nt, nt_amps, t0, t1, nruns, depth;
    int stim_q;
    double stim_area, v_init;
    const int *q_meta;                   /* [NQ] mechanism mask (0 for padding); bit 15: a
                                            run-time-mask position */
    const int *q_link;                   /* [4*NQ] row, parent row, first- and second-son position
                                            (-1 = none); row -1 marks padding */
    const double *q_cm;                  /* [NQ] */
    const double *row_ef;                /* [2N + NG_ROW_PAD]: e, f interleaved by Hines row, then zeros */
    const double *pval;                  /* [NPARAM*NQ] value, or tagged NaN = varying column */
    const double *pv;                    /* [nsets*nvar] varying parameter values */
    unsigned pconst_mask;                /* bit j: parameter j has one value wherever it is used */
    const double *pconst;                /* [NPARAM] those values */
    const int *msched;                   /* mechanism-pass segment queue: the count, then the
                                            segments chunk << 8 | r0 << 4 | n, or k << 8 for
                                            packed segment k, longest first */
    const int *gen_q;                    /* [ngen] the run-time-mask positions (packed segments) */
    int ngen;
    const int *chains, *lvl_ptr;         /* chains: (i0, i1, tA, tB, pe); lvl_ptr[depth+2] */
    const int *fath, *flvl_ptr, *fchild; /* fathers with their own phase: (J, c0, c1); flvl_ptr[depth+1] */
    int lds_staging;                     /* row coefficients e, f and the runs' varying
                                            parameter values staged in LDS (else read
                                            from global memory); selects the kernel's
                                            STAGE instantiation */
    int block_sync;                      /* batch kernel: synchronise at workgroup scope */
    int nchains, nfath, nfchild;         /* schedule sizes (staged into LDS as int16) */
    int tail_level;                      /* levels tail_level..depth hold one chain each and no
                                            fathers of their own (depth + 1: none) */
    int fast_solve;                      /* reciprocal-form elimination (well-scaled trees);
                                            0: the reference's operation order */
    const double *dts, *amps;
    const int *rec_q;
    double *states;                      /* [nruns][NSTATE + 2][NQ]: mechanism states, v, dv */
    double *vhot;                        /* [nruns][nrec][nt] */
};

/* Synchronisation of a group.  sync() orders LDS accesses; mem() also makes
 * the group's global stores visible to its other lanes (the dv that a
 * parent reads from its children). */
struct SyncWave {
    __device__ __forceinline__ void operator()() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __device__ __forceinline__ void mem() const
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
};
struct SyncBlock {
    __device__ __forceinline__ void operator()() const { __syncthreads(); }
    __device__ __forceinline__ void mem() const { __syncthreads(); }
};
struct SyncDyn {
    bool block;
    __device__ __forceinline__ void operator()() const { if (block) __syncthreads(); else SyncWave{}(); }
    __device__ __forceinline__ void mem() const { if (block) __syncthreads(); else SyncWave{}.mem(); }
};

/* LDS layout of a workgroup, in this order:
 *   EF[2N]                 e, f interleaved by Hines row, when staged
 *   PAD[16]                eight rows of padding (read, never used: the solve's
 *                          batches of eight rows are not clamped to the chain)
 *   per run (runs_per_block of them): XU[2(N+1)], PV[npv], REC[64*nrec]
 *                          doubles; XU holds (b, u), then (PX, PF), then
 *                          (x, PF) interleaved by row, so that a row is one
 *                          16-byte access; npv = nvar rounded up to even when
 *                          staged, else 0 (regions stay 16-byte aligned)
 *   PAD[16]                eight rows of padding
 *   QL[NQ]                 short4: row, parent row, son positions
 *   QM[NQ]                 unsigned short: mechanism mask; bit 15: a run-time-mask
 *                          position, run by the packed segments
 *   CH[8*nchains], FA[3*nfath], FC[nfchild], LP[depth+2], FLP[depth+1] short */
/* Doubles of padding before and after the runs' regions in LDS (and rows of
 * zeros after the row coefficients in global memory): eight rows, the most a
 * solve batch reads past either end of a chain. */
#define NG_ROW_PAD 16
__host__ __device__ inline size_t ng_lds_per_run(int N, int npv, int nrec)
{
    return (size_t)(2 * (N + 1) + ((npv + 1) & ~1) + 64 * nrec);
}
__host__ __device__ inline size_t ng_lds_bytes(int N, int NQ, int runs, int npv, int nrec, int staged,
                                               int nchains, int nfath, int nfchild, int depth)
{
    const size_t nshort = 5 * (size_t)NQ + 8 * (size_t)nchains + 3 * (size_t)nfath + (size_t)nfchild +
                          (size_t)(depth + 2) + (size_t)(depth + 1);
    const size_t dbl = (staged ? 2 * (size_t)N : 0) + 2 * NG_ROW_PAD + (size_t)runs * ng_lds_per_run(N, staged ? npv : 0, nrec);
    return dbl * sizeof(double) + ((nshort * sizeof(short) + 7) / 8) * 8;
}

/* Element idx of a global array as base + a 32-bit byte offset (a scalar
 * base and one offset register per lane; every array indexed so holds less
 * than 4 GiB). */
template <class T>
__device__ __forceinline__ T &ng_at(T *base, int idx)
{
    return *(T *)((char *)base + (unsigned)idx * (unsigned)sizeof(T));
}

/* Parameter tables hold either the value or, for a run-varying column k, a
 * NaN whose high word is NG_PTAG and low word k (decoded from the run's
 * column values).  Genuine NaN parameters (quiet NaN payloads) pass through. */
#define NG_PTAG 0x7FF40000u

/* Parameter j at position q, read where a mechanism uses it: a value shared
 * by every segment comes from the constant table (a scalar load), otherwise
 * the per-position table gives the value or the tag of a run-varying column.
 * The tag is resolved without a branch (the run's first column is read when
 * there is no tag), so that the loads of several parameters are issued
 * together instead of each waiting for its table entry. */
struct ParamFetch {
    const double *pconst, *pval, *pv;
    unsigned mask;
    int q, NQ;
    __device__ __forceinline__ double operator()(int j) const
    {
        if (mask & (1u << j)) return pconst[j];
        const double x = ng_at(pval, j * NQ + q);
        const bool tag = (unsigned)__double2hiint(x) == NG_PTAG;
        const double y = pv[tag ? (unsigned)__double2loint(x) : 0u];
        return tag ? y : x;
    }
};

/* The parameter entries of position q (value or tag of a run-varying
 * column), read once for all runs of a position; with a fixed mechanism mask
 * only the entries its mechanisms use are loaded. */
template <int NP>
struct ParamEntries {
    double x[NP];
    unsigned cmask;
    __device__ __forceinline__ ParamEntries(const SimArgs &a, int q) : cmask(a.pconst_mask)
    {
#pragma unroll
        for (int j = 0; j < NP; j++) x[j] = (cmask & (1u << j)) ? a.pconst[j] : ng_at(a.pval, j * a.NQ + q);
    }
};
/* Parameter j of one run at that position: a tag takes the run's column
 * value (the run's first value is read when there is no tag, no branch). */
template <int NP>
struct ParamRun {
    const ParamEntries<NP> &e;
    const double *pv;
    __device__ __forceinline__ double operator()(int j) const
    {
        const double x = e.x[j];
        if (e.cmask & (1u << j)) return x;
        const bool tag = (unsigned)__double2hiint(x) == NG_PTAG;
        const double y = pv[tag ? (unsigned)__double2loint(x) : 0u];
        return tag ? y : x;
    }
};

/* Pointers into the block-shared part of the LDS layout; EF points to LDS or
 * to global memory depending on the STAGE instantiation, so that every access
 * compiles to an LDS or a global load (never a flat one). */
struct Tables {
    const double2 *EF;                   /* (e, f) by Hines row */
    double *runs;                        /* first per-run region */
    const short4 *QL;
    const unsigned short *QM;
    const short *CH, *FA, *FC, *LP, *FLP;
};

/* Stage the block-shared tables of morphology a for `runs` per-run regions. */
template <bool STAGE>
__device__ __forceinline__ Tables stage_tables(const SimArgs &a, double *lds, int runs)
{
    const int N = a.N, NQ = a.NQ;
    const int npv = STAGE ? a.nvar : 0;
    Tables T;
    if constexpr (STAGE) {
        for (int i = threadIdx.x; i < 2 * N; i += blockDim.x) lds[i] = a.row_ef[i];
        T.EF = (const double2 *)lds;
        T.runs = lds + 2 * N + NG_ROW_PAD;
    } else {
        T.EF = (const double2 *)a.row_ef;
        T.runs = lds + NG_ROW_PAD;
    }
    short4 *QL = (short4 *)(T.runs + (size_t)runs * ng_lds_per_run(N, npv, a.nrec) + NG_ROW_PAD);
    unsigned short *QM = (unsigned short *)(QL + NQ);
    short *CH = (short *)(QM + NQ);
    short *FA = CH + 8 * a.nchains, *FC = FA + 3 * a.nfath, *LP = FC + a.nfchild, *FLP = LP + (a.depth + 2);
    for (int i = threadIdx.x; i < NQ; i += blockDim.x) {
        QL[i] = make_short4((short)a.q_link[4 * i], (short)a.q_link[4 * i + 1], (short)a.q_link[4 * i + 2],
                            (short)a.q_link[4 * i + 3]);
        QM[i] = (unsigned short)a.q_meta[i];
    }
    for (int i = threadIdx.x; i < 8 * a.nchains; i += blockDim.x) {
        const int k = i >> 3, m = i & 7;
        CH[i] = m < 5 ? (short)a.chains[5 * k + m] : (short)0;
    }
    for (int i = threadIdx.x; i < 3 * a.nfath; i += blockDim.x) FA[i] = (short)a.fath[i];
    for (int i = threadIdx.x; i < a.nfchild; i += blockDim.x) FC[i] = (short)a.fchild[i];
    for (int i = threadIdx.x; i < a.depth + 2; i += blockDim.x) LP[i] = (short)a.lvl_ptr[i];
    for (int i = threadIdx.x; i < a.depth + 1; i += blockDim.x) FLP[i] = (short)a.flvl_ptr[i];
    T.QL = QL; T.QM = QM; T.CH = CH; T.FA = FA; T.FC = FC; T.LP = LP; T.FLP = FLP;
    return T;
}

/* Eliminate the final row (ut, bt) with coefficients e, f into the running
 * father values (uJ, bJ) with the reference's arithmetic
 * (CudaStuff.cu:53-64): uJ -= f (e / ut), bJ -= (bt e) / ut. */
__device__ __forceinline__ void ng_fold(double ut, double bt, double e, double f, double &uJ, double &bJ)
{
    uJ = uJ - f * (e / ut);
    bJ = bJ - (bt * e) / ut;
}

/* Steps t0..t1 of runs run0 .. run0 + R - 1 by the LW lanes of a group
 * (lane g).  RG is the first of the group's R per-run LDS regions; qctr is
 * the group's segment counter in LDS (the mechanism pass's queue).  Runs from
 * nruns up to the allocated count are padding (their parameters are the last
 * set's): partial workgroups simulate them instead of leaving lanes idle. */
/* tk / n for 0 <= tk < 8 n and 1 <= n < 32768 (a task's run): (tk + 1/2) / n
 * lies at least 1/(2n) > 1.5e-5 from an integer, and the single-precision
 * reciprocal and product are within 2e-6 of it below 8, so truncation gives
 * the quotient exactly. */
__device__ __forceinline__ int ng_task_run(int tk, int n)
{
    return (int)(((float)tk + 0.5f) * __builtin_amdgcn_rcpf((float)n));
}

/* A chain task of a level: run r and the chain's (i0, i1, tA, tB, pe), read
 * from the chain table (eight shorts a chain, two 8-byte loads). */
struct NgChain {
    int r, i0, i1, tA, tB, pe;
};
__device__ __forceinline__ NgChain ng_chain(const short *CH, int tk, int k0, int nc)
{
    NgChain c;
    c.r = ng_task_run(tk, nc);
    const int k = k0 + tk - c.r * nc;
    const short4 a = ((const short4 *)CH)[2 * k], b = ((const short4 *)CH)[2 * k + 1];
    c.i0 = a.x; c.i1 = a.y; c.tA = a.z; c.tB = a.w; c.pe = b.x;
    return c;
}

/* Position p = r * n + q of a strided pass (p = g, g + LW, ...), stepped
 * without an integer division per iteration. */
struct NgPos {
    int r, q, dr, dq, n;
    __device__ __forceinline__ NgPos(int p, int stride, int n_) : n(n_)
    {
        r = p / n; q = p - r * n; dr = stride / n; dq = stride - dr * n;
    }
    __device__ __forceinline__ void step()
    {
        q += dq; r += dr;
        if (q >= n) { q -= n; r++; }
    }
};

template <class SET, bool STAGE, class SYNC>
__device__ __forceinline__ void sim_runs(const SimArgs &a, const Tables &T, double *RG, int run0, int g,
                                         const SYNC &sync, unsigned *qctr)
{
    const int LW = a.LW, N = a.N, NQ = a.NQ, R = a.runs_per_group;
    constexpr int NS = SET::NSTATE + 2, SV = SET::NSTATE, SDV = SET::NSTATE + 1;
    const int npv = STAGE ? ((a.nvar + 1) & ~1) : 0;
    const size_t PR = ng_lds_per_run(N, STAGE ? a.nvar : 0, a.nrec);
    const int P = R * NQ;                                 /* positions of the group */
    const double2 *EF = T.EF;
    const short4 *QL = T.QL;
    const unsigned short *QM = T.QM;
    const short *CH = T.CH, *FA = T.FA, *FC = T.FC, *LP = T.LP, *FLP = T.FLP;
    const int nsets = a.nruns / a.nstim;
    /* The positions of a pass are p = g, g + LW, ... with LW and NQ multiples
     * of 64, so the 64 positions of a wave always belong to one run: the
     * run index of the position passes is wave-uniform (scalar registers, so
     * that the run's state and LDS bases are scalar and the loads take a
     * 32-bit offset). */
    auto wave_run = [](int r) { return __builtin_amdgcn_readfirstlane(r); };
    auto region = [&](int r) { return RG + (size_t)r * PR; };
    auto states = [&](int r) { return a.states + (size_t)(run0 + r) * NS * NQ; };
    auto pvof = [&](int r) -> const double * {
        if constexpr (STAGE) return region(r) + 2 * (N + 1);
        else return a.pv + (size_t)min((run0 + r) / a.nstim, nsets - 1) * a.nvar;
    };

    if (g == 0) *qctr = 0u;
    if constexpr (STAGE)
        for (int i = g; i < R * a.nvar; i += LW) {
            const int r = i / a.nvar, k = i - r * a.nvar;
            region(r)[2 * (N + 1) + k] = a.pv[(size_t)min((run0 + r) / a.nstim, nsets - 1) * a.nvar + k];
        }
    sync();
    if (a.t0 == 0) {
        /* INITIAL: states at steady state, v = v_init, dv = 0 */
        NgPos ps(g, LW, NQ);
#pragma unroll 1
        for (int p = g; p < P; p += LW, ps.step()) {
            const int r = wave_run(ps.r), q = ps.q;
            if (q >= N) continue;
            double *S = states(r);
            const unsigned m = QM[q] & 0x7fffu;
            const ParamFetch Pf{a.pconst, a.pval, pvof(r), a.pconst_mask, q, NQ};
            double st[SET::NSTATE];
#pragma unroll
            for (int k = 0; k < SET::NSTATE; k++) st[k] = 0.0;
            SET::init(m, a.v_init, st, Pf);
            const unsigned sb = SET::state_bits(m);
#pragma unroll
            for (int k = 0; k < SET::NSTATE; k++)
                if (sb & (1u << k)) ng_at(S, k * NQ + q) = st[k];
            ng_at(S, SV * NQ + q) = a.v_init;
            ng_at(S, SDV * NQ + q) = 0.0;
        }
    }
    sync.mem();

    double dtp = a.t0 > 0 ? a.dts[a.t0 - 1] : 0.0;
#if NG_PHASE_CYCLES
    unsigned long long ph[6] = {0, 0, 0, 0, 0, 0};
#endif

    for (int t = a.t0; t < a.t1; t++) {
        const double dt = a.dts[t];
        const double cm_dt = ng_rcp(dt * 1000);           /* cm / (dt * 1000) as cm * cm_dt */
        const int slot_t = (t - a.t0) & 63;
        NG_STAMP(T0);

        /* mechanism pass: fused cnexp(t-1) + BREAKPOINT(t) + Hines row
         * (CudaStuff.cu:263-305).  eca is stored by INITIAL only: at t = 0 it
         * is loaded, afterwards it is recomputed from cai after the cnexp. */
        const unsigned keep = (t > 0) ? ~(1u << SET::ECA) : ~0u;
        /* The waves of the group take the segments (chunk c of 64 positions,
         * runs r0 .. r0+n-1) from the group's queue, longest first, each the
         * next one when it has finished its last (an LDS counter), so that
         * they end the pass together.  A wave walks the runs of a chunk in
         * turn with the chunk's mechanism mask, loading the next run's state
         * while it computes the current one.  Parameter entries are the same
         * for every run and are read once per segment.  Positions whose mask
         * has no specialised code skip their chunk's segments: packed
         * segments run them, each lane one (position, run) pair, so the
         * run-time-mask code (which costs the union of its lanes' masks)
         * runs over 64 pairs at a time instead of once per chunk and run. */
        const int nseg = a.msched[0];
#pragma unroll 1
        for (;;) {
            const unsigned lane = __lane_id(), lead = __builtin_amdgcn_readfirstlane(lane);
            unsigned sv = 0u;
            if (lane == lead) sv = __hip_atomic_fetch_add(qctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            const int si = (int)__builtin_amdgcn_readfirstlane(sv);
            if (si >= nseg) break;
            const int d = a.msched[1 + si];
            /* segment code for position q, runs r0 .. r0+n-1: specialised for
             * the set's common mechanism masks (straight-line code, independent
             * gates interleaved), else the run-time-mask code */
            auto seg = [&](auto mask, const int q, const int r0, const int n) {
                constexpr bool fixed = !std::is_same<decltype(mask), unsigned>::value;
                const unsigned m = mask;
                const short4 l = QL[q];                       /* row, parent row, sons */
                const int h = l.x;
                const double cmq = ng_at(a.q_cm, q);
                const double fh = EF[h].y;
                const double e1 = l.z >= 0 ? EF[QL[l.z].x].x : 0.0, e2 = l.w >= 0 ? EF[QL[l.w].x].x : 0.0;
                const unsigned sbn = (SET::state_bits(m) & keep) | (t > 0 ? SET::ica_bits(m) : 0u);
                const unsigned sb = SET::state_bits(m) & keep;
                const unsigned ib = SET::ica_bits(m);
                struct In { double st[SET::NSTATE], vq, dvq, dv1, dv2; };
                auto load = [&](int r, int qq, In &x) {
                    const double *S = states(r);
#pragma unroll
                    for (int k = 0; k < SET::NSTATE; k++) x.st[k] = (sbn & (1u << k)) ? ng_at(S, k * NQ + qq) : 0.0;
                    x.vq = ng_at(S, SV * NQ + qq);
                    x.dvq = ng_at(S, SDV * NQ + qq);
                    x.dv1 = l.z >= 0 ? ng_at(S, SDV * NQ + l.z) : 0.0;
                    x.dv2 = l.w >= 0 ? ng_at(S, SDV * NQ + l.w) : 0.0;
                };
                auto step = [&](int r, int qq, In &cur, const auto &Pf) {
                    double *S = states(r);
                    double *XU = region(r);
                    double *st = cur.st;
                    const double vq = cur.vq;
                    /* record v before the step (CudaStuff.cu:257-261) */
                    for (int k = 0; k < a.nrec; k++)
                        if (q == a.rec_q[k]) XU[2 * (N + 1) + npv + k * 64 + slot_t] = vq;
                    if (t > 0) {
                        SET::cnexp(m, vq, dtp, st, ib ? st[SET::ICA] : 0.0, Pf);
                        SET::eca_after_cnexp(m, st);
#pragma unroll
                        for (int k = 0; k < SET::NSTATE; k++)
                            if (sb & (1u << k)) ng_at(S, k * NQ + qq) = st[k];
                    }
                    double sumDv = 0.0, sum = 0.0, icaDv = 0.0, icaNow = 0.0;
                    SET::current(m, vq + 0.001, st, Pf, sumDv, icaDv, false);
                    SET::current(m, vq, st, Pf, sum, icaNow, true);
                    if (ib) ng_at(S, SET::ICA * NQ + qq) = icaNow;
                    double istim = 0.0;
                    if (q == a.stim_q) {
                        const int stim = (run0 + r) % a.nstim;
                        istim = 100 * a.amps[(size_t)stim * a.nt_amps + t] / a.stim_area;
                    }
                    const double gModel = (sumDv - sum) * (1.0 / NG_EPS_V);
                    double rhs = istim - sum;
                    double D = gModel + cmq * cm_dt;
                    D -= fh;
                    rhs -= fh * cur.dvq;
                    /* the children's terms, first son then second son
                     * (CudaStuff.cu:296-305, ordered) */
                    if (l.z >= 0) { rhs += e1 * cur.dv1; D -= e1; }
                    if (l.w >= 0) { rhs += e2 * cur.dv2; D -= e2; }
                    ((double2 *)XU)[h] = make_double2(rhs, D);
                };
                if constexpr (fixed) {
                    const ParamEntries<SET::NPARAM> pe(a, q);
                    In cur;
                    load(r0, q, cur);
#pragma unroll 1
                    for (int j = 0; j < n; j++) {
                        In nxt = cur;
                        if (j + 1 < n) load(r0 + j + 1, q, nxt);
                        step(r0 + j, q, cur, ParamRun<SET::NPARAM>{pe, pvof(r0 + j)});
                        cur = nxt;
                    }
                } else {
                    /* run-time masks (rare segments): one run at a time, the
                     * parameters read where used; q passes through an empty asm
                     * so that the addresses of the state slots and parameter
                     * entries are not hoisted out of the run loop (which would
                     * hold one for every slot and parameter of the set) */
#pragma unroll 1
                    for (int j = 0; j < n; j++) {
                        int qp = q;
                        __asm__ volatile("" : "+v"(qp));
                        In cur;
                        load(r0 + j, qp, cur);
                        step(r0 + j, qp, cur, ParamFetch{a.pconst, a.pval, pvof(r0 + j), a.pconst_mask, qp, NQ});
                    }
                }
            };
            if ((d & 15) == 0) {
                /* packed segment: lane i takes (position, run) pair 64 k + i of
                 * the group's run-time-mask positions */
                const int p = (d >> 8) * 64 + (g & 63);
                if (p >= a.ngen * R) continue;
                const int r = p / a.ngen;
                const int q = a.gen_q[p - r * a.ngen];
                seg(QM[q] & 0x7fffu, q, r, 1);
            } else {
                const int q = ((d >> 8) << 6) + (g & 63);
                if (q >= N || (QM[q] & 0x8000u)) continue;    /* run-time-mask positions: packed */
                const int r0 = (d >> 4) & 15, n = d & 15;
                SET::dispatch(QM[q] & 0x7fffu, [&](auto mask) { seg(mask, q, r0, n); });
            }
        }
        sync();
        if (g == 0) *qctr = 0u;                           /* every wave has taken its last segment */

        NG_STAMP(T1);
        /* triangularisation by levels (BeforeLU, CudaStuff.cu:31-69), with the
         * reference's arithmetic.  Level L first folds the fathers of level
         * L-1 that have their own phase, then eliminates its chains; a chain
         * (i0, i1, tA, tB, pe) folds up to two related rows tA, tB (-1 = none)
         * into its first row i0-1, then eliminates rows i0-1 .. i1-2 each into
         * the next.  A task is (run, chain); every load that does not depend
         * on the chain's own results is issued up front, and the rows' values
         * are loaded two rows ahead (indices clamped to the chain; a clamped
         * value is not used).  The divisions are the IEEE ones: a shared
         * refined reciprocal is as fast, but its operand-range test and branch
         * cost more per row than the two divisions.  A lane's first task of
         * the next level is read from the (constant) chain table before the
         * level's barrier. */
        const int D = a.depth, Lt = min(a.tail_level, D + 1);
        /* a chain task: fold the related rows into the first row, then
         * eliminate the rows in order */
        auto elim = [&](const NgChain &c) {
            double2 *XU2 = (double2 *)region(c.r);
            const int i0 = c.i0, i1 = c.i1, tA = c.tA, tB = c.tB;
            const int ta = max(tA, 0), tb = max(tB, 0), j0 = min(i0, i1 - 1);
            double2 p0 = XU2[i0 - 1], pa = XU2[ta], ea = EF[ta], pb = XU2[tb], eb = EF[tb];
            const int j1 = min(i0 + 1, i1 - 1);
            double2 ef = EF[i0 - 1], pn = XU2[j0], ef1 = EF[max(j1 - 1, 0)], pn1 = XU2[j1];
            /* all of them in flight together (not sunk into the branches) */
            __asm__ volatile("" : "+v"(p0.x), "+v"(p0.y), "+v"(pa.x), "+v"(pa.y), "+v"(ea.x), "+v"(ea.y),
                             "+v"(pb.x), "+v"(pb.y), "+v"(eb.x), "+v"(eb.y));
            double u = p0.y, b = p0.x;
            if (tA >= 0) {
                ng_fold(pa.y, pa.x, ea.x, ea.y, u, b);
                if (tB >= 0) ng_fold(pb.y, pb.x, eb.x, eb.y, u, b);
                XU2[i0 - 1] = make_double2(b, u);
            }
            for (int i = i0; i < i1; i++) {
                const int j = min(i + 2, i1 - 1);
                const double2 ef2 = EF[j - 1], pn2 = XU2[j];
                const double unew = pn.y - ef.y * (ef.x / u);
                const double bnew = pn.x - (b * ef.x) / u;
                XU2[i] = make_double2(bnew, unew);
                u = unew; b = bnew;
                ef = ef1; pn = pn1; ef1 = ef2; pn1 = pn2;
            }
        };
        /* The reciprocal form of the same elimination (fast_solve, for trees
         * whose coefficients are well scaled): row i's pivot quotients
         * PX = b w and PF = -f w with w = 1/u (the hardware reciprocal refined
         * by two Newton steps) replace (b, u), and the next row is
         * u' = d + e PF, b' = rhs - e PX, i.e. u_i -= f (e / u) and
         * b_i -= (b e) / u with the quotients taken once.  Related rows are
         * folded through their quotients the same way.  The results equal the
         * reference order's up to rounding. */
        auto elim_fast = [&](const NgChain &c) {
            double2 *XU2 = (double2 *)region(c.r);
            const int i0 = c.i0, last = c.i1 - 1, tA = c.tA, tB = c.tB;
            const int ta = max(tA, 0), tb = max(tB, 0);
            const double2 x = XU2[i0 - 1], pa = XU2[ta], ea = EF[ta], pb = XU2[tb], eb = EF[tb];
            /* rows i .. i+7 of the chain at a time: row i needs its (e, f)
             * and the next row's (rhs, d); all eight loads are issued before
             * the first row is computed, at immediate offsets from one
             * address (past the chain's last row they read other rows or the
             * padding after the regions and the coefficients; those values
             * are not used) */
            int i = i0 - 1;
            double2 ef[8], xn[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                ef[k] = EF[i + k];
                xn[k] = XU2[i + k + 1];
            }
            double u = x.y, b = x.x;
            if (tA >= 0) {
                u = fma(ea.x, pa.y, u);
                b = fma(-ea.x, pa.x, b);
                if (tB >= 0) {
                    u = fma(eb.x, pb.y, u);
                    b = fma(-eb.x, pb.x, b);
                }
            }
            /* row i: w = 1/u, PX = b w, PF = -f_i w replace (b, u); the next
             * row is u' = d + e_i PF, b' = rhs - e_i PX */
            for (;;) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const double w = ng_rcp(u);
                    const double PX = b * w, PF = -(ef[k].y * w);
                    XU2[i] = make_double2(PX, PF);
                    if (i >= last) return;
                    u = fma(ef[k].x, PF, xn[k].y);
                    b = fma(-ef[k].x, PX, xn[k].x);
                    i++;
                }
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    ef[k] = EF[i + k];
                    xn[k] = XU2[i + k + 1];
                }
            }
        };
        const bool fast = a.fast_solve != 0;
        int k0 = LP[0], nc = LP[1] - k0;
        int k0n = LP[min(1, D)], ncn = LP[min(1, D) + 1] - k0n;
        NgChain first = ng_chain(CH, min(g, R * nc - 1), k0, nc);
        for (int L = 0; L < Lt; L++) {
            /* level L+2's range and level L+1's first task, read ahead */
            const int k0nn = LP[min(L + 2, D)], ncnn = LP[min(L + 2, D) + 1] - k0nn;
            const NgChain next = ng_chain(CH, min(g, R * ncn - 1), k0n, ncn);
            if (a.nfath > 0 && L > 0 && FLP[L] > FLP[L - 1]) {
                const int f0 = FLP[L - 1], nf = FLP[L] - f0;
                for (int tk = g; tk < R * nf; tk += LW) {
                    const int r = ng_task_run(tk, nf), k = f0 + tk - r * nf;
                    double2 *XU2 = (double2 *)region(r);
                    const int J = FA[3 * k], c0 = FA[3 * k + 1], c1 = FA[3 * k + 2];
                    double2 pJ = XU2[J];
                    for (int cc = c0; cc < c1; cc++) {
                        const int tr = FC[cc];
                        const double2 pt = XU2[tr], et = EF[tr];
                        if (fast) {
                            /* (b, u) of J from the quotients (PX, PF) of row tr */
                            pJ.y = fma(et.x, pt.y, pJ.y);
                            pJ.x = fma(-et.x, pt.x, pJ.x);
                        } else {
                            ng_fold(pt.y, pt.x, et.x, et.y, pJ.y, pJ.x);
                        }
                    }
                    XU2[J] = pJ;
                }
                sync();
            }
            for (int tk = g; tk < R * nc; tk += LW) {
                NgChain c = first;
                if (tk != g) c = ng_chain(CH, tk, k0, nc);
                if (fast) elim_fast(c);
                else elim(c);
            }
            sync();
            first = next; k0 = k0n; nc = ncn; k0n = k0nn; ncn = ncnn;
        }
        /* the tail: one chain per level, each depending only on the chain
         * below it of the same run, so lane r takes run r's chains of all
         * tail levels in level order without a barrier between them */
        if (Lt <= D) {
            if (g < R)
                for (int L = Lt; L <= D; L++) {
                    if (fast) elim_fast(ng_chain(CH, g, LP[L], 1));
                    else elim(ng_chain(CH, g, LP[L], 1));
                }
            sync();
        }

        NG_STAMP(T2);
        /* pivot quotients of every row, in parallel: PX = b/u and PF = -f/u
         * (BKSUB1's first step, CudaStuff.cu:75-78) replace b and u; the
         * reciprocal-form elimination has stored them already */
        if (!fast) {
            NgPos rw(g, LW, N);
            for (int i = g; i < R * N; i += LW, rw.step()) {
                const int r = rw.r, h = rw.q;
                double2 *XU2 = (double2 *)region(r);
                const double2 ph = XU2[h];
                const double rc = ng_rcp(ph.y);
                XU2[h] = make_double2(ph.x * rc, -(EF[h].y * rc));
            }
            sync();
        }

        NG_STAMP(T3);
        /* back-substitution by chains, highest level first: a chain's last
         * row takes x of its parent row pe (a row of a higher level; pe = N
         * marks the root, x = PX), then x_i = PX_i + PF_i x_{i+1} down to the
         * chain's first row, the quotients of eight rows loaded at a time;
         * x replaces PX */
        auto bsub = [&](const NgChain &c) {
            double *XU = region(c.r);
            const double2 *XU2 = (const double2 *)XU;
            const int lo = c.i0 - 1;
            int i = c.i1 - 1;
            /* rows i .. i-7 loaded at once, at immediate offsets from one
             * address (below the chain's first row they read other rows or
             * the padding before the regions; those values are not used);
             * the root's parent value is 0 */
            double2 v[8];
#pragma unroll
            for (int k = 0; k < 8; k++) v[k] = XU2[i - k];
            double x = c.pe < N ? XU[2 * c.pe] : 0.0;
            for (;;) {
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    x = fma(v[k].y, x, v[k].x);
                    XU[2 * i] = x;
                    if (--i < lo) return;
                }
#pragma unroll
                for (int k = 0; k < 8; k++) v[k] = XU2[i - k];
            }
        };
        /* the tail first (top of the tree), lane r for run r, then the
         * levels below it */
        if (Lt <= D) {
            if (g < R)
                for (int L = D; L >= Lt; L--) bsub(ng_chain(CH, g, LP[L], 1));
            sync();
        }
        const int Ls = min(Lt - 1, D);
        if (Ls >= 0) {
            k0 = LP[Ls]; nc = LP[Ls + 1] - k0;
            k0n = LP[max(Ls - 1, 0)]; ncn = LP[max(Ls - 1, 0) + 1] - k0n;
            first = ng_chain(CH, min(g, R * nc - 1), k0, nc);
        }
        for (int L = Ls; L >= 0; L--) {
            const int k0nn = LP[max(L - 2, 0)], ncnn = LP[max(L - 2, 0) + 1] - k0nn;
            const NgChain next = ng_chain(CH, min(g, R * ncn - 1), k0n, ncn);
            for (int tk = g; tk < R * nc; tk += LW) {
                NgChain c = first;
                if (tk != g) c = ng_chain(CH, tk, k0, nc);
                bsub(c);
            }
            sync();
            first = next; k0 = k0n; nc = ncn; k0n = k0nn; ncn = ncnn;
        }

        NG_STAMP(T4);
        /* v += x (CudaStuff.cu:312-315) and dv += x[parent] - x[self] for the
         * next step (CudaStuff.cu:287) */
        NgPos pu(g, LW, NQ);
#pragma unroll 4
        for (int p = g; p < P; p += LW, pu.step()) {
            /* padding positions read a real one and store nothing, so that
             * the loads of several positions can be in flight together */
            const int r = wave_run(pu.r), q = pu.q, qr = min(q, N - 1);
            const double *XU = region(r);
            double *S = states(r);
            const short4 l = QL[qr];
            const double xh = XU[2 * l.x], xpar = XU[2 * l.y];
            const double vn = ng_at(S, SV * NQ + qr) + xh, dvn = ng_at(S, SDV * NQ + qr) + (xpar - xh);
            if (q < N) { ng_at(S, SV * NQ + q) = vn; ng_at(S, SDV * NQ + q) = dvn; }
        }
        dtp = dt;

        /* flush the recorded chunk */
        if (slot_t == 63 || t == a.t1 - 1) {
            const int base = t - slot_t, nrec64 = a.nrec * 64;
            for (int i = g; i < R * nrec64; i += LW) {
                const int r = i / nrec64, k = i - r * nrec64, j = k & 63;
                if (j <= slot_t)
                    a.vhot[((size_t)(run0 + r) * a.nrec + (k >> 6)) * a.nt + base + j] = region(r)[2 * (N + 1) + npv + k];
            }
        }
        sync.mem();
#if NG_PHASE_CYCLES
        NG_STAMP(T5);
        ph[0] += T1 - T0; ph[1] += T2 - T1; ph[2] += T3 - T2; ph[3] += T4 - T3; ph[4] += T5 - T4; ph[5] += 1;
#endif
    }
#if NG_PHASE_CYCLES
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) atomicAdd(&ng_phase_cycles[k], ph[k]);
#endif
}

#ifndef NG_WAVES_PER_EU
#define NG_WAVES_PER_EU 2
#endif

/* Most groups a workgroup holds (16 one-wave groups of 1,024 threads). */
#define NG_MAX_GROUPS 16

/* One morphology.  Workgroup b holds G = blockDim.x / LW groups of R runs:
 * runs b * G * R + group * R + (0 .. R-1).  MAXT is the largest workgroup
 * the instantiation accepts: 512 (up to two waves per SIMD, up to 256 VGPRs
 * a lane) or 1024 (16 waves, four per SIMD, 128 VGPRs a lane; what does not
 * fit is spilled, in practice in the code of the rare many-mechanism masks). */
template <class SET, bool MULTI, bool STAGE, int MAXT>
__global__ void __launch_bounds__(MAXT) __attribute__((amdgpu_waves_per_eu(MAXT > 512 ? 4 : NG_WAVES_PER_EU)))
ngpu_sim_kernel(SimArgs a)
{
    extern __shared__ double lds[];
    __shared__ unsigned qctr[NG_MAX_GROUPS];              /* the groups' segment counters */
    const int LW = a.LW, R = a.runs_per_group;
    const int G = blockDim.x / LW;
    const int group = __builtin_amdgcn_readfirstlane(threadIdx.x / LW);  /* LW is a multiple of 64 */
    const int g = threadIdx.x - group * LW;
    const Tables T = stage_tables<STAGE>(a, lds, G * R);
    __syncthreads();
    const int run0 = (blockIdx.x * G + group) * R;
    double *RG = T.runs + (size_t)group * R * ng_lds_per_run(a.N, STAGE ? a.nvar : 0, a.nrec);
    if constexpr (MULTI) sim_runs<SET, STAGE>(a, T, RG, run0, g, SyncBlock{}, qctr + group);
    else sim_runs<SET, STAGE>(a, T, RG, run0, g, SyncWave{}, qctr + group);
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

template <class SET>
__global__ void __launch_bounds__(NG_BATCH_THREADS) __attribute__((amdgpu_waves_per_eu(NG_WAVES_PER_EU)))
ngpu_batch_kernel(BatchArgs b)
{
    extern __shared__ double lds[];
    __shared__ int s_item;
    __shared__ unsigned qctr[NG_MAX_GROUPS];
    for (;;) {
        __syncthreads();                 /* the previous item is done with LDS */
        if (threadIdx.x == 0) s_item = atomicAdd(b.head, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= b.nitems) break;
        const int4 item = b.items[it];
        const int mi = __builtin_amdgcn_readfirstlane(item.x);
        const int run0 = __builtin_amdgcn_readfirstlane(item.y);
        const SimArgs &a = b.morph[mi];
        if (threadIdx.x == 0) b.stamps[2 * it] = wall_clock64();
        const int LW = a.LW, R = a.runs_per_group, G = NG_BATCH_THREADS / LW;
        const int group = __builtin_amdgcn_readfirstlane(threadIdx.x / LW);
        const int g = threadIdx.x - group * LW;
        const SyncDyn sync{a.block_sync != 0};
        /* groups beyond the item's runs simulate padding runs */
        if (a.lds_staging) {
            const Tables T = stage_tables<true>(a, lds, G * R);
            __syncthreads();
            double *RG = T.runs + (size_t)group * R * ng_lds_per_run(a.N, a.nvar, a.nrec);
            sim_runs<SET, true>(a, T, RG, run0 + group * R, g, sync, qctr + group);
        } else {
            const Tables T = stage_tables<false>(a, lds, G * R);
            __syncthreads();
            double *RG = T.runs + (size_t)group * R * ng_lds_per_run(a.N, 0, a.nrec);
            sim_runs<SET, false>(a, T, RG, run0 + group * R, g, sync, qctr + group);
        }
        __syncthreads();
        if (threadIdx.x == 0) b.stamps[2 * it + 1] = wall_clock64();
    }
}

/* Host-side launchers of one mechanism set's kernels (instantiated once per
 * set, in sim_<set>.hip).  multi: groups of more than one wave. */
template <class SET, bool MULTI, bool STAGE, int MAXT = 512>
static hipError_t ng_launch_sim(dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    auto k = ngpu_sim_kernel<SET, MULTI, STAGE, MAXT>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, grid, block, lds, s, a);
    return hipGetLastError();
}

template <class SET>
static hipError_t ng_launch_set(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a)
{
    if (block.x > 512) return a.lds_staging ? ng_launch_sim<SET, true, true, 1024>(grid, block, lds, s, a)
                                            : ng_launch_sim<SET, true, false, 1024>(grid, block, lds, s, a);
    if (multi) return a.lds_staging ? ng_launch_sim<SET, true, true>(grid, block, lds, s, a)
                                    : ng_launch_sim<SET, true, false>(grid, block, lds, s, a);
    return a.lds_staging ? ng_launch_sim<SET, false, true>(grid, block, lds, s, a)
                         : ng_launch_sim<SET, false, false>(grid, block, lds, s, a);
}

template <class SET>
static hipError_t ng_launch_batch_set(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b)
{
    auto k = ngpu_batch_kernel<SET>;
    if (lds > 65536) {
        hipError_t e = hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, grid, dim3(NG_BATCH_THREADS), lds, s, b);
    return hipGetLastError();
}

#endif
