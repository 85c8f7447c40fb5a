/*
 * ngpu_host.cpp -- host side of libneurogpu: context, model compilation for the
 * device layout, parameter compression, launches (optionally captured into a
 * hipGraph), timing, and the reference's data-file boundary.
 *
 * Replaces the reference's host runtime:
 *   initFrameWork / callKernel / stEfork2Main  CANON/CudaStuff.cu:372-587
 *   ReadParallelNeuronData                     CANON/MainC.cu:335-551
 *   ReadAllParams / ReadCSVStim / ReadSimData  CANON/Util.cu:210-432
 *   SaveArrayToFile                            CANON/Util.cu:443-456
 * (CANON = GUI/BBP_TTPC_EXAMPLE/pyNeuroGPU_unix/src)
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/neurogpu.h"
#include "sim_kernel.h"
#include "gen_ttpc.h"
#include "gen_mainen.h"
#include "schedule.h"
#include "treemap.h"

/* Per-set launchers, one translation unit each (sim_<set>.hip). */
hipError_t ngpu_launch_passive(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a);
hipError_t ngpu_launch_ttpc(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a);
hipError_t ngpu_launch_features(const double *vhot, int nrec, int nt, int nruns, int nstim, const double *dts,
                                const double *t_on, const double *t_off, double thr, double *out, hipStream_t s);
hipError_t ngpu_features_scratch(int device, double **ptr, size_t *cap);
hipError_t ngpu_launch_scatter_genomes(const double *genome, int nsets, int nfree, const int *src, int nvar, double *pv,
                                       hipStream_t stream);
hipError_t ngpu_launch_mainen(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a);
hipError_t ngpu_launch_batch_passive(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b);
hipError_t ngpu_launch_batch_ttpc(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b);
hipError_t ngpu_launch_batch_mainen(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b);
hipError_t ngpu_launch_ttpcgen(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a);
hipError_t ngpu_launch_batch_ttpcgen(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b);
hipError_t ngpu_launch_mainengen(bool multi, dim3 grid, dim3 block, size_t lds, hipStream_t s, const SimArgs &a);
hipError_t ngpu_launch_batch_mainengen(dim3 grid, size_t lds, hipStream_t s, const BatchArgs &b);

static std::string g_last_error;

/* Largest |e| or |f| of a tree that takes the reciprocal-form elimination by
 * default.  Generated morphologies stay below 100; the reference's exported
 * trees hold zero-area nodes (cm = 0) coupled with coefficients near 1e21,
 * where only the reference's own operation order reproduces its results. */
static const double kScaledCoef = 1e6;

/* First parameter index of each mechanism of a set, plus the total (the
 * parameter layout of the generated AllModels.cuh). */
static const int kPBaseTTPC[14] = {0, 1, 2, 6, 8, 9, 10, 11, 12, 13, 14, 16, 18, 19};
static const int kPBaseMainen[8] = {0, 2, 2, 6, 11, 16, 28, 30};
static const int kPBasePassive[2] = {0, 2};

/* The mechanism sets: counts and parameter layout (a generated set carries
 * its own PBASE, equal to the hand-written set's for the same model). */
struct SetInfo {
    int nmech, nparam, nstate;
    const int *pbase;
};
static const SetInfo *set_info(int ms)
{
    static const SetInfo passive{SetPassive::NMECH, SetPassive::NPARAM, SetPassive::NSTATE, kPBasePassive};
    static const SetInfo ttpc{SetTTPC::NMECH, SetTTPC::NPARAM, SetTTPC::NSTATE, kPBaseTTPC};
    static const SetInfo mainen{SetMainen::NMECH, SetMainen::NPARAM, SetMainen::NSTATE, kPBaseMainen};
    static const SetInfo ttpc_gen{SetGen_ttpc::NMECH, SetGen_ttpc::NPARAM, SetGen_ttpc::NSTATE, SetGen_ttpc::PBASE};
    static const SetInfo mainen_gen{SetGen_mainen::NMECH, SetGen_mainen::NPARAM, SetGen_mainen::NSTATE,
                                    SetGen_mainen::PBASE};
    switch (ms) {
    case NGPU_MECH_PASSIVE: return &passive;
    case NGPU_MECH_TTPC: return &ttpc;
    case NGPU_MECH_MAINEN: return &mainen;
    case NGPU_MECH_TTPC_GEN: return &ttpc_gen;
    case NGPU_MECH_MAINEN_GEN: return &mainen_gen;
    default: return nullptr;
    }
}
static int mech_nmodels(int ms) { return set_info(ms) ? set_info(ms)->nmech : -1; }
static int mech_nparams(int ms) { return set_info(ms) ? set_info(ms)->nparam : -1; }
static int mech_nstate(int ms) { return set_info(ms) ? set_info(ms)->nstate : SetPassive::NSTATE; }
static int param_owner(int ms, int j)
{
    const SetInfo *si = set_info(ms);
    if (!si) return -1;
    for (int m = 0; m < si->nmech; m++)
        if (j >= si->pbase[m] && j < si->pbase[m + 1]) return m;
    return -1;
}

template <class T> struct DevArr {
    T *p = nullptr;
    size_t n = 0;
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
};

struct ngpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    /* model */
    bool have_tree = false;
    int N = 0, ncomps = 0, mechset = 0, nmodels = 0, depth = 0, logn = 0;
    std::vector<double> e, f, cm;
    std::vector<int> par, son, seg_to_comp;
    std::vector<unsigned> mask;
    std::vector<int> chains, lvl_ptr, fath, flvl_ptr, fchild;
    /* sim */
    bool have_sim = false;
    double v_init = -65.0;
    std::vector<int> rec;
    /* params */
    bool have_params = false;
    int nsets = 0, nparams = 0, nvar = 0;
    std::vector<int> pref;        /* [nparams*N]: >=0 varying column, else -(uniform index)-1 */
    std::vector<double> puni, pv;
    int nfree = 0;                 /* > 0: parameters from a template and genome rows */
    bool pv_on_device = false;     /* d_pv written on the device (ngpu_set_genomes) */
    /* stimuli */
    bool have_stim = false;
    int nstim = 0, nt_amps = 0, nt = 0, stim_seg = -1;
    double area = 1.0;
    std::vector<double> amps, dts;
    /* launch configuration */
    int cfg_w = 0, cfg_r = 0, cfg_spl = 0, cfg_graph = 0;   /* waves and runs per group, 0 = auto */
    int cfg_stage = -1;            /* row coefficients: -1 auto, 0 global memory, 1 LDS */
    int cfg_solve = NGPU_SOLVE_AUTO;
    bool well_scaled = false;      /* every cm > 0 and |e|, |f| <= kScaledCoef: the reciprocal-form
                                      elimination is as accurate as the reference's order */
    /* device state */
    int lay_W = 0, lay_R = 0, lay_NQ = 0;
    int ncu = 0;                   /* compute units of the device, once queried */
    bool tables_dirty = true, params_dirty = true, stim_dirty = true, graph_dirty = true;
    DevArr<int> d_qlink, d_qmeta, d_chains, d_lvl, d_fath, d_flvl, d_fchild, d_rec, d_msched, d_gsrc, d_genq;
    int ngen = 0;                  /* run-time-mask positions (packed mechanism-pass segments) */
    std::vector<int> qmask;        /* mechanism mask by position */
    std::vector<int> msched;       /* mechanism-pass schedule of the last (W, R) */
    int msched_R = 0;              /* 0: no schedule uploaded */
    DevArr<double> d_qcm, d_ef, d_pval, d_pconst, d_pv, d_dts, d_amps, d_states, d_vhot, d_fwin, d_feat, d_genome;
    int stim_q = -1;
    std::vector<int> rec_q;
    unsigned pconst_mask = 0;
    std::vector<double> pconst;
    hipGraphExec_t gexec = nullptr;
    bool ran = false;
    int t_done = 0;               /* steps [0, t_done) simulated; the next window starts there */
    int last_launches = 0;
    int last_G = 1;               /* groups per workgroup of the last launch */
    int last_stage = 1;
};

static int fail(ngpu_ctx *c, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    g_last_error = buf;
    return code;
}

/* Host-side C++ exceptions (allocation failures, malformed input caught by
 * the standard library) must not cross the C ABI: the entry points that
 * allocate are function-try-blocks ending in NGPU_CATCH. */
#define NGPU_CATCH(ctx, code)                                                                     \
    catch (const std::bad_alloc &) { return fail((ctx), NGPU_E_NOMEM, "host allocation failed"); } \
    catch (const std::exception &ex_) { return fail((ctx), (code), "%s", ex_.what()); }

#define HIPCHK(c, call)                                                                           \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return fail((c), NGPU_E_HIP, "%s failed: %s", #call, hipGetErrorString(_e));          \
    } while (0)

template <class T>
static int upload(ngpu_ctx *c, DevArr<T> &d, const std::vector<T> &h)
{
    size_t n = std::max<size_t>(h.size(), 1);
    if (d.n != n) {
        d.release();
        HIPCHK(c, hipMalloc((void **)&d.p, n * sizeof(T)));
        d.n = n;
    }
    if (!h.empty()) HIPCHK(c, hipMemcpyAsync(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
    return 0;
}

template <class T>
static int alloc(ngpu_ctx *c, DevArr<T> &d, size_t n)
{
    n = std::max<size_t>(n, 1);
    if (d.n == n) return 0;
    d.release();
    hipError_t e = hipMalloc((void **)&d.p, n * sizeof(T));
    if (e != hipSuccess) return fail(c, NGPU_E_NOMEM, "device allocation of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    d.n = n;
    return 0;
}

/* Drops the captured launch sequence; every change that calls this (tree,
 * record sites, parameters, stimuli, launch shape) also invalidates the state
 * a windowed run carries between calls. */
static void drop_graph(ngpu_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->graph_dirty = true;
    c->t_done = 0;
}

extern "C" int ngpu_mechset_nparams(int ms) { return mech_nparams(ms) < 0 ? NGPU_E_ARG : mech_nparams(ms); }

/* Device part of a context: device, stream and events. */
static int init_device(ngpu_ctx *c, int device)
{
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return fail(c, NGPU_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return fail(c, NGPU_E_ARG, "device %d out of range (0..%d)", device, ndev - 1);
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess)
        return fail(c, NGPU_E_HIP, "could not create stream/events on device %d", device);
    return NGPU_OK;
}

extern "C" int ngpu_create(int device, ngpu_ctx **out)
{
    if (!out) return fail(nullptr, NGPU_E_ARG, "ngpu_create: out is NULL");
    *out = nullptr;
    ngpu_ctx *c = new ngpu_ctx();
    int rc = init_device(c, device);
    if (rc) {
        g_last_error = c->err;
        ngpu_close(c);
        return rc;
    }
    *out = c;
    return NGPU_OK;
}

extern "C" void ngpu_close(ngpu_ctx *c)
{
    if (!c) return;
    if (c->stream) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
    }
    drop_graph(c);
    for (auto *d : {&c->d_qlink, &c->d_qmeta, &c->d_chains, &c->d_lvl, &c->d_fath, &c->d_flvl, &c->d_fchild, &c->d_rec, &c->d_msched, &c->d_gsrc, &c->d_genq}) d->release();
    for (auto *d : {&c->d_qcm, &c->d_ef, &c->d_pval, &c->d_pconst, &c->d_pv, &c->d_dts, &c->d_amps, &c->d_states, &c->d_vhot, &c->d_fwin, &c->d_feat, &c->d_genome}) d->release();
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *ngpu_last_error(const ngpu_ctx *c) { return c ? c->err.c_str() : g_last_error.c_str(); }

/* ------------------------------------------------------------------ tree */

extern "C" int ngpu_set_tree(ngpu_ctx *c, const ngpu_tree *t)
try {
    if (!c || !t) return fail(c, NGPU_E_ARG, "ngpu_set_tree: NULL argument");
    const int N = t->N;
    if (N < 1 || t->ncomps < 1) return fail(c, NGPU_E_ARG, "N=%d and ncomps=%d must be positive", N, t->ncomps);
    if (N > 32000) return fail(c, NGPU_E_ARG, "N=%d exceeds the 32000-segment limit of the int16 schedule", N);
    if (mech_nmodels(t->mechset) < 0) return fail(c, NGPU_E_ARG, "unknown mechanism set %d", t->mechset);
    if (t->nmodels != mech_nmodels(t->mechset))
        return fail(c, NGPU_E_ARG, "mechanism set %d has %d mechanisms, tree gives NModels=%d", t->mechset, mech_nmodels(t->mechset), t->nmodels);
    if (!t->e || !t->f || !t->cm || !t->ks || !t->seg_to_comp || !t->bool_model || !t->son_no)
        return fail(c, NGPU_E_ARG, "ngpu_set_tree: missing array");
    if (t->depth < 0 || t->logn_depth < 0 || t->n_lrel != t->depth + 1 || t->n_flrel != t->depth)
        return fail(c, NGPU_E_ARG, "inconsistent Depth=%d nLRel=%d nFLRel=%d", t->depth, t->n_lrel, t->n_flrel);
    std::vector<int> par(N, -1), son(N);
    for (int s = 1; s < N; s++) {
        int r = N - s, kr = t->ks[r];
        if (kr <= r || kr > N) return fail(c, NGPU_E_ARG, "Ks[%d]=%d is not a row above %d", r, kr, r);
        par[s] = N - kr;
    }
    for (int s = 0; s < N; s++) {
        if (t->seg_to_comp[s] < 0 || t->seg_to_comp[s] >= t->ncomps) return fail(c, NGPU_E_ARG, "SegToComp[%d]=%d out of range", s, t->seg_to_comp[s]);
        son[s] = t->son_no[s];
    }
    std::vector<unsigned> mask(N, 0u);
    for (int m = 0; m < t->nmodels; m++)
        for (int s = 0; s < N; s++)
            if (t->bool_model[(size_t)m * N + s]) mask[s] |= 1u << m;
    /* Level schedule: unique runs per level, unique fathers per father-level. */
    std::vector<int> chains, lvl_ptr{0}, fath, flvl_ptr{0}, fchild;
    std::vector<int> elim(N, 0);
    auto row_ok = [&](int r) { return r >= 0 && r < N; };
    for (int L = 0; L <= t->depth; L++) {
        int lo = t->lrel_starts[L], hi = t->lrel_ends[L];
        if (lo < 0 || hi >= t->comp_depth || lo > hi) return fail(c, NGPU_E_ARG, "LRel range %d..%d invalid", lo, hi);
        std::vector<int> seen;
        for (int i = lo * 32; i < (hi + 1) * 32; i++) {
            int J = t->comp_by_level32[i] - 1;
            if (J < 0 || J > t->n_call) return fail(c, NGPU_E_ARG, "CompByLevel32 entry %d out of range", J + 1);
            if (std::find(seen.begin(), seen.end(), J) != seen.end()) continue;
            seen.push_back(J);
            int i0 = t->seg_start[J] - 1, i1 = t->seg_end[J];
            if (i0 < 1 || i1 > N) return fail(c, NGPU_E_ARG, "run %d spans rows %d..%d outside the tree", J, i0, i1);
            for (int r = i0; r < i1; r++) elim[r - 1]++;
            chains.push_back(i0);
            chains.push_back(i1);
        }
        lvl_ptr.push_back((int)chains.size() / 2);
        if (L < t->depth) {
            int flo = t->flrel_starts[L], fhi = t->flrel_ends[L];
            if (flo < 0 || fhi >= t->comp_fdepth || flo > fhi) return fail(c, NGPU_E_ARG, "FLRel range %d..%d invalid", flo, fhi);
            std::vector<int> fseen;
            for (int i = flo * 32; i < (fhi + 1) * 32; i++) {
                int B = t->comp_by_flevel32[i] - 1;
                if (B < 0 || B >= t->n_fathers) return fail(c, NGPU_E_ARG, "CompByFLevel32 entry %d out of range", B + 1);
                if (std::find(fseen.begin(), fseen.end(), B) != fseen.end()) continue;
                fseen.push_back(B);
                int J = t->fathers[B] - 1;
                if (!row_ok(J)) return fail(c, NGPU_E_ARG, "father row %d out of range", J + 1);
                fath.push_back(J);
                fath.push_back((int)fchild.size());
                for (int j = t->rel_starts[B]; j <= t->rel_ends[B]; j++) {
                    if (j < 1 || j > t->n_call) return fail(c, NGPU_E_ARG, "RelStarts/RelEnds out of range");
                    int tr = t->rel_vec[j - 1] - 2;
                    if (!row_ok(tr) || tr >= J) return fail(c, NGPU_E_ARG, "related row %d of father %d invalid", tr + 1, J + 1);
                    elim[tr]++;
                    fchild.push_back(tr);
                }
                fath.push_back((int)fchild.size());
            }
            flvl_ptr.push_back((int)fath.size() / 3);
        }
    }
    for (int r = 0; r < N - 1; r++)
        if (elim[r] != 1) return fail(c, NGPU_E_ARG, "tree maps eliminate row %d %d times (expected once)", r + 1, elim[r]);
    /* Chains become (i0, i1, tA, tB, pe): a father J folded after level L
     * whose row is the one the chain at level L+1 starts from (J = i0 - 1),
     * that no other chain of that level touches and that has at most two
     * related rows, is folded by that chain's lane just before the chain
     * (tA, tB: its related rows in order, -1 = none).  The same operations
     * in the same order, one barrier fewer per level.  Other fathers keep
     * their own phase.  pe is the parent row of the chain's last row (N for
     * the root), where the back-substitution of the chain starts. */
    {
        const int nlev = (int)lvl_ptr.size() - 1;
        std::vector<int> ch4(chains.size() / 2 * 5);
        for (size_t k = 0; k < chains.size() / 2; k++) {
            ch4[5 * k] = chains[2 * k]; ch4[5 * k + 1] = chains[2 * k + 1];
            ch4[5 * k + 2] = -1; ch4[5 * k + 3] = -1;
            const int end = chains[2 * k + 1] - 1;          /* row of the chain's last segment */
            ch4[5 * k + 4] = end == N - 1 ? N : N - 1 - par[N - 1 - end];
        }
        std::vector<int> rest, rest_ptr{0};
        for (int L = 0; L + 1 < nlev; L++) {
            const int k0 = lvl_ptr[L + 1], k1 = lvl_ptr[L + 2];
            for (int f = flvl_ptr[L]; f < flvl_ptr[L + 1]; f++) {
                const int J = fath[3 * f], c0 = fath[3 * f + 1], c1 = fath[3 * f + 2];
                int owner = -1, touching = 0;
                bool clash = false;
                for (int k = k0; k < k1; k++) {
                    const int i0 = chains[2 * k], i1 = chains[2 * k + 1];
                    if (J >= i0 - 1 && J < i1) { touching++; if (J == i0 - 1 && ch4[5 * k + 2] < 0) owner = k; }
                    for (int cc = c0; cc < c1; cc++)
                        if (fchild[cc] >= i0 && fchild[cc] < i1) clash = true;
                }
                if (c1 == c0) continue;                 /* nothing to fold */
                if (owner >= 0 && touching == 1 && !clash && c1 - c0 <= 2) {
                    ch4[5 * owner + 2] = fchild[c0];
                    ch4[5 * owner + 3] = c1 - c0 == 2 ? fchild[c0 + 1] : -1;
                } else {
                    rest.push_back(J); rest.push_back(c0); rest.push_back(c1);
                }
            }
            rest_ptr.push_back((int)rest.size() / 3);
        }
        chains.swap(ch4);
        fath.swap(rest);
        flvl_ptr.swap(rest_ptr);
    }
    /* The back-substitution walks each chain from its last row down to its
     * first: the chains' row ranges [i0-1, i1-1] must cover every row once,
     * each row's parent inside a chain must be the next row, and the parent
     * of a chain's last row must lie in a chain of a higher level. */
    {
        std::vector<int> lvl_of(N, -1);
        const int nlev = (int)lvl_ptr.size() - 1;
        for (int L = 0; L < nlev; L++)
            for (int k = lvl_ptr[L]; k < lvl_ptr[L + 1]; k++)
                for (int r = chains[5 * k] - 1; r < chains[5 * k + 1]; r++) {
                    if (lvl_of[r] >= 0) return fail(c, NGPU_E_ARG, "row %d lies in two segment runs", r + 1);
                    lvl_of[r] = L;
                }
        for (int r = 0; r < N; r++)
            if (lvl_of[r] < 0) return fail(c, NGPU_E_ARG, "row %d lies in no segment run", r + 1);
        for (int L = 0; L < nlev; L++)
            for (int k = lvl_ptr[L]; k < lvl_ptr[L + 1]; k++) {
                const int i0 = chains[5 * k], i1 = chains[5 * k + 1], pe = chains[5 * k + 4];
                for (int r = i0 - 1; r < i1 - 1; r++)
                    if (N - 1 - par[N - 1 - r] != r + 1)
                        return fail(c, NGPU_E_ARG, "segment run %d..%d is not a path to the root", i0, i1);
                if (pe < N && lvl_of[pe] <= L)
                    return fail(c, NGPU_E_ARG, "the parent of row %d is not on a higher level", i1);
            }
    }
    for (size_t i = 0; i < (size_t)t->logn_depth * N; i++) {
        int a = t->fidxs[i] - 1;
        if (a < 0 || a > N) return fail(c, NGPU_E_ARG, "FIdxs entry %d out of range", a + 1);
    }
    c->N = N; c->ncomps = t->ncomps; c->mechset = t->mechset; c->nmodels = t->nmodels;
    c->depth = t->depth; c->logn = t->logn_depth;
    c->e.assign(t->e, t->e + N); c->f.assign(t->f, t->f + N); c->cm.assign(t->cm, t->cm + N);
    c->par = par; c->son = son; c->seg_to_comp.assign(t->seg_to_comp, t->seg_to_comp + N); c->mask = mask;
    c->chains = chains; c->lvl_ptr = lvl_ptr; c->fath = fath; c->flvl_ptr = flvl_ptr; c->fchild = fchild;
    c->well_scaled = true;
    for (int i = 0; i < N; i++)
        if (!(t->cm[i] > 0) || !(std::fabs(t->e[i]) <= kScaledCoef) || !(std::fabs(t->f[i]) <= kScaledCoef))
            c->well_scaled = false;
    c->have_tree = true;
    c->have_params = false;          /* parameter layout depends on the tree */
    c->lay_W = 0; c->tables_dirty = true;
    drop_graph(c);
    return NGPU_OK;
}
NGPU_CATCH(c, NGPU_E_ARG)

extern "C" int ngpu_set_solve(ngpu_ctx *c, int mode)
{
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (mode != NGPU_SOLVE_AUTO && mode != NGPU_SOLVE_REFERENCE && mode != NGPU_SOLVE_FAST)
        return fail(c, NGPU_E_ARG, "solve mode %d is not NGPU_SOLVE_AUTO, _REFERENCE or _FAST", mode);
    c->cfg_solve = mode;
    drop_graph(c);
    return NGPU_OK;
}

/* The elimination a run of c uses: the reciprocal form where chosen, or
 * by default where the tree is well scaled. */
static bool fast_solve(const ngpu_ctx *c)
{
    return c->cfg_solve == NGPU_SOLVE_FAST || (c->cfg_solve == NGPU_SOLVE_AUTO && c->well_scaled);
}

extern "C" int ngpu_set_sim(ngpu_ctx *c, double v_init, int nrec, const int *rec_sites)
try {
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (!c->have_tree) return fail(c, NGPU_E_STATE, "ngpu_set_sim before ngpu_set_tree");
    if (nrec < 0 || (nrec > 0 && !rec_sites) || nrec > 64) return fail(c, NGPU_E_ARG, "nrec=%d invalid (0..64)", nrec);
    for (int r = 0; r < nrec; r++)
        if (rec_sites[r] < 0 || rec_sites[r] >= c->N) return fail(c, NGPU_E_ARG, "record site %d out of range", rec_sites[r]);
    if (!std::isfinite(v_init)) return fail(c, NGPU_E_ARG, "v_init must be finite");
    c->v_init = v_init;
    c->rec.assign(rec_sites, rec_sites + nrec);
    c->have_sim = true;
    c->tables_dirty = true;
    drop_graph(c);
    return NGPU_OK;
}
NGPU_CATCH(c, NGPU_E_ARG)

extern "C" int ngpu_set_params(ngpu_ctx *c, const double *P, int nsets, int ncomps, int nparams)
try {
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (!c->have_tree) return fail(c, NGPU_E_STATE, "ngpu_set_params before ngpu_set_tree");
    if (nsets < 1 || !P) return fail(c, NGPU_E_ARG, "nsets=%d must be positive", nsets);
    if (ncomps != c->ncomps) return fail(c, NGPU_E_ARG, "ncomps=%d but the tree has NComps=%d", ncomps, c->ncomps);
    if (nparams != mech_nparams(c->mechset)) return fail(c, NGPU_E_ARG, "nparams=%d but mechanism set %d has %d", nparams, c->mechset, mech_nparams(c->mechset));
    const int N = c->N;
    const size_t row = (size_t)nparams * ncomps;
    /* Column (p, comp) over sets: uniform columns become constants, varying
     * columns are de-duplicated and stored per set. */
    std::map<std::string, int> seen;
    std::vector<int> colref(row);
    std::vector<double> puni;
    std::vector<std::vector<double>> vcols;
    std::vector<double> col(nsets);
    for (size_t k = 0; k < row; k++) {
        bool uni = true;
        for (int s = 0; s < nsets; s++) {
            col[s] = P[(size_t)s * row + k];
            if (std::memcmp(&col[s], &col[0], sizeof(double)) != 0) uni = false;
        }
        if (uni) {
            colref[k] = -(int)puni.size() - 1;
            puni.push_back(col[0]);
        } else {
            std::string key((const char *)col.data(), col.size() * sizeof(double));
            auto it = seen.find(key);
            if (it == seen.end()) {
                it = seen.emplace(key, (int)vcols.size()).first;
                vcols.push_back(col);
            }
            colref[k] = it->second;
        }
    }
    c->nvar = (int)vcols.size();
    c->pv.assign((size_t)nsets * std::max(c->nvar, 1), 0.0);
    for (int k = 0; k < c->nvar; k++)
        for (int s = 0; s < nsets; s++) c->pv[(size_t)s * c->nvar + k] = vcols[k][s];
    c->pref.assign((size_t)nparams * N, 0);
    for (int j = 0; j < nparams; j++)
        for (int s = 0; s < N; s++) c->pref[(size_t)j * N + s] = colref[(size_t)j * ncomps + c->seg_to_comp[s]];
    c->puni = puni;
    c->nsets = nsets;
    c->nparams = nparams;
    c->nfree = 0;
    c->pv_on_device = false;
    c->have_params = true;
    c->params_dirty = true;
    c->tables_dirty = true;
    drop_graph(c);
    return NGPU_OK;
}
NGPU_CATCH(c, NGPU_E_ARG)

/* Sentinel of free parameter i in a template row (extractModel_mappings.py:258-270). */
static const double kSentinel = 12300.0;

extern "C" int ngpu_set_param_template(ngpu_ctx *c, const double *tpl, int ncomps, int nparams, int nfree)
try {
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (!c->have_tree) return fail(c, NGPU_E_STATE, "ngpu_set_param_template before ngpu_set_tree");
    if (!tpl) return fail(c, NGPU_E_ARG, "ngpu_set_param_template: NULL template");
    if (ncomps != c->ncomps) return fail(c, NGPU_E_ARG, "ncomps=%d but the tree has NComps=%d", ncomps, c->ncomps);
    if (nparams != mech_nparams(c->mechset)) return fail(c, NGPU_E_ARG, "nparams=%d but mechanism set %d has %d", nparams, c->mechset, mech_nparams(c->mechset));
    if (nfree < 1 || nfree > 4096) return fail(c, NGPU_E_ARG, "nfree=%d not in 1..4096", nfree);
    const int N = c->N;
    const size_t row = (size_t)nparams * ncomps;
    std::vector<int> colref(row), used(nfree, 0);
    std::vector<double> puni;
    for (size_t k = 0; k < row; k++) {
        const double x = tpl[k] - kSentinel;
        const int i = (int)x;
        if (x >= 0 && x < nfree && x == (double)i) {
            colref[k] = i;
            used[i] = 1;
        } else {
            colref[k] = -(int)puni.size() - 1;
            puni.push_back(tpl[k]);
        }
    }
    for (int i = 0; i < nfree; i++)
        if (!used[i]) return fail(c, NGPU_E_ARG, "free parameter %d has no position (no template entry holds %g)", i, kSentinel + i);
    c->pref.assign((size_t)nparams * N, 0);
    for (int j = 0; j < nparams; j++)
        for (int s = 0; s < N; s++) c->pref[(size_t)j * N + s] = colref[(size_t)j * ncomps + c->seg_to_comp[s]];
    c->puni = puni;
    c->nparams = nparams;
    c->nvar = nfree;
    c->nfree = nfree;
    c->have_params = false;        /* until genome rows are given */
    c->pv_on_device = false;
    c->tables_dirty = true;
    drop_graph(c);
    return NGPU_OK;
}
NGPU_CATCH(c, NGPU_E_ARG)

extern "C" int ngpu_set_genomes(ngpu_ctx *c, const double *x, int nsets)
try {
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (c->nfree < 1) return fail(c, NGPU_E_STATE, "ngpu_set_genomes needs ngpu_set_param_template first");
    if (!x || nsets < 1) return fail(c, NGPU_E_ARG, "ngpu_set_genomes: NULL rows or nsets=%d", nsets);
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<int> src(c->nvar);
    for (int k = 0; k < c->nvar; k++) src[k] = k;            /* varying column k is free parameter k */
    int rc;
    if ((rc = upload(c, c->d_gsrc, src)) || (rc = alloc(c, c->d_genome, (size_t)nsets * c->nfree)) ||
        (rc = alloc(c, c->d_pv, (size_t)nsets * c->nvar)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_genome.p, x, (size_t)nsets * c->nfree * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipError_t e = ngpu_launch_scatter_genomes(c->d_genome.p, nsets, c->nfree, c->d_gsrc.p, c->nvar, c->d_pv.p, c->stream);
    if (e != hipSuccess) return fail(c, NGPU_E_HIP, "genome scatter launch failed: %s", hipGetErrorString(e));
    if (nsets != c->nsets) drop_graph(c);
    c->nsets = nsets;
    c->pv_on_device = true;
    c->have_params = true;
    c->params_dirty = false;
    c->ran = false;
    c->t_done = 0;
    return NGPU_OK;
}
NGPU_CATCH(c, NGPU_E_ARG)

extern "C" int ngpu_set_stimuli(ngpu_ctx *c, const double *amps, int nstim, int nt_amps, const double *dts, int nt, int stim_seg, double area)
try {
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (!c->have_tree) return fail(c, NGPU_E_STATE, "ngpu_set_stimuli before ngpu_set_tree");
    if (nstim < 1 || nt < 1 || nt_amps < nt || !amps || !dts) return fail(c, NGPU_E_ARG, "invalid stimulus sizes nstim=%d nt=%d nt_amps=%d", nstim, nt, nt_amps);
    if (stim_seg < -1 || stim_seg >= c->N) return fail(c, NGPU_E_ARG, "stimulus segment %d out of range", stim_seg);
    if (!(area > 0)) return fail(c, NGPU_E_ARG, "stimulus area must be positive");
    for (int i = 0; i < nt; i++)
        if (!(dts[i] > 0) || !std::isfinite(dts[i])) return fail(c, NGPU_E_ARG, "dt[%d]=%g must be positive and finite", i, dts[i]);
    c->amps.assign(amps, amps + (size_t)nstim * nt_amps);
    c->dts.assign(dts, dts + nt);
    c->nstim = nstim; c->nt_amps = nt_amps; c->nt = nt; c->stim_seg = stim_seg; c->area = area;
    c->have_stim = true;
    c->stim_dirty = true;
    c->tables_dirty = true;
    drop_graph(c);
    return NGPU_OK;
}
NGPU_CATCH(c, NGPU_E_ARG)

extern "C" int ngpu_set_launch(ngpu_ctx *c, int w, int r, int spl, int graph)
{
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (!(w == 0 || w == 1 || w == 2 || w == 4 || w == 8 || w == 16) || r < 0 || r > 8 || spl < 0 ||
        (graph != 0 && graph != 1))
        return fail(c, NGPU_E_ARG, "invalid launch configuration w=%d r=%d spl=%d graph=%d", w, r, spl, graph);
    c->cfg_w = w; c->cfg_r = r; c->cfg_spl = spl; c->cfg_graph = graph;
    c->tables_dirty = true;
    drop_graph(c);
    return NGPU_OK;
}

extern "C" long long ngpu_vhot_size(const ngpu_ctx *c)
{
    if (!c || !c->have_params || !c->have_stim) return 0;
    return (long long)c->nsets * c->nstim * (long long)c->rec.size() * c->nt;
}

/* ------------------------------------------------------------------ run */

/* Positions per run: segments rounded up to whole waves. */
static int positions(int N) { return (N + 63) / 64 * 64; }

/* Most runs per group: four per one-wave group (on the 200-segment cell,
 * at equal occupancy, 9.2e9 compartment-steps/s against 8.7e9 with two and
 * 7.4e9 with one; the 500-segment cell fits two in LDS), six per larger
 * group (the 1,200-segment cell). */
static int max_runs_per_group(int W) { return W == 1 ? 4 : 6; }

static int build_tables(ngpu_ctx *c)
{
    const int N = c->N, NQ = positions(N);
    /* positions: segments sorted by mechanism mask, stable */
    std::vector<int> order(N);
    for (int s = 0; s < N; s++) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return c->mask[a] < c->mask[b]; });
    std::vector<int> pos(N), qlink(4 * (size_t)NQ, -1), qmeta(NQ, 0);
    std::vector<double> qcm(NQ, 0.0);
    for (int q = 0; q < N; q++) {
        const int s = order[q];
        pos[s] = q;
        qlink[4 * q] = N - 1 - s;                             /* row */
        qlink[4 * q + 1] = s == 0 ? N - 1 : N - 1 - c->par[s]; /* parent row (the root: itself) */
        qmeta[q] = (int)c->mask[s];
        qcm[q] = c->cm[s];
    }
    for (int q = N; q < NQ; q++) qlink[4 * q + 1] = 0;
    c->qmask = qmeta;
    c->msched_R = 0;
    /* run-time-mask positions: run by the packed segments of the mechanism
     * pass (ng_mech_schedule); bit 15 of the kernel's mask table marks them */
    const std::vector<int> genq = ng_generic_positions(qmeta, N, c->mechset);
    for (int q : genq) qmeta[q] |= 0x8000;
    /* sons in son order: the reference adds the children's terms to the
     * parent's row first for SonNoVec 1, then for SonNoVec 2
     * (CANON/CudaStuff.cu:296-305); other values add nothing */
    for (int s = 1; s < N; s++) {
        const int sn = c->son[s];
        if (sn != 1 && sn != 2) continue;
        int &slot = qlink[4 * (size_t)pos[c->par[s]] + 1 + sn];
        if (slot >= 0) return fail(c, NGPU_E_ARG, "segment %d has two children with SonNoVec %d", c->par[s], sn);
        slot = pos[s];
    }
    const int np = c->nparams;
    /* A parameter is a kernel constant when every segment that uses it holds
     * the same uniform value; otherwise the table holds the value, or a tagged
     * NaN naming the run-varying column. */
    c->pconst_mask = 0;
    c->pconst.assign(32, 0.0);
    for (int j = 0; j < np; j++) {
        int owner = param_owner(c->mechset, j);
        bool uni = true, any = false;
        double val = 0.0;
        for (int s = 0; s < N && uni; s++) {
            if (owner < 0 || !(c->mask[s] & (1u << owner))) continue;
            int ref = c->pref[(size_t)j * N + s];
            if (ref >= 0) { uni = false; break; }
            double x = c->puni[-ref - 1];
            if (!any) { val = x; any = true; }
            else if (std::memcmp(&x, &val, sizeof x) != 0) uni = false;
        }
        if (uni) { c->pconst_mask |= 1u << j; c->pconst[j] = val; }
    }
    std::vector<double> pval((size_t)np * NQ, 0.0);
    for (int j = 0; j < np; j++)
        for (int q = 0; q < N; q++) {
            int ref = c->pref[(size_t)j * N + order[q]];
            double x;
            if (ref >= 0) {
                unsigned long long bits = ((unsigned long long)NG_PTAG << 32) | (unsigned)ref;
                std::memcpy(&x, &bits, sizeof x);
            } else {
                x = c->puni[-ref - 1];
            }
            pval[(size_t)j * NQ + q] = x;
        }
    c->rec_q.resize(c->rec.size());
    for (size_t r = 0; r < c->rec.size(); r++) c->rec_q[r] = pos[c->rec[r]];
    c->stim_q = c->stim_seg >= 0 ? pos[c->stim_seg] : -1;
    std::vector<double> ef(2 * (size_t)N + NG_ROW_PAD, 0.0); /* e, f interleaved by Hines row, then
                                                               eight rows of zeros (read past a
                                                               chain's end, not used) */
    for (int h = 0; h < N; h++) { ef[2 * h] = c->e[h]; ef[2 * h + 1] = c->f[h]; }
    int rc;
    if ((rc = upload(c, c->d_pconst, c->pconst)) || (rc = upload(c, c->d_qlink, qlink)) || (rc = upload(c, c->d_qmeta, qmeta)) ||
        (rc = upload(c, c->d_qcm, qcm)) || (rc = upload(c, c->d_ef, ef)) ||
        (rc = upload(c, c->d_pval, pval)) || (rc = upload(c, c->d_chains, c->chains)) ||
        (rc = upload(c, c->d_lvl, c->lvl_ptr)) || (rc = upload(c, c->d_fath, c->fath)) || (rc = upload(c, c->d_flvl, c->flvl_ptr)) ||
        (rc = upload(c, c->d_fchild, c->fchild)) || (rc = upload(c, c->d_rec, c->rec_q)) ||
        (rc = upload(c, c->d_genq, genq)))
        return rc;
    c->ngen = (int)genq.size();
    c->lay_NQ = NQ;
    c->tables_dirty = false;
    return NGPU_OK;
}

static hipError_t launch_one(ngpu_ctx *c, const SimArgs &a, dim3 grid, dim3 block, size_t lds)
{
    bool multi = c->lay_W > 1;
    switch (c->mechset) {
    case NGPU_MECH_TTPC: return ngpu_launch_ttpc(multi, grid, block, lds, c->stream, a);
    case NGPU_MECH_MAINEN: return ngpu_launch_mainen(multi, grid, block, lds, c->stream, a);
    case NGPU_MECH_TTPC_GEN: return ngpu_launch_ttpcgen(multi, grid, block, lds, c->stream, a);
    case NGPU_MECH_MAINEN_GEN: return ngpu_launch_mainengen(multi, grid, block, lds, c->stream, a);
    default: return ngpu_launch_passive(multi, grid, block, lds, c->stream, a);
    }
}

/* LDS a workgroup may request: the CU's 160 KiB less the kernels' static
 * part (the groups' segment counters and the batch kernel's item slot). */
static const size_t kLdsLimit = 160 * 1024 - 128;

/* LDS staging (row coefficients and the run's varying parameter values) is
 * considered for up to this many varying values per run. */
static const int kStageMaxVar = 512;

static size_t lds_need(const ngpu_ctx *c, int runs, int staged)
{
    return ng_lds_bytes(c->N, positions(c->N), runs, std::max(c->nvar, 1), (int)c->rec.size(), staged,
                        (int)c->chains.size() / 5, (int)c->fath.size() / 3, (int)c->fchild.size(), c->depth);
}

/* Runs per group R (0 = as many as fit, up to max_runs_per_group), groups
 * per workgroup G (as many as fit in LDS, at most 8 waves) and LDS staging
 * (-1 = when it fits) for groups of W waves. */
static int plan(ngpu_ctx *c, int W, int &R, int &G, int &stage)
{
    if (stage < 0) stage = c->nvar <= kStageMaxVar && lds_need(c, std::max(R, 1), 1) <= kLdsLimit ? 1 : 0;
    if (R <= 0) {
        R = 1;
        while (R < max_runs_per_group(W) && lds_need(c, R + 1, stage) <= kLdsLimit) R++;
    }
    if (lds_need(c, R, stage) > kLdsLimit)
        return fail(c, NGPU_E_ARG, "%d runs of a %d-segment tree need %zu bytes of LDS per workgroup (limit 160 KiB)", R,
                    c->N, lds_need(c, R, stage));
    G = 1;
    while ((G + 1) * W <= 8 && lds_need(c, (G + 1) * R, stage) <= kLdsLimit) G++;
    return NGPU_OK;
}

/* Kernel arguments of a prepared context. */
static void fill_args(const ngpu_ctx *c, SimArgs &a, int W, int R, int staged, int block_sync)
{
    a = SimArgs{};
    a.N = c->N; a.NQ = c->lay_NQ; a.LW = 64 * W; a.runs_per_group = R;
    a.nvar = std::max(c->nvar, 1); a.nrec = (int)c->rec.size(); a.nstim = c->nstim; a.nt = c->nt; a.nt_amps = c->nt_amps;
    a.nruns = c->nsets * c->nstim; a.depth = c->depth; a.stim_q = c->stim_q;
    a.stim_area = c->area; a.v_init = c->v_init;
    a.q_meta = c->d_qmeta.p; a.q_link = c->d_qlink.p; a.q_cm = c->d_qcm.p;
    a.row_ef = c->d_ef.p; a.pval = c->d_pval.p; a.pv = c->d_pv.p;
    a.pconst_mask = c->pconst_mask;
    a.pconst = c->d_pconst.p;
    a.msched = c->d_msched.p;
    a.gen_q = c->d_genq.p; a.ngen = c->ngen;
    a.chains = c->d_chains.p; a.lvl_ptr = c->d_lvl.p; a.fath = c->d_fath.p; a.flvl_ptr = c->d_flvl.p;
    a.fchild = c->d_fchild.p; a.lds_staging = staged; a.block_sync = block_sync;
    a.fast_solve = fast_solve(c) ? 1 : 0;
    /* the tail of the level schedule: levels with one chain and no fathers
     * of their own, from the top down */
    a.tail_level = c->depth + 1;
    for (int L = c->depth; L >= 0; L--) {
        const bool one = c->lvl_ptr[L + 1] - c->lvl_ptr[L] == 1;
        const bool nof = L == 0 || c->flvl_ptr[L] == c->flvl_ptr[L - 1];
        if (!(one && nof)) break;
        a.tail_level = L;
    }
    a.nchains = (int)c->chains.size() / 5; a.nfath = (int)c->fath.size() / 3; a.nfchild = (int)c->fchild.size();
    a.dts = c->d_dts.p; a.amps = c->d_amps.p; a.rec_q = c->d_rec.p;
    a.states = c->d_states.p; a.vhot = c->d_vhot.p;
    a.t0 = 0; a.t1 = c->nt;
}

static int enqueue(ngpu_ctx *c, int t_begin, int t_end, int &launches)
{
    const int W = c->lay_W, nruns = c->nsets * c->nstim;
    int R = c->lay_R, G = 1, stage = c->cfg_stage;
    int rc = plan(c, W, R, G, stage);
    if (rc) return rc;
    const size_t lds = lds_need(c, G * R, stage);
    SimArgs a;
    fill_args(c, a, W, R, stage, W > 1);
    c->last_G = G;
    c->last_stage = stage;
    dim3 block(64 * W * G), grid((nruns + G * R - 1) / (G * R));
    int spl = c->cfg_spl > 0 ? c->cfg_spl : c->nt;
    launches = 0;
    for (int t0 = t_begin; t0 < t_end; t0 += spl) {
        a.t0 = t0;
        a.t1 = std::min(t_end, t0 + spl);
        hipError_t e = launch_one(c, a, grid, block, lds);
        if (e != hipSuccess) return fail(c, NGPU_E_HIP, "kernel launch failed: %s", hipGetErrorString(e));
        launches++;
    }
    return NGPU_OK;
}

static int prepare_layout(ngpu_ctx *c, int W, int R);

/* Device, default record site and the morphology's tables, once. */
static int ensure_tables(ngpu_ctx *c)
{
    if (!c->have_sim) { c->rec.assign(1, 0); c->have_sim = true; }
    HIPCHK(c, hipSetDevice(c->device));
    if (c->tables_dirty) {
        int rc = build_tables(c);
        if (rc) return rc;
        drop_graph(c);
    }
    return NGPU_OK;
}

/* The mechanism pass's segment queue for groups of R runs
 * (ng_mech_schedule, schedule.h), uploaded once per R. */
static int mech_schedule(ngpu_ctx *c, int R)
{
    if (c->msched_R == R) return NGPU_OK;
    c->msched = ng_mech_schedule(c->qmask, c->N, c->lay_NQ, c->mechset, R);
    int rc = upload(c, c->d_msched, c->msched);
    if (rc) return rc;
    c->msched_R = R;
    return NGPU_OK;
}

/* Groups per workgroup for groups of W waves and R runs (as plan() chooses
 * them), 0 if R runs do not fit in LDS; stage as plan() resolves it. */
static int fit_groups(const ngpu_ctx *c, int W, int R, int stage)
{
    if (stage < 0) stage = c->nvar <= kStageMaxVar && lds_need(c, R, 1) <= kLdsLimit ? 1 : 0;
    if (lds_need(c, R, stage) > kLdsLimit) return 0;
    int G = 1;
    while ((G + 1) * W <= 8 && lds_need(c, (G + 1) * R, stage) <= kLdsLimit) G++;
    return G;
}

/* Waves W and runs R per group where the caller left them at 0
 * (ng_choose_shape, schedule.h): candidate waves 1, 2 and 4 up to 512
 * segments, 16 above, or the caller's W; runs 1 up to max_runs_per_group, or
 * the caller's R. */
static void choose_layout(ngpu_ctx *c, int &W, int &R)
{
    if (c->cfg_w > 0 && c->cfg_r > 0) { W = c->cfg_w; R = c->cfg_r; return; }
    if (c->ncu <= 0 && hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
        c->ncu = 0;
    std::vector<int> ws;
    if (c->cfg_w > 0) ws.push_back(c->cfg_w);
    else if (c->N <= 512) ws = {1, 2, 4};
    else ws.push_back(16);
    ng_choose_shape((long)c->nsets * c->nstim, c->ncu, ws, c->cfg_r,
                    [&](int w, int r) { return fit_groups(c, w, r, c->cfg_stage); }, max_runs_per_group, W, R);
    if (R == 0) R = c->cfg_r;            /* nothing fits: plan() reports it */
}

static int prepare(ngpu_ctx *c)
{
    if (!c->have_tree || !c->have_params || !c->have_stim)
        return fail(c, NGPU_E_STATE, "ngpu_run needs ngpu_set_tree, ngpu_set_params and ngpu_set_stimuli first");
    /* the layout choice reads the tables' sizes */
    int rc = ensure_tables(c);
    if (rc) return rc;
    int W, R;
    choose_layout(c, W, R);
    return prepare_layout(c, W, R);
}

/* Padding runs beyond nruns: a workgroup always simulates whole groups. */
static const size_t kPadRuns = 64;

/* Tables, parameters, stimuli and buffers on the device for groups of W waves
 * and R runs (0 = as many as fit). */
static int prepare_layout(ngpu_ctx *c, int W, int R)
{
    if (!c->have_tree || !c->have_params || !c->have_stim)
        return fail(c, NGPU_E_STATE, "a run needs ngpu_set_tree, ngpu_set_params and ngpu_set_stimuli first");
    int rc = ensure_tables(c);
    if (rc) return rc;
    if (W != c->lay_W || R != c->lay_R) drop_graph(c);
    c->lay_W = W; c->lay_R = R;
    if (c->params_dirty && !c->pv_on_device) {
        if ((rc = upload(c, c->d_pv, c->pv))) return rc;
        c->params_dirty = false;
    }
    if (c->stim_dirty) {
        if ((rc = upload(c, c->d_dts, c->dts)) || (rc = upload(c, c->d_amps, c->amps))) return rc;
        c->stim_dirty = false;
    }
    const size_t nruns = (size_t)c->nsets * c->nstim + kPadRuns, NQ = c->lay_NQ;
    if ((rc = alloc(c, c->d_states, nruns * (mech_nstate(c->mechset) + 2) * NQ)) ||
        (rc = alloc(c, c->d_vhot, nruns * c->rec.size() * c->nt)))
        return rc;
    /* the schedule of the runs per group a launch will use (a shape that does
     * not fit is reported by the launch) */
    int Rp = R, G = 1, stage = c->cfg_stage;
    if (plan(c, W, Rp, G, stage) == NGPU_OK && (rc = mech_schedule(c, Rp))) return rc;
    return NGPU_OK;
}

/* Steps [t_begin, t_end) of every run (t_begin 0: from the initial state;
 * otherwise from the state the previous call left at t_begin).  Whole runs
 * may replay a captured hipGraph; windows are launched directly. */
static int run_range(ngpu_ctx *c, int t_begin, int t_end, double *vhot, ngpu_stats *st)
{
    hipEvent_t h0, h1;
    int rc = prepare(c);
    if (rc) return rc;
    if (t_begin < 0 || t_end > c->nt || t_begin >= t_end)
        return fail(c, NGPU_E_ARG, "step window [%d, %d) is not inside [0, %d)", t_begin, t_end, c->nt);
    if (t_begin != 0 && t_begin != c->t_done)
        return fail(c, NGPU_E_STATE, "a window starting at step %d needs the state at step %d (the last run ended at %d)",
                    t_begin, t_begin, c->t_done);
    const bool whole = t_begin == 0 && t_end == c->nt;
    HIPCHK(c, hipEventCreate(&h0));
    HIPCHK(c, hipEventCreate(&h1));
    HIPCHK(c, hipEventRecord(h0, c->stream));
    int launches = 0;
    if (c->cfg_graph && whole) {
        if (!c->gexec) {
            hipGraph_t g;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue(c, 0, c->nt, launches);
            hipError_t ce = hipStreamEndCapture(c->stream, &g);
            if (rc) return rc;
            if (ce != hipSuccess) return fail(c, NGPU_E_HIP, "graph capture failed: %s", hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) return fail(c, NGPU_E_HIP, "graph instantiation failed: %s", hipGetErrorString(ie));
            c->last_launches = launches;
        }
        launches = c->last_launches;
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    } else {
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
        if ((rc = enqueue(c, t_begin, t_end, launches))) return rc;
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    }
    hipError_t se = hipStreamSynchronize(c->stream);
    if (se != hipSuccess) return fail(c, NGPU_E_HIP, "simulation failed: %s", hipGetErrorString(se));
    c->t_done = t_end;
    c->ran = t_end == c->nt;
    if (vhot) {
        HIPCHK(c, hipMemcpy(vhot, c->d_vhot.p, (size_t)ngpu_vhot_size(c) * sizeof(double), hipMemcpyDeviceToHost));
    }
    HIPCHK(c, hipEventRecord(h1, c->stream));
    HIPCHK(c, hipEventSynchronize(h1));
    if (st) {
        float mk = 0, mt = 0;
        HIPCHK(c, hipEventElapsedTime(&mk, c->ev0, c->ev1));
        HIPCHK(c, hipEventElapsedTime(&mt, h0, h1));
        st->ms_kernel = mk;
        st->ms_total = mt;
        st->comp_steps = (double)c->nsets * c->nstim * (double)c->N * (t_end - t_begin);
        st->comp_steps_per_s = mk > 0 ? st->comp_steps / (mk * 1e-3) : 0.0;
        int R = c->lay_R, G = 1, stage = c->cfg_stage;
        (void)plan(c, c->lay_W, R, G, stage);
        st->waves_per_run = c->lay_W;
        st->slots_per_lane = (R * c->lay_NQ + 64 * c->lay_W - 1) / (64 * c->lay_W);
        st->runs_per_block = c->last_G * R;
        st->runs_per_group = R;
        st->launches = launches;
        st->nvar = c->nvar;
        st->nsets = c->nsets; st->nstim = c->nstim; st->nrec = (int)c->rec.size(); st->nt = c->nt; st->N = c->N;
        st->solve = fast_solve(c) ? NGPU_SOLVE_FAST : NGPU_SOLVE_REFERENCE;
    }
    (void)hipEventDestroy(h0);
    (void)hipEventDestroy(h1);
    return NGPU_OK;
}

extern "C" int ngpu_run(ngpu_ctx *c, double *vhot, ngpu_stats *st)
try {
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (!c->have_stim) return fail(c, NGPU_E_STATE, "ngpu_run needs ngpu_set_tree, ngpu_set_params and ngpu_set_stimuli first");
    return run_range(c, 0, c->nt, vhot, st);
}
NGPU_CATCH(c, NGPU_E_ARG)

extern "C" int ngpu_run_steps(ngpu_ctx *c, int t_begin, int t_end, ngpu_stats *st)
try {
    if (!c) return fail(c, NGPU_E_ARG, "NULL context");
    if (!c->have_stim) return fail(c, NGPU_E_STATE, "ngpu_run_steps needs ngpu_set_tree, ngpu_set_params and ngpu_set_stimuli first");
    return run_range(c, t_begin, t_end, nullptr, st);
}
NGPU_CATCH(c, NGPU_E_ARG)

extern "C" int ngpu_get_sizes(const ngpu_ctx *c, int *nsets, int *nstim, int *nrec, int *nt, int *N)
{
    if (!c || !nsets || !nstim || !nrec || !nt || !N) return fail(nullptr, NGPU_E_ARG, "ngpu_get_sizes: NULL argument");
    *nsets = c->have_params ? c->nsets : 0;
    *nstim = c->have_stim ? c->nstim : 0;
    *nrec = c->have_sim ? (int)c->rec.size() : (c->have_tree ? 1 : 0);
    *nt = c->have_stim ? c->nt : 0;
    *N = c->have_tree ? c->N : 0;
    return NGPU_OK;
}

/* ------------------------------------------------------------------ batches */

struct ngpu_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    std::vector<ngpu_ctx *> mem;
    std::vector<ngpu_batch_launch> cfg;
    std::vector<int> lay_w, lay_r, lay_sync, lay_stage;
    std::vector<double> t_first, t_last, cs;   /* per member, ms from the first item start */
    int4 *d_items = nullptr;
    int *d_head = nullptr;
    unsigned long long *d_stamps = nullptr;
    size_t cap_items = 0;
};

static int bfail(ngpu_batch *b, int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    g_last_error = buf;
    return code;
}

#define BHIPCHK(b, call)                                                                          \
    do {                                                                                          \
        hipError_t _e = (call);                                                                   \
        if (_e != hipSuccess)                                                                     \
            return bfail((b), NGPU_E_HIP, "%s failed: %s", #call, hipGetErrorString(_e));         \
    } while (0)

extern "C" int ngpu_batch_create(int device, ngpu_batch **out)
{
    if (!out) return bfail(nullptr, NGPU_E_ARG, "ngpu_batch_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return bfail(nullptr, NGPU_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return bfail(nullptr, NGPU_E_ARG, "device %d out of range (0..%d)", device, ndev - 1);
    ngpu_batch *b = new ngpu_batch();
    b->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
        hipMalloc((void **)&b->d_head, sizeof(int)) != hipSuccess) {
        ngpu_batch_close(b);
        return bfail(nullptr, NGPU_E_HIP, "could not create the batch's stream, events or counter on device %d", device);
    }
    *out = b;
    return NGPU_OK;
}

extern "C" void ngpu_batch_close(ngpu_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void *p : {(void *)b->d_items, (void *)b->d_head, (void *)b->d_stamps})
        if (p) (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" const char *ngpu_batch_last_error(const ngpu_batch *b) { return b ? b->err.c_str() : g_last_error.c_str(); }

extern "C" int ngpu_batch_add(ngpu_batch *b, ngpu_ctx *c)
{
    if (!b || !c) return bfail(b, NGPU_E_ARG, "ngpu_batch_add: NULL argument");
    if (c->device != b->device) return bfail(b, NGPU_E_ARG, "member on device %d, batch on device %d", c->device, b->device);
    if (!c->have_tree) return bfail(b, NGPU_E_STATE, "member has no tree");
    if ((int)b->mem.size() >= NG_BATCH_MAX_MORPH)
        return bfail(b, NGPU_E_ARG, "a batch holds at most %d contexts", NG_BATCH_MAX_MORPH);
    if (!b->mem.empty() && c->mechset != b->mem[0]->mechset)
        return bfail(b, NGPU_E_ARG, "all members of a batch need the same mechanism set");
    for (auto *m : b->mem)
        if (m == c) return bfail(b, NGPU_E_ARG, "context already in the batch");
    b->mem.push_back(c);
    b->cfg.push_back(ngpu_batch_launch{0, -1, -1, 0});
    return (int)b->mem.size() - 1;
}

extern "C" int ngpu_batch_set_launch(ngpu_batch *b, int member, const ngpu_batch_launch *cfg)
{
    if (!b || !cfg) return bfail(b, NGPU_E_ARG, "ngpu_batch_set_launch: NULL argument");
    if (member < 0 || member >= (int)b->mem.size()) return bfail(b, NGPU_E_ARG, "member %d out of range", member);
    const int w = cfg->waves_per_run;
    if (!(w == 0 || w == 1 || w == 2 || w == 4 || w == 8)) return bfail(b, NGPU_E_ARG, "waves_per_run %d not in {0,1,2,4,8}", w);
    if (cfg->block_sync < -1 || cfg->block_sync > 1 || cfg->lds_staging < -1 || cfg->lds_staging > 1)
        return bfail(b, NGPU_E_ARG, "block_sync and lds_staging take -1, 0 or 1");
    if (w > 1 && cfg->block_sync == 0) return bfail(b, NGPU_E_ARG, "runs of several waves synchronise at workgroup scope");
    if (cfg->runs_per_group < 0 || cfg->runs_per_group > 8) return bfail(b, NGPU_E_ARG, "runs_per_group %d not in 0..8", cfg->runs_per_group);
    b->cfg[member] = *cfg;
    return NGPU_OK;
}

extern "C" int ngpu_batch_run(ngpu_batch *b, ngpu_batch_stats *st)
try {
    if (!b) return bfail(b, NGPU_E_ARG, "NULL batch");
    if (b->mem.empty()) return bfail(b, NGPU_E_STATE, "empty batch");
    BHIPCHK(b, hipSetDevice(b->device));
    const int nm = (int)b->mem.size();
    std::vector<SimArgs> args(nm);
    std::vector<int4> items;
    std::vector<double> cost;
    size_t lds = 0;
    b->lay_w.assign(nm, 0); b->lay_r.assign(nm, 0); b->lay_sync.assign(nm, 0); b->lay_stage.assign(nm, 0);
    for (int i = 0; i < nm; i++) {
        ngpu_ctx *c = b->mem[i];
        if (!c->have_params || !c->have_stim) return bfail(b, NGPU_E_STATE, "member %d lacks parameters or stimuli", i);
        int W = b->cfg[i].waves_per_run;
        if (W == 0) W = c->N <= 512 ? 1 : 8;
        int R = b->cfg[i].runs_per_group;
        int rc = prepare_layout(c, W, R);
        if (rc) return bfail(b, rc, "member %d: %s", i, c->err.c_str());
        /* the workgroup is 8 waves: G = 8 / W groups of R runs */
        const int G = 8 / W;
        int stage = b->cfg[i].lds_staging;
        if (stage < 0) stage = c->nvar <= kStageMaxVar && lds_need(c, G * std::max(R, 1), 1) <= kLdsLimit ? 1 : 0;
        if (R <= 0) {
            R = 1;
            while (R < max_runs_per_group(W) && lds_need(c, G * (R + 1), stage) <= kLdsLimit) R++;
        }
        int sync = b->cfg[i].block_sync;
        if (sync < 0) sync = W > 1 ? 1 : 0;
        const size_t need = lds_need(c, G * R, stage);
        if (need > kLdsLimit) return bfail(b, NGPU_E_ARG, "member %d needs %zu bytes of LDS per workgroup (limit 160 KiB)", i, need);
        lds = std::max(lds, need);
        if ((rc = mech_schedule(c, R))) return bfail(b, rc, "member %d: %s", i, c->err.c_str());
        fill_args(c, args[i], W, R, stage, sync);
        b->lay_w[i] = W; b->lay_r[i] = R; b->lay_sync[i] = sync; b->lay_stage[i] = stage;
        const int nruns = c->nsets * c->nstim, per_item = G * R;
        for (int r0 = 0; r0 < nruns; r0 += per_item) {
            items.push_back(make_int4(i, r0, std::min(per_item, nruns - r0), 0));
            cost.push_back((double)c->N * c->nt * std::min(per_item, nruns - r0));
        }
    }
    /* most work first (segments x steps x runs of the item), so that the
     * last items to start are the shortest and no CU idles long at the end */
    std::vector<int> order(items.size());
    for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<int4> sorted(items.size());
    for (size_t k = 0; k < order.size(); k++) sorted[k] = items[order[k]];
    if (b->cap_items < sorted.size()) {
        if (b->d_items) (void)hipFree(b->d_items);
        if (b->d_stamps) (void)hipFree(b->d_stamps);
        b->d_items = nullptr; b->d_stamps = nullptr; b->cap_items = 0;
        BHIPCHK(b, hipMalloc((void **)&b->d_items, sorted.size() * sizeof(int4)));
        BHIPCHK(b, hipMalloc((void **)&b->d_stamps, sorted.size() * 2 * sizeof(unsigned long long)));
        b->cap_items = sorted.size();
    }
    /* the members' uploads were queued on their own streams */
    for (auto *c : b->mem) BHIPCHK(b, hipStreamSynchronize(c->stream));
    BHIPCHK(b, hipMemcpyAsync(b->d_items, sorted.data(), sorted.size() * sizeof(int4), hipMemcpyHostToDevice, b->stream));
    BHIPCHK(b, hipMemsetAsync(b->d_head, 0, sizeof(int), b->stream));
    int ncu = 0;
    BHIPCHK(b, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
    BatchArgs ba{};
    for (int i = 0; i < nm; i++) ba.morph[i] = args[i];
    ba.items = b->d_items; ba.nitems = (int)sorted.size(); ba.head = b->d_head; ba.stamps = b->d_stamps;
    const int grid = std::max(1, std::min(ncu, (int)sorted.size()));
    BHIPCHK(b, hipEventRecord(b->ev0, b->stream));
    hipError_t le;
    switch (b->mem[0]->mechset) {
    case NGPU_MECH_TTPC: le = ngpu_launch_batch_ttpc(dim3(grid), lds, b->stream, ba); break;
    case NGPU_MECH_MAINEN: le = ngpu_launch_batch_mainen(dim3(grid), lds, b->stream, ba); break;
    case NGPU_MECH_TTPC_GEN: le = ngpu_launch_batch_ttpcgen(dim3(grid), lds, b->stream, ba); break;
    case NGPU_MECH_MAINEN_GEN: le = ngpu_launch_batch_mainengen(dim3(grid), lds, b->stream, ba); break;
    default: le = ngpu_launch_batch_passive(dim3(grid), lds, b->stream, ba); break;
    }
    if (le != hipSuccess) return bfail(b, NGPU_E_HIP, "batch kernel launch failed: %s", hipGetErrorString(le));
    BHIPCHK(b, hipEventRecord(b->ev1, b->stream));
    hipError_t se = hipStreamSynchronize(b->stream);
    if (se != hipSuccess) return bfail(b, NGPU_E_HIP, "batch simulation failed: %s", hipGetErrorString(se));
    float ms = 0.f;
    BHIPCHK(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
    std::vector<unsigned long long> stamps(sorted.size() * 2);
    BHIPCHK(b, hipMemcpy(stamps.data(), b->d_stamps, stamps.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    int khz = 0;
    BHIPCHK(b, hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
    unsigned long long t0 = ~0ull;
    for (size_t k = 0; k < sorted.size(); k++) t0 = std::min(t0, stamps[2 * k]);
    b->t_first.assign(nm, 1e300); b->t_last.assign(nm, 0.0); b->cs.assign(nm, 0.0);
    double total = 0;
    for (size_t k = 0; k < sorted.size(); k++) {
        const int i = sorted[k].x;
        const double s0 = (double)(stamps[2 * k] - t0) / khz, s1 = (double)(stamps[2 * k + 1] - t0) / khz;
        b->t_first[i] = std::min(b->t_first[i], s0);
        b->t_last[i] = std::max(b->t_last[i], s1);
        const double cs = (double)b->mem[i]->N * b->mem[i]->nt * sorted[k].z;
        b->cs[i] += cs;
        total += cs;
    }
    for (auto *c : b->mem) c->ran = true;
    if (st) {
        st->ms_kernel = ms;
        st->comp_steps = total;
        st->nitems = (int)sorted.size();
        st->workgroups = grid;
        st->lds_bytes = (int)lds;
    }
    return NGPU_OK;
}
catch (const std::bad_alloc &) { return bfail(b, NGPU_E_NOMEM, "host allocation failed"); }
catch (const std::exception &ex_) { return bfail(b, NGPU_E_ARG, "%s", ex_.what()); }

extern "C" int ngpu_batch_member_info(const ngpu_batch *b, int member, ngpu_batch_member *out)
{
    if (!b || !out) return bfail(nullptr, NGPU_E_ARG, "ngpu_batch_member_info: NULL argument");
    if (member < 0 || member >= (int)b->mem.size()) return bfail(nullptr, NGPU_E_ARG, "member %d out of range", member);
    if (b->t_first.size() != b->mem.size()) return bfail(nullptr, NGPU_E_STATE, "no completed batch run");
    out->first_start_ms = b->t_first[member];
    out->last_end_ms = b->t_last[member];
    out->comp_steps = b->cs[member];
    out->waves_per_run = b->lay_w[member];
    out->runs_per_item = NG_BATCH_THREADS / (64 * b->lay_w[member]) * b->lay_r[member];
    out->runs_per_group = b->lay_r[member];
    out->block_sync = b->lay_sync[member];
    out->lds_staging = b->lay_stage[member];
    return NGPU_OK;
}

extern "C" int ngpu_features(ngpu_ctx *c, const double *t_on, const double *t_off, double thr, double *out)
try {
    if (!c || !t_on || !t_off || !out) return fail(c, NGPU_E_ARG, "ngpu_features: NULL argument");
    if (!c->ran) return fail(c, NGPU_E_STATE, "ngpu_features needs a completed ngpu_run");
    if (c->rec.empty()) return fail(c, NGPU_E_STATE, "no record site");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<double> win(2 * c->nstim);
    for (int k = 0; k < c->nstim; k++) { win[k] = t_on[k]; win[c->nstim + k] = t_off[k]; }
    const int nruns = c->nsets * c->nstim;
    int rc;
    if ((rc = upload(c, c->d_fwin, win)) || (rc = alloc(c, c->d_feat, (size_t)nruns * NGPU_NFEAT)))
        return rc;
    hipError_t e = ngpu_launch_features(c->d_vhot.p, (int)c->rec.size(), c->nt, nruns, c->nstim, c->d_dts.p, c->d_fwin.p,
                                        c->d_fwin.p + c->nstim, thr, c->d_feat.p, c->stream);
    if (e != hipSuccess) return fail(c, NGPU_E_HIP, "feature kernel launch failed: %s", hipGetErrorString(e));
    HIPCHK(c, hipMemcpyAsync(out, c->d_feat.p, (size_t)nruns * NGPU_NFEAT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return NGPU_OK;
}
NGPU_CATCH(c, NGPU_E_ARG)

/* The features through the device's fixed scratch array instead of buffers
 * the context allocates: [0, nstim) t_on, [nstim, 2 nstim) t_off, then the
 * features from the next 256-byte boundary.  No allocation of any kind. */
extern "C" int ngpu_features_noalloc(ngpu_ctx *c, const double *t_on, const double *t_off, double thr, double *out)
{
    if (!c || !t_on || !t_off || !out) return fail(c, NGPU_E_ARG, "ngpu_features_noalloc: NULL argument");
    HIPCHK(c, hipSetDevice(c->device));
    double *scr = nullptr;
    size_t cap = 0;
    hipError_t e = ngpu_features_scratch(c->device, &scr, &cap);
    if (e != hipSuccess) return fail(c, NGPU_E_HIP, "device address of the feature scratch: %s", hipGetErrorString(e));
    const size_t nstim = (size_t)std::max(c->nstim, 0), nruns = (size_t)std::max(c->nsets, 0) * nstim;
    const size_t feat0 = (2 * nstim + 31) / 32 * 32;
    if (feat0 + nruns * NGPU_NFEAT > cap)
        return fail(c, NGPU_E_ARG, "%zu runs x %d features and %zu stimulus windows need %zu doubles of scratch; "
                    "ngpu_features_noalloc holds %zu (use ngpu_features)", nruns, NGPU_NFEAT, nstim,
                    feat0 + nruns * NGPU_NFEAT, cap);
    if (!c->ran) return fail(c, NGPU_E_STATE, "ngpu_features_noalloc needs a completed ngpu_run");
    if (c->rec.empty()) return fail(c, NGPU_E_STATE, "no record site");
    HIPCHK(c, hipMemcpyAsync(scr, t_on, nstim * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(scr + nstim, t_off, nstim * sizeof(double), hipMemcpyHostToDevice, c->stream));
    e = ngpu_launch_features(c->d_vhot.p, (int)c->rec.size(), c->nt, (int)nruns, (int)nstim, c->d_dts.p, scr, scr + nstim,
                             thr, scr + feat0, c->stream);
    if (e != hipSuccess) return fail(c, NGPU_E_HIP, "feature kernel launch failed: %s", hipGetErrorString(e));
    HIPCHK(c, hipMemcpyAsync(out, scr + feat0, nruns * NGPU_NFEAT * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return NGPU_OK;
}

extern "C" int ngpu_fetch_vhot(ngpu_ctx *c, double *vhot)
{
    if (!c || !vhot) return fail(c, NGPU_E_ARG, "NULL argument");
    if (!c->ran) return fail(c, NGPU_E_STATE, "no completed run");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpy(vhot, c->d_vhot.p, (size_t)ngpu_vhot_size(c) * sizeof(double), hipMemcpyDeviceToHost));
    return NGPU_OK;
}

/* ------------------------------------------------------------------ files */

static bool read_lines(const std::string &path, std::vector<std::string> &lines)
{
    std::ifstream in(path);
    if (!in) return false;
    std::string l;
    while (std::getline(in, l)) lines.push_back(l);
    return true;
}

static std::vector<double> tokens(const std::string &line)
{
    std::vector<double> out;
    std::stringstream ss(line);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        size_t b = tok.find_first_not_of(" \t\r\n");
        if (b == std::string::npos) continue;
        out.push_back(std::strtod(tok.c_str() + b, nullptr));
    }
    return out;
}

struct ParsedTree {
    ngpu_tree t{};
    std::vector<double> e, f, cm;
    std::vector<int> ks, s2c, bm, son, rs, re, rv, ss, se, fa, fx, cbl, cbfl, lrs, lre, flrs, flre;
};

static int parse_tree(ngpu_ctx *c, const std::string &path, ParsedTree &p)
{
    std::vector<std::string> L;
    if (!read_lines(path, L)) return fail(c, NGPU_E_FILE, "cannot open %s", path.c_str());
    if (L.size() < 31) return fail(c, NGPU_E_FILE, "%s: expected at least 31 lines, found %zu", path.c_str(), L.size());
    size_t li = 0;
    auto dvec = [&](size_t n, std::vector<double> &out, const char *what) -> int {
        std::vector<double> v = tokens(L[li++]);
        if (v.size() != n) return fail(c, NGPU_E_FILE, "%s: %s has %zu values, expected %zu", path.c_str(), what, v.size(), n);
        out = v;
        return 0;
    };
    auto ivec = [&](size_t n, std::vector<int> &out, const char *what) -> int {
        std::vector<double> v;
        int r = dvec(n, v, what);
        if (r) return r;
        out.resize(n);
        for (size_t i = 0; i < n; i++) out[i] = (int)v[i];   /* atoi truncation, as the reference */
        return 0;
    };
    auto ione = [&](int &x, const char *what) -> int {
        std::vector<int> v;
        int r = ivec(1, v, what);
        if (!r) x = v[0];
        return r;
    };
    int N, NC, NM, depth, logn, nf, ncall, cd, cfd, nl, nfl, rc;
    if ((rc = ione(N, "N")) || (rc = ione(NC, "NComps"))) return rc;
    if (N < 1 || N > 32000 || NC < 1 || NC > N) return fail(c, NGPU_E_FILE, "%s: N=%d and NComps=%d out of range", path.c_str(), N, NC);
    if ((rc = dvec(N, p.e, "e")) || (rc = dvec(N, p.f, "f")) || (rc = ivec(N, p.ks, "Ks")) || (rc = ivec(N, p.s2c, "SegToComp")) ||
        (rc = dvec(N, p.cm, "Cms")) || (rc = ione(NM, "NModels")))
        return rc;
    if (NM < 1 || NM > 32) return fail(c, NGPU_E_FILE, "%s: NModels=%d out of range", path.c_str(), NM);
    if ((rc = ivec((size_t)NM * N, p.bm, "boolModel")) ||
        (rc = ivec(N, p.son, "SonNoVec")) || (rc = ione(depth, "Depth")) || (rc = ione(logn, "LognDepth")) ||
        (rc = ione(nf, "nFathers")) || (rc = ione(ncall, "nCallForFather")) || (rc = ivec(nf, p.rs, "RelStarts")) ||
        (rc = ivec(nf, p.re, "RelEnds")) || (rc = ivec(ncall, p.rv, "RelVec")) || (rc = ivec(ncall + 1, p.ss, "SegStartI")) ||
        (rc = ivec(ncall + 1, p.se, "SegEndI")) || (rc = ivec(nf, p.fa, "Fathers")) || (rc = ivec((size_t)logn * N, p.fx, "FIdxs")) ||
        (rc = ione(cd, "CompDepth")) || (rc = ione(cfd, "CompFDepth")) || (rc = ivec((size_t)cd * 32, p.cbl, "CompByLevel32")) ||
        (rc = ivec((size_t)cfd * 32, p.cbfl, "CompByFLevel32")) || (rc = ione(nl, "nLRel")) || (rc = ivec(nl, p.lrs, "LRelStarts")) ||
        (rc = ivec(nl, p.lre, "LRelEnds")) || (rc = ione(nfl, "nFLRel")) || (rc = ivec(nfl, p.flrs, "FLRelStarts")) ||
        (rc = ivec(nfl, p.flre, "FLRelEnds")))
        return rc;
    ngpu_tree &t = p.t;
    t.N = N; t.ncomps = NC; t.nmodels = NM; t.depth = depth; t.logn_depth = logn; t.n_fathers = nf; t.n_call = ncall;
    t.e = p.e.data(); t.f = p.f.data(); t.cm = p.cm.data(); t.ks = p.ks.data(); t.seg_to_comp = p.s2c.data();
    t.bool_model = p.bm.data(); t.son_no = p.son.data(); t.rel_starts = p.rs.data(); t.rel_ends = p.re.data();
    t.rel_vec = p.rv.data(); t.seg_start = p.ss.data(); t.seg_end = p.se.data(); t.fathers = p.fa.data(); t.fidxs = p.fx.data();
    t.comp_depth = cd; t.comp_fdepth = cfd; t.comp_by_level32 = p.cbl.data(); t.comp_by_flevel32 = p.cbfl.data();
    t.n_lrel = nl; t.n_flrel = nfl; t.lrel_starts = p.lrs.data(); t.lrel_ends = p.lre.data();
    t.flrel_starts = p.flrs.data(); t.flrel_ends = p.flre.data();
    return NGPU_OK;
}

/* A value of a one-value-per-line file (Sim.csv, Stim_meta.csv): the first
 * token of line k, or false when the line is missing or empty. */
static bool first_token(const std::vector<std::string> &L, size_t k, double &x)
{
    if (k >= L.size()) return false;
    std::vector<double> v = tokens(L[k]);
    if (v.empty()) return false;
    x = v[0];
    return true;
}

/* Counts read from data files become allocation sizes: bounded here. */
static bool count_ok(double x, double lo, double hi) { return std::isfinite(x) && x >= lo && x <= hi && x == std::floor(x); }

static int open_into(ngpu_ctx *c, const std::string &d, const std::string &sfx, int mechset)
{
    const char *dir = d.c_str();
    int rc;
    ParsedTree pt;
    if ((rc = parse_tree(c, d + "/BasicConstSegP.csv", pt))) return rc;
    if (mechset < 0) {
        int nm = pt.t.nmodels;
        mechset = nm == SetTTPC::NMECH ? NGPU_MECH_TTPC : nm == SetMainen::NMECH ? NGPU_MECH_MAINEN : nm == 1 ? NGPU_MECH_PASSIVE : -1;
        if (mechset < 0) return fail(c, NGPU_E_FILE, "cannot infer a mechanism set from NModels=%d", nm);
    }
    pt.t.mechset = mechset;
    if ((rc = ngpu_set_tree(c, &pt.t))) return rc;
    /* Sim.csv */
    std::vector<std::string> L;
    double v0, nrd;
    if (!read_lines(d + "/Sim.csv", L) || !first_token(L, 0, v0) || !first_token(L, 2, nrd))
        return fail(c, NGPU_E_FILE, "cannot read %s/Sim.csv (v_init on line 1, nrecsite on line 3)", dir);
    if (!count_ok(nrd, 0, 64)) return fail(c, NGPU_E_FILE, "Sim.csv: nrecsite %g is not a count in 0..64", nrd);
    std::vector<int> rec;
    if ((int)nrd > 0) {
        std::vector<double> rs = L.size() > 3 ? tokens(L[3]) : std::vector<double>();
        if ((int)rs.size() != (int)nrd) return fail(c, NGPU_E_FILE, "Sim.csv: %d record sites declared, %zu given", (int)nrd, rs.size());
        for (double r : rs) {
            if (!count_ok(r, 0, pt.t.N - 1)) return fail(c, NGPU_E_FILE, "Sim.csv: record site %g outside the tree", r);
            rec.push_back((int)r);
        }
    }
    if ((rc = ngpu_set_sim(c, v0, (int)rec.size(), rec.data()))) return rc;
    /* AllParams.csv */
    L.clear();
    if (!read_lines(d + "/AllParams.csv", L) || L.empty()) return fail(c, NGPU_E_FILE, "cannot read %s/AllParams.csv", dir);
    std::vector<double> hdr = tokens(L[0]);
    const int np = mech_nparams(mechset);
    if (hdr.size() != 1 || !count_ok(hdr[0], 1, (double)L.size() - 1))
        return fail(c, NGPU_E_FILE, "AllParams.csv: the NSets header is not a count matching the %zu rows", L.size() - 1);
    const int nsets = (int)hdr[0];
    std::vector<double> P;
    P.reserve((size_t)nsets * np * pt.t.ncomps);
    for (int s = 0; s < nsets; s++) {
        std::vector<double> row = tokens(L[1 + s]);
        if ((int)row.size() != np * pt.t.ncomps) return fail(c, NGPU_E_FILE, "AllParams.csv row %d has %zu values, expected %d", s, row.size(), np * pt.t.ncomps);
        P.insert(P.end(), row.begin(), row.end());
    }
    if ((rc = ngpu_set_params(c, P.data(), nsets, pt.t.ncomps, np))) return rc;
    /* Stim_meta.csv: NStimuli, Nt, comp, loc (1-based segment), area */
    L.clear();
    double fnstim, fnt, floc, area;
    if (!read_lines(d + "/Stim_meta.csv", L) || !first_token(L, 0, fnstim) || !first_token(L, 1, fnt) ||
        !first_token(L, 3, floc) || !first_token(L, 4, area))
        return fail(c, NGPU_E_FILE, "cannot read %s/Stim_meta.csv (NStimuli, Nt, comp, loc, area on lines 1-5)", dir);
    if (!count_ok(fnstim, 1, 1 << 20) || !count_ok(fnt, 1, 1 << 26) || !count_ok(floc, 0, pt.t.N))
        return fail(c, NGPU_E_FILE, "Stim_meta.csv: NStimuli=%g Nt=%g loc=%g out of range", fnstim, fnt, floc);
    if (!std::isfinite(area) || !(area > 0)) return fail(c, NGPU_E_FILE, "Stim_meta.csv: area %g must be positive and finite", area);
    const int nstim = (int)fnstim, nt = (int)fnt, loc = (int)floc;
    std::vector<std::string> R;
    if (!read_lines(d + "/Stim_raw" + sfx + ".csv", R) || (int)R.size() < nstim) return fail(c, NGPU_E_FILE, "cannot read Stim_raw%s.csv with %d rows", sfx.c_str(), nstim);
    std::vector<double> amps((size_t)nstim * nt);
    for (int i = 0; i < nstim; i++) {
        std::vector<double> row = tokens(R[i]);
        if ((int)row.size() < nt) return fail(c, NGPU_E_FILE, "Stim_raw row %d has %zu values, expected >= %d", i, row.size(), nt);
        std::copy(row.begin(), row.begin() + nt, amps.begin() + (size_t)i * nt);
    }
    std::vector<std::string> T;
    if (!read_lines(d + "/times" + sfx + ".csv", T) && !read_lines(d + "/times.csv", T)) return fail(c, NGPU_E_FILE, "cannot read times.csv");
    std::vector<double> dts = T.empty() ? std::vector<double>() : tokens(T[0]);
    if ((int)dts.size() < nt) return fail(c, NGPU_E_FILE, "times.csv has %zu values, expected >= %d", dts.size(), nt);
    dts.resize(nt);
    return ngpu_set_stimuli(c, amps.data(), nstim, nt, dts.data(), nt, loc - 1, area);
}

/* The files are parsed and checked into a host-only context first (no device
 * is touched, so a malformed directory is reported the same way on any
 * machine); the device part is set up afterwards. */
/* ---- tree maps */
namespace {
struct TreeHolder {
    ngpu_tree t{};
    NgTreeMaps m;
    std::vector<double> e, f, cm;
    std::vector<int> seg_to_comp, bool_model;
};
std::mutex g_trees_mu;
std::map<const ngpu_tree *, std::unique_ptr<TreeHolder>> g_trees;

int tree_fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}
}

extern "C" int ngpu_tree_build(int N, const int *parent, const double *e, const double *f, const double *cm,
                               const int *seg_to_comp, int ncomps, int nmodels, const int *bool_model,
                               ngpu_tree **out)
{
    if (!out) return tree_fail(NGPU_E_ARG, "ngpu_tree_build: out is NULL");
    *out = nullptr;
    if (N < 1 || ncomps < 1 || nmodels < 1 || !parent || !e || !f || !cm || !seg_to_comp || !bool_model)
        return tree_fail(NGPU_E_ARG, "ngpu_tree_build: N, ncomps and nmodels must be positive and every array given");
    try {
        auto h = std::make_unique<TreeHolder>();
        h->m = ng_tree_maps(std::vector<int>(parent, parent + N));
        h->e.assign(e, e + N); h->f.assign(f, f + N); h->cm.assign(cm, cm + N);
        h->seg_to_comp.assign(seg_to_comp, seg_to_comp + N);
        h->bool_model.assign(bool_model, bool_model + (size_t)nmodels * N);
        for (int s = 0; s < N; s++)
            if (seg_to_comp[s] < 0 || seg_to_comp[s] >= ncomps)
                return tree_fail(NGPU_E_ARG, "ngpu_tree_build: seg_to_comp[" + std::to_string(s) + "] out of range");
        const NgTreeMaps &m = h->m;
        ngpu_tree &t = h->t;
        t.N = N; t.ncomps = ncomps; t.mechset = 0; t.nmodels = nmodels;
        t.e = h->e.data(); t.f = h->f.data(); t.cm = h->cm.data();
        t.ks = m.ks.data(); t.seg_to_comp = h->seg_to_comp.data(); t.bool_model = h->bool_model.data();
        t.son_no = m.son_no.data();
        t.depth = m.depth; t.logn_depth = m.logn_depth;
        t.n_fathers = (int)m.fathers.size(); t.n_call = (int)m.rel_vec.size();
        t.rel_starts = m.rel_starts.data(); t.rel_ends = m.rel_ends.data(); t.rel_vec = m.rel_vec.data();
        t.seg_start = m.seg_start.data(); t.seg_end = m.seg_end.data(); t.fathers = m.fathers.data();
        t.fidxs = m.fidxs.data();
        t.comp_depth = m.comp_depth; t.comp_fdepth = m.comp_fdepth;
        t.comp_by_level32 = m.comp_by_level32.data(); t.comp_by_flevel32 = m.comp_by_flevel32.data();
        t.n_lrel = (int)m.lrel_starts.size(); t.n_flrel = (int)m.flrel_starts.size();
        t.lrel_starts = m.lrel_starts.data(); t.lrel_ends = m.lrel_ends.data();
        t.flrel_starts = m.flrel_starts.data(); t.flrel_ends = m.flrel_ends.data();
        std::lock_guard<std::mutex> lk(g_trees_mu);
        *out = &h->t;
        g_trees[&h->t] = std::move(h);
        return NGPU_OK;
    } catch (const std::invalid_argument &ex) {
        return tree_fail(NGPU_E_ARG, std::string("ngpu_tree_build: ") + ex.what());
    } catch (const std::bad_alloc &) {
        return tree_fail(NGPU_E_NOMEM, "ngpu_tree_build: out of host memory");
    }
}

extern "C" void ngpu_tree_free(ngpu_tree *tree)
{
    std::lock_guard<std::mutex> lk(g_trees_mu);
    g_trees.erase(tree);
}

extern "C" int ngpu_tree_write(const ngpu_tree *t, const char *path)
{
    if (!t || !path) return tree_fail(NGPU_E_ARG, "ngpu_tree_write: tree and path must be given");
    if (t->N < 1 || t->depth < 0 || t->n_lrel != t->depth + 1 || t->comp_depth < 1)
        return tree_fail(NGPU_E_ARG, "ngpu_tree_write: inconsistent tree sizes");
    /* the file's last record: Ks - 1 (uint16: -1 as 65535) and the end of
     * the first run of the deepest level, less one */
    const int lrow = t->lrel_starts[t->depth];
    if (lrow < 0 || lrow >= t->comp_depth) return tree_fail(NGPU_E_ARG, "ngpu_tree_write: bad level table");
    const int last_run = t->comp_by_level32[(size_t)lrow * 32] - 1;
    if (last_run < 0 || last_run > t->n_call) return tree_fail(NGPU_E_ARG, "ngpu_tree_write: bad level table");
    std::vector<int> ksb(t->N + 1);
    for (int i = 0; i <= t->N; i++) {
        const int v = (i < t->N ? t->ks[i] : t->seg_end[last_run]) - 1;
        ksb[i] = v < 0 ? 65535 : v;
    }
    std::FILE *fp = std::fopen(path, "w");
    if (!fp) return tree_fail(NGPU_E_FILE, std::string("ngpu_tree_write: cannot open ") + path);
    const size_t N = t->N;
    ng_write_scalar(fp, t->N);
    ng_write_scalar(fp, t->ncomps);
    ng_write_record(fp, t->e, N);
    ng_write_record(fp, t->f, N);
    ng_write_record(fp, t->ks, N);
    ng_write_record(fp, t->seg_to_comp, N);
    ng_write_record(fp, t->cm, N);
    ng_write_scalar(fp, t->nmodels);
    ng_write_record(fp, t->bool_model, (size_t)t->nmodels * N);
    ng_write_record(fp, t->son_no, N);
    ng_write_scalar(fp, t->depth);
    ng_write_scalar(fp, t->logn_depth);
    ng_write_scalar(fp, t->n_fathers);
    ng_write_scalar(fp, t->n_call);
    ng_write_record(fp, t->rel_starts, t->n_fathers);
    ng_write_record(fp, t->rel_ends, t->n_fathers);
    ng_write_record(fp, t->rel_vec, t->n_call);
    ng_write_record(fp, t->seg_start, t->n_call + 1);
    ng_write_record(fp, t->seg_end, t->n_call + 1);
    ng_write_record(fp, t->fathers, t->n_fathers);
    ng_write_record(fp, t->fidxs, (size_t)t->logn_depth * N);
    ng_write_scalar(fp, t->comp_depth);
    ng_write_scalar(fp, t->comp_fdepth);
    ng_write_record(fp, t->comp_by_level32, (size_t)t->comp_depth * 32);
    ng_write_record(fp, t->comp_by_flevel32, (size_t)t->comp_fdepth * 32);
    ng_write_scalar(fp, t->n_lrel);
    ng_write_record(fp, t->lrel_starts, t->n_lrel);
    ng_write_record(fp, t->lrel_ends, t->n_lrel);
    ng_write_scalar(fp, t->n_flrel);
    ng_write_record(fp, t->flrel_starts, t->n_flrel);
    ng_write_record(fp, t->flrel_ends, t->n_flrel);
    ng_write_record(fp, ksb);
    const bool bad = std::ferror(fp) != 0;
    if (std::fclose(fp) != 0 || bad) return tree_fail(NGPU_E_FILE, std::string("ngpu_tree_write: error writing ") + path);
    return NGPU_OK;
}

extern "C" int ngpu_tree_disperse(int N, const int *parent, int cap, int *new_parent, int *src, int *new_n)
{
    if (N < 1 || !parent || !new_parent || !src || !new_n || cap < 0)
        return tree_fail(NGPU_E_ARG, "ngpu_tree_disperse: N must be positive and every array given");
    try {
        std::vector<int> np, sv;
        ng_disperse(std::vector<int>(parent, parent + N), np, sv);
        *new_n = (int)np.size();
        if ((int)np.size() > cap)
            return tree_fail(NGPU_E_ARG, "ngpu_tree_disperse: the dispersed tree has " + std::to_string(np.size()) +
                                             " segments, more than cap = " + std::to_string(cap));
        std::copy(np.begin(), np.end(), new_parent);
        std::copy(sv.begin(), sv.end(), src);
        return NGPU_OK;
    } catch (const std::invalid_argument &ex) {
        return tree_fail(NGPU_E_ARG, std::string("ngpu_tree_disperse: ") + ex.what());
    } catch (const std::bad_alloc &) {
        return tree_fail(NGPU_E_NOMEM, "ngpu_tree_disperse: out of host memory");
    }
}

extern "C" int ngpu_open(const char *dir, const char *suffix, int mechset, int device, ngpu_ctx **out)
{
    if (!dir || !out) return fail(nullptr, NGPU_E_ARG, "ngpu_open: NULL argument");
    *out = nullptr;
    ngpu_ctx *c = new (std::nothrow) ngpu_ctx();
    if (!c) return fail(nullptr, NGPU_E_NOMEM, "host allocation failed");
    int rc;
    try {
        rc = open_into(c, std::string(dir), std::string(suffix ? suffix : ""), mechset);
    } catch (const std::bad_alloc &) {
        rc = fail(c, NGPU_E_NOMEM, "host allocation failed while reading %s", dir);
    } catch (const std::exception &ex) {
        rc = fail(c, NGPU_E_FILE, "malformed data in %s: %s", dir, ex.what());
    }
    if (!rc) rc = init_device(c, device);
    if (rc) {
        g_last_error = c->err;
        ngpu_close(c);
        return rc;
    }
    *out = c;
    return NGPU_OK;
}

extern "C" int ngpu_write_vhot(ngpu_ctx *c, const char *path)
try {
    if (!c || !path) return fail(c, NGPU_E_ARG, "NULL argument");
    if (!c->ran) return fail(c, NGPU_E_STATE, "no completed run");
    long long n = ngpu_vhot_size(c);
    if (n > 0x7fffffffLL) return fail(c, NGPU_E_ARG, "%lld values exceed the VHotP int32 header", n);
    std::vector<double> h((size_t)n);
    int rc = ngpu_fetch_vhot(c, h.data());
    if (rc) return rc;
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(c, NGPU_E_FILE, "cannot write %s", path);
    int n32 = (int)n, prec = 3;
    bool ok = std::fwrite(&n32, sizeof n32, 1, fp) == 1 && std::fwrite(&prec, sizeof prec, 1, fp) == 1 &&
              std::fwrite(h.data(), sizeof(double), h.size(), fp) == h.size();
    ok = (std::fclose(fp) == 0) && ok;
    return ok ? NGPU_OK : fail(c, NGPU_E_FILE, "short write to %s", path);
}
NGPU_CATCH(c, NGPU_E_ARG)
