/*
 * neurogpu.h -- C ABI of libneurogpu, the MI355X (gfx950) engine for
 * NeuroGPU's compartmental simulation loop.
 *
 * The reference has no FFI: its boundary is the `neuroGPU <i>` process plus
 * files.  Each entry point below replaces one step of that process; the
 * reference function it stands in for is cited (paths relative to the
 * reference repository root, CANON = GUI/BBP_TTPC_EXAMPLE/pyNeuroGPU_unix/src).
 *
 * Conventions: 0 on success, a negative NGPU_E* code on failure, with the
 * message in ngpu_last_error().  The context owns all device memory; callers
 * own host buffers.  One context per thread; contexts are not re-entrant.
 * No entry point retains a caller pointer after it returns.
 */
#ifndef NEUROGPU_H
#define NEUROGPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define NGPU_OK 0
#define NGPU_E_ARG (-1)      /* invalid argument or inconsistent sizes        */
#define NGPU_E_FILE (-2)     /* missing or malformed data file                */
#define NGPU_E_HIP (-3)      /* HIP runtime error                             */
#define NGPU_E_STATE (-4)    /* call order: model/params/stimuli not set      */
#define NGPU_E_NOMEM (-5)    /* host or device allocation failed              */

/* Mechanism sets compiled into the library. */
#define NGPU_MECH_PASSIVE 0  /* pas                                               */
#define NGPU_MECH_TTPC 1     /* BBP TTPC, 13 mechanisms (Code/OneKernel/src/AllModels.cu) */
#define NGPU_MECH_MAINEN 2   /* ca cad kca km kv na pas (Figures/FigureS4_mainen/CParsed) */
#define NGPU_MECH_TTPC_GEN 3 /* the TTPC set generated by neurogpu_amd.nmodl from the model's
                                .mod files in Figures/Figure2_BBP_TTPC_clean (as
                                PYS/extractModel.py generates AllModels.cu): the mechanisms,
                                parameters and tree files of NGPU_MECH_TTPC */
#define NGPU_MECH_MAINEN_GEN 4 /* the Mainen set generated the same way from
                                  Figures/FigureS4_mainen's .mod files: the mechanisms,
                                  parameters and tree files of NGPU_MECH_MAINEN */

typedef struct ngpu_ctx ngpu_ctx;

/* Tree data = the contents of BasicConstSegP.csv, i.e. the reference's HMat
 * (CANON/Util.h:70-110) as filled by ReadParallelNeuronData (CANON/MainC.cu:335-551).
 * Index conventions are the file's (1-based rows; e, f, ks, fidxs in Hines-row order). */
typedef struct {
    int N, ncomps, mechset, nmodels;
    const double *e, *f, *cm;
    const int *ks, *seg_to_comp, *bool_model, *son_no;
    int depth, logn_depth, n_fathers, n_call;
    const int *rel_starts, *rel_ends, *rel_vec, *seg_start, *seg_end, *fathers, *fidxs;
    int comp_depth, comp_fdepth;
    const int *comp_by_level32, *comp_by_flevel32;
    int n_lrel, n_flrel;
    const int *lrel_starts, *lrel_ends, *flrel_starts, *flrel_ends;
} ngpu_tree;

typedef struct {
    double ms_total;          /* upload + kernels + download                         */
    double ms_kernel;         /* HIP events around the simulation launches            */
    double comp_steps;        /* sum over runs of N * Nt                              */
    double comp_steps_per_s;  /* comp_steps / (ms_kernel / 1000)                      */
    int waves_per_run, slots_per_lane, runs_per_block, launches;
    int nvar;                 /* distinct run-varying parameter columns               */
    int nsets, nstim, nrec, nt, N;
    int runs_per_group;       /* runs sharing one group of waves_per_run waves        */
    int solve;                /* elimination used: NGPU_SOLVE_REFERENCE or NGPU_SOLVE_FAST */
} ngpu_stats;

/* Create a context on HIP device `device` (replaces cudaSetDevice in CANON/kernel.cu:24-31). */
int ngpu_create(int device, ngpu_ctx **out);

/* Model tree and mechanism placement (replaces initFrameWork's
 * cudaMemcpyToSymbol of the tree maps, CANON/CudaStuff.cu:372-437). */
int ngpu_set_tree(ngpu_ctx *ctx, const ngpu_tree *tree);

/* v_init and record sites (replaces ReadSimData, CANON/Util.cu:394-432). */
int ngpu_set_sim(ngpu_ctx *ctx, double v_init, int nrec, const int *rec_sites);

/* Parameter matrix [nsets][nparams*ncomps], index p*ncomps + comp
 * (replaces ReadAllParams, CANON/Util.cu:210-234). */
int ngpu_set_params(ngpu_ctx *ctx, const double *allparams, int nsets, int ncomps, int nparams);

/* Parameters assembled on the device, for callers that change only a few
 * free parameters per set (a GA).  tpl is one parameter-matrix row
 * [nparams*ncomps]; free parameter i sits at every entry that holds the
 * sentinel 12300 + i (the convention of the GA evaluators'
 * create_params_to_all_params / my_assemble_allparams,
 * Figures/Figure4_DEAP/extractModel_mappings.py:158-171, :258-270); every
 * other entry is the value all sets share.  Each free parameter needs at
 * least one entry.  Then ngpu_set_genomes uploads the free parameters of
 * nsets sets, x [nsets][nfree], and a device kernel scatters them into the
 * per-set parameter table (the host never builds the full matrix).  Equal to
 * ngpu_set_params with the template filled in on the host, bit for bit. */
int ngpu_set_param_template(ngpu_ctx *ctx, const double *tpl, int ncomps, int nparams, int nfree);
int ngpu_set_genomes(ngpu_ctx *ctx, const double *x, int nsets);

/* Stimuli: amps [nstim][nt_amps] in nA (nt_amps >= nt), dts [nt] in ms,
 * stim_seg is 0-based (Stim_meta loc - 1), area in um^2
 * (replaces ReadCSVStim, CANON/Util.cu:337-390). */
int ngpu_set_stimuli(ngpu_ctx *ctx, const double *amps, int nstim, int nt_amps,
                     const double *dts, int nt, int stim_seg, double area);

/* Launch shape (the autotuner's knobs; 0 = choose automatically): a group of
 * waves_per_run waves (1, 2, 4, 8 or 16) advances runs_per_group runs (1..8)
 * whose positions and solve tasks share the group's lanes; a workgroup holds
 * as many groups as fit in LDS and 8 waves (one group of 16 waves: a
 * 1024-thread workgroup, four waves per SIMD).  steps_per_launch (0 = all steps
 * in one launch), use_graph (capture the launch sequence into a hipGraph and
 * replay it). No reference counterpart: the reference fixes the launch at
 * NSets x (32, NStim) (CANON/CudaStuff.cu:487-507). */
int ngpu_set_launch(ngpu_ctx *ctx, int waves_per_run, int runs_per_group,
                    int steps_per_launch, int use_graph);

/* The form of the Hines elimination (BeforeLU, CANON/CudaStuff.cu:31-69).
 * NGPU_SOLVE_REFERENCE keeps the reference's operations in its order
 * (u_i -= f (e / u), b_i -= (b e) / u, two IEEE divisions a row), which its
 * exported trees need: they couple zero-area nodes with coefficients near
 * 1e21.  NGPU_SOLVE_FAST takes each row's pivot quotients once, from one
 * refined reciprocal of the pivot, and folds the next row through them;
 * results equal the reference order's up to rounding on well-scaled trees.
 * NGPU_SOLVE_AUTO
 * (the default) takes the fast form when every cm > 0 and every |e|, |f| is
 * at most 1e6.  No reference counterpart (the reference has one form). */
#define NGPU_SOLVE_AUTO 0
#define NGPU_SOLVE_REFERENCE 1
#define NGPU_SOLVE_FAST 2
int ngpu_set_solve(ngpu_ctx *ctx, int mode);

/* Run every (set, stimulus) for all Nt steps and copy the traces to vhot
 * [nsets][nstim][nrec][nt] (caller-owned; may be NULL to keep them on the
 * device).  Replaces callKernel/stEfork2Main (CANON/CudaStuff.cu:443-587). */
int ngpu_run(ngpu_ctx *ctx, double *vhot, ngpu_stats *stats);

/* Steps [t_begin, t_end) of every (set, stimulus): t_begin = 0 starts from
 * the initial state, any other t_begin continues from the state the previous
 * call left there (NGPU_E_STATE otherwise; changing the tree, record sites,
 * parameters, stimuli or launch shape discards that state).  The traces of
 * the window's steps are written to the context's trace buffer, so after the
 * window that ends at Nt ngpu_fetch_vhot / ngpu_features see the whole run.
 * Replaces the time loop of the reference kernel, which runs all Nt steps in
 * one launch (CANON/CudaStuff.cu:228-338); windows let a caller interleave
 * other work or bound the length of one launch. */
int ngpu_run_steps(ngpu_ctx *ctx, int t_begin, int t_end, ngpu_stats *stats);

/* Sizes the context holds (0 where not yet set): parameter sets, stimuli,
 * record sites, steps, segments -- e.g. after ngpu_open, to size the buffers
 * of ngpu_run / ngpu_features.  The reference reads them from the files'
 * headers (CANON/MainC.cu:335-551, CANON/Util.cu:337-432). */
int ngpu_get_sizes(const ngpu_ctx *ctx, int *nsets, int *nstim, int *nrec, int *nt, int *N);

/* Spike features of every trace of the last run (record site 0), computed on
 * the device: out [nsets][nstim][NGPU_NFEAT] = spike count, first-spike
 * latency (ms after t_on[stim]), mean AP amplitude above the pre-stimulus
 * baseline (mV), mean AP width at thr (ms), mean ISI (ms); NaN where undefined.
 * Only spikes whose threshold crossing lies in [t_on[stim], t_off[stim]) count.
 * A trace with a NaN sample (a diverged run) has every feature NaN, the spike
 * count included (the evaluator's NaN-trace case, PYS/efel_ext.py:47-49, :90-95).
 * Replaces the per-individual eFEL pass of the GA evaluator (PYS/efel_ext.py:23-97,
 * threshold at :15). */
#define NGPU_NFEAT 5
int ngpu_features(ngpu_ctx *ctx, const double *t_on, const double *t_off, double thr, double *out);

/* ngpu_features without any allocation: no counterpart in the reference.  It
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The stimulus windows and
 * the features are staged in a fixed device array of 2^20 doubles (8 MiB),
 * which each device holds from the first call until the process ends:
 * 2 * nstim (rounded up to 32) + 5 * nsets * nstim must fit in it, else
 * NGPU_E_ARG.  Calls on one device share that array, so calls that overlap
 * on one device race; its contents after a call are unspecified.  Same
 * arguments and results as ngpu_features. */
int ngpu_features_noalloc(ngpu_ctx *ctx, const double *t_on, const double *t_off, double thr, double *out);

/* Copy the traces of the last run to vhot (same layout as ngpu_run). */
int ngpu_fetch_vhot(ngpu_ctx *ctx, double *vhot);

/* ---- Tree maps (the reference builds them in Python at model export,
 * NeuroGPU_Base/scripts/pyscripts/create_auxilliary_data_3.py:22-315 and
 * make_tree_from_parent_vec.py:16-59, and writes BasicConstSegP.csv at
 * :229-272).  parent[s] is the parent segment of segment s (parent[0] = -1,
 * parent[s] < s), at most two children a segment and one on the root
 * (ngpu_tree_disperse first otherwise); e, f in Hines-row order (row
 * N-1-s for segment s); bool_model [nmodels][N].  *out gets every map of
 * BasicConstSegP.csv, owned by the library until ngpu_tree_free; its
 * mechset is 0 (set it before ngpu_set_tree).  NGPU_E_ARG with a message
 * (ngpu_last_error(NULL)) for a parent vector the schedule cannot take. */
int ngpu_tree_build(int N, const int *parent, const double *e, const double *f, const double *cm,
                    const int *seg_to_comp, int ncomps, int nmodels, const int *bool_model, ngpu_tree **out);
void ngpu_tree_free(ngpu_tree *tree);
/* Write a tree as BasicConstSegP.csv, "%.9f," values, one record a line
 * (create_auxilliary_data_3.py:229-272). */
int ngpu_tree_write(const ngpu_tree *tree, const char *path);
/* Wide branch points dispersed into extension segments, as the reference's
 * model export does (hinesDisperseBranching, NeuroGPU_Base/runModel.hoc:158-215):
 * a segment with k > 2 children (the root: k > 1) gets k - 1 extension
 * segments carrying its children two at a time; the tree is renumbered depth
 * first.  new_parent and src hold up to cap entries (2N always suffices);
 * src[i] = the original segment of new segment i, or -(h + 1) for an
 * extension of segment h; *new_n = the new size. */
int ngpu_tree_disperse(int N, const int *parent, int cap, int *new_parent, int *src, int *new_n);

/* Parse a reference data directory (BasicConstSegP.csv, Sim.csv,
 * AllParams.csv, Stim_meta.csv, Stim_raw<suffix>.csv, times<suffix>.csv) and
 * load it into a new context (replaces RunByModelP's Init,
 * CANON/MainC.cu:37-80 and :592-626).  mechset < 0 infers the set from NModels. */
int ngpu_open(const char *data_dir, const char *suffix, int mechset, int device, ngpu_ctx **out);

/* Write the last traces as VHotP<dev>.dat: int32 N, int32 3, N doubles
 * (replaces SaveArrayToFile, CANON/Util.cu:443-456). */
int ngpu_write_vhot(ngpu_ctx *ctx, const char *path);

/* Number of doubles ngpu_run writes (nsets * nstim * nrec * nt). */
long long ngpu_vhot_size(const ngpu_ctx *ctx);

/* Parameters per compartment of a mechanism set (NPARAMS of the generated
 * AllModels.cuh), or a negative code. */
int ngpu_mechset_nparams(int mechset);

/* Message of the last failed call on ctx (or of the last failed
 * ngpu_create / ngpu_open when ctx is NULL). */
const char *ngpu_last_error(const ngpu_ctx *ctx);

void ngpu_close(ngpu_ctx *ctx);

/* ---- Batches: several contexts (morphologies of uneven size and step count,
 * one mechanism set) simulated by one persistent launch.  No reference
 * counterpart: the reference runs one morphology per process and launch
 * (CANON/CudaStuff.cu:487-507).  At most 8 contexts per batch.  Each
 * workgroup of 8 waves takes items from a device work queue, most work first;
 * an item is (8 / waves_per_run) * runs_per_group runs of one member.  Traces and features stay
 * with each member context (ngpu_fetch_vhot, ngpu_features). */
typedef struct ngpu_batch ngpu_batch;
typedef struct {
    int waves_per_run;        /* 1, 2, 4 or 8; 0 = 1 up to 512 segments, else 8        */
    int block_sync;           /* levels synchronised per workgroup (1) or per wave (0,
                                 waves_per_run 1 only); -1 = per wave when possible   */
    int lds_staging;          /* Hines row coefficients e, f and the runs' varying
                                 parameter values staged in LDS (1) or read from
                                 global memory (0); -1 = in LDS when they fit         */
    int runs_per_group;       /* 1..8 runs per group of waves; 0 = as many as fit    */
} ngpu_batch_launch;
typedef struct {
    double ms_kernel;         /* HIP events around the persistent launch              */
    double comp_steps;        /* all members                                          */
    int nitems, workgroups, lds_bytes;
} ngpu_batch_stats;
typedef struct {
    double first_start_ms, last_end_ms;   /* from the first item's start (device clock) */
    double comp_steps;
    int waves_per_run, runs_per_item, block_sync, lds_staging;   /* what ran          */
    int runs_per_group;
} ngpu_batch_member;
int ngpu_batch_create(int device, ngpu_batch **out);
/* Add a configured context (tree, parameters, stimuli); returns its member
 * index.  The caller keeps ownership and must keep it alive. */
int ngpu_batch_add(ngpu_batch *batch, ngpu_ctx *ctx);
int ngpu_batch_set_launch(ngpu_batch *batch, int member, const ngpu_batch_launch *cfg);
int ngpu_batch_run(ngpu_batch *batch, ngpu_batch_stats *stats);
int ngpu_batch_member_info(const ngpu_batch *batch, int member, ngpu_batch_member *out);
const char *ngpu_batch_last_error(const ngpu_batch *batch);
void ngpu_batch_close(ngpu_batch *batch);

/* ---- Surrogate network (no reference counterpart: the reference GA has no
 * surrogate).  An ensemble of `members` MLPs nin -> H -> H -> H -> nout with
 * ReLU, BF16 weights and activations on MFMA with FP32 accumulation, FP32
 * master weights and Adam.  Parameters of one member, in order: W1[H][nin],
 * b1[H], W2[H][H], b2[H], W3[H][H], b3[H], W4[nout][H], b4[nout]; members are
 * concatenated (1..64 members; every launch covers all of them). */
typedef struct ngpu_mlp ngpu_mlp;
int ngpu_mlp_create(int device, int nin, int hidden, int nout, int members, ngpu_mlp **out);
long long ngpu_mlp_nparams(const ngpu_mlp *mlp);                      /* per member */
int ngpu_mlp_set_weights(ngpu_mlp *mlp, const float *w);              /* all members; resets Adam */
int ngpu_mlp_get_weights(ngpu_mlp *mlp, float *w);
/* Mean-squared-error training, mini-batch Adam (beta 0.9/0.999, eps 1e-8);
 * losses [members][epochs] (may be NULL). */
int ngpu_mlp_train(ngpu_mlp *mlp, const float *x, const float *t, int n, int epochs, int batch, float lr,
                   unsigned seed, float *losses);
/* out [members][n][nout] */
int ngpu_mlp_predict(ngpu_mlp *mlp, const float *x, int n, float *out);
/* Loss and gradient of member `member` on one batch (for checks against a CPU restatement). */
int ngpu_mlp_gradients(ngpu_mlp *mlp, int member, const float *x, const float *t, int n, float *grad, float *loss);
const char *ngpu_mlp_last_error(const ngpu_mlp *mlp);
void ngpu_mlp_close(ngpu_mlp *mlp);
/* C[M][N] = op(A) op(B) on the device, A/B bf16 bit patterns (op = transpose
 * when ta/tb), FP32 result; a test entry point for the MFMA GEMM. */
int ngpu_gemm_bf16(int device, int ta, int tb, int M, int N, int K, const unsigned short *A,
                   const unsigned short *B, float *C);

#ifdef __cplusplus
}
#endif
#endif
