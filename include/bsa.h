/*
 * bsa.h — C ABI of the MI355X engine for the likelihood of codeml's
 * branch-site model A (model = 2, NSsites = 2).
 *
 * One engine handle holds one gene (tree + compressed codon alignment) on
 * one GPU.  All arrays cross the boundary as plain host pointers; every call
 * returns a status code (0 = success) and never exits the process; the text
 * of the last error on the calling thread is in bsa_last_error().
 *
 * Reference interface replaced (the reference passes everything through the
 * globals com/tree/nodes, codeml.h:87-105):
 *   bsa_create      <- CUDA_Init(double *x)                cuda-codeml.cu:1106
 *                      (+ deviceInit 762-837, prepareTree 851-919)
 *   bsa_destroy     <- CUDA_FreeAll(void)                  cuda-codeml.cu:1180
 *   bsa_lnl         <- CUDA_lfundG(double x[], int np)     treesub.c:7397
 *                      (CUDA_fx_r treesub.c:7554 and its four
 *                      CUDA_ScheduleConditionalPNode calls, cuda-codeml.cu:1336)
 *   bsa_class_lnl   <- com.fhK[ir*npatt+h] filled by CUDA_ScheduleConditionalPNode
 *                      (cuda-codeml.cu:1328) and the lfundG site loop
 *                      (dfsites, treesub.c:7387)
 *   bsa_lnl_grad_fd <- gradientB (tools.c:6326-6350), the finite-difference
 *                      gradient ming2 uses because the reference passes
 *                      dfun = NULL (codeml.c:688)
 */
#ifndef BSA_H
#define BSA_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BSA_OK            0
#define BSA_EINVAL        1   /* bad argument / inconsistent input       */
#define BSA_ENOMEM        2   /* host or device allocation failed        */
#define BSA_EDEVICE       3   /* HIP runtime error                       */
#define BSA_ENUMERIC      4   /* eigen solver failed / non-finite value  */

/* precision of the conditional likelihoods */
#define BSA_FP64          0   /* FP64 partials                             */
#define BSA_FP32          1   /* FP32 partials (FP64 arithmetic, stochastic
                                 rounding on store), power-of-two scaling  */

/* kernel variant */
#define BSA_KERNEL_MFMA   0   /* 16x16x4 FP64 matrix-core contractions      */
#define BSA_KERNEL_VALU   1   /* pruning by plain vector FMA contraction    */
#define BSA_KERNEL_MFMA32 2   /* with BSA_FP32 only: as BSA_KERNEL_MFMA, with
                                 the gradient sweep's products on the FP32
                                 matrix cores: the child adjoints (P - I in
                                 FP32, the identity term added in FP64) and
                                 the interior children's branch matrices
                                 (FP32 per 16-pattern group, summed in
                                 FP64); the pruning keeps FP64 products    */
#define BSA_KERNEL_FP32MC 3   /* with BSA_FP32 only: as BSA_KERNEL_MFMA32,
                                 with the pruning's products on the FP32
                                 matrix cores too: per pattern row the two
                                 largest child terms and the identity term
                                 in FP64, the rest of P - I in FP32 (jobs
                                 with more than two interior children keep
                                 FP64 products)                            */

typedef struct bsa_engine bsa_engine;

typedef struct {
    /* tree: tips are nodes 0..ns-1 (alignment order), interior ns..nnode-1 */
    int ns;
    int nnode;
    int root;
    const int *father;        /* [nnode], father[root] = -1                 */
    const int *label;         /* [nnode], branch type of the edge above
                                 each node: 0 background, 1 foreground      */
    const int *ibranch;       /* [nnode], index of each edge in the branch-
                                 length part of x (codeml's tree.branches
                                 order); root -1                            */
    /* alignment, compressed to site patterns */
    int npatt;
    const unsigned char *z;   /* [ns*npatt], codes 0..60 = sense codons,
                                 61..ncodes-1 = ambiguity codes             */
    int ncodes;
    const unsigned char *code_mask; /* [ncodes*61], 1 where code allows the
                                       sense codon (CharaMap, treesub.c:1146) */
    const double *fpatt;      /* [npatt] pattern counts                     */
    const double *pi;         /* [61] codon frequencies                     */
    /* options */
    int precision;            /* BSA_FP64 or BSA_FP32                       */
    int kernel;               /* BSA_KERNEL_MFMA, _VALU, _MFMA32, _FP32MC   */
    int device;               /* HIP device ordinal                         */
    int tile;                 /* site patterns per pruning block: 64, 128,
                                 or 0 for the default                       */
} bsa_problem;

/* Parameter vector x in codeml's optimisation variables (codeml.c:1980-1981,
 * 2080-2098):  x = [t_0 .. t_{nbranch-1}, kappa, x_p0, x_p1, w0 (, w2)]
 * with  p0 = e^{x_p0}/(1+e^{x_p0}+e^{x_p1}),  p1 = e^{x_p1}/(...)
 * (f_and_x, tools.c:1280-1305).  With fix_omega the vector has no w2 and
 * w2 = omega_fix (H0 uses omega_fix = 1, codeml.c:1648-1649, 2425).
 * With p_transformed = 0, x carries p0, p1 themselves (codeml's LASTROUND,
 * used by in.codeml files starting with -1). */
typedef struct {
    int fix_omega;
    double omega_fix;
    int p_transformed;
} bsa_model_opts;

int  bsa_create(const bsa_problem *prob, bsa_engine **out);
int  bsa_destroy(bsa_engine *eng);
int  bsa_set_model(bsa_engine *eng, const bsa_model_opts *opts);
int  bsa_num_params(const bsa_engine *eng);          /* np, or -1          */

/* -lnL at x (np values), as codeml's lfundG returns it */
int  bsa_lnl(bsa_engine *eng, const double *x, int np, double *neg_lnl);

/* -lnL and its gradient d(-lnL)/dx in the same variables (ming2's dfun
 * contract, tools.c:6361, 6408, 6461-6462), from one forward pass and one
 * adjoint (preorder) sweep: branch-length derivatives through
 * dP/dt = Qfactor U diag(lambda e^{lambda t}) V, rate-matrix parameters through
 * the eigen derivative of P(t), and the chain rule through Qfactor, the class
 * proportions and the logistic p0, p1.  Replaces the finite-difference
 * gradient the reference uses because it passes dfun = NULL (codeml.c:688). */
int  bsa_lnl_grad(bsa_engine *eng, const double *x, int np, double *neg_lnl, double *grad);

/* -lnL and its finite-difference gradient d(-lnL)/dx in the same variables,
 * as gradientB (tools.c:6326-6350): step small*(|x_i|+1) (small <= 0 means
 * 1e-6, Small_Diff), forward differences (np+1 evaluations) or, with
 * central != 0, central differences with step^0.67 (2np+1 evaluations) */
int  bsa_lnl_grad_fd(bsa_engine *eng, const double *x, int np, int central, double small,
                     double *neg_lnl, double *grad);

/* after bsa_lnl / bsa_lnl_grad / bsa_lnl_grad_fd: per-class log f(x_h | class) [4*npatt]
 * (class-major, as com.fhK with node scaling) and the mixture log f_h
 * [npatt]; either pointer may be NULL */
int  bsa_class_lnl(bsa_engine *eng, double *fhK, double *site_lnf);

/* Bayes empirical Bayes posteriors of the four site classes for model A at
 * x (normally the H1 estimates): branch lengths, kappa and the two Qfactors
 * fixed at x; w0 ~ U(0, 1), w2 ~ U(1, 11) and (p0, p1, p2) ~ Dirichlet(1, 1,
 * 1) on a 10-point grid per dimension (lfunNSsites_ACD, codeml.c:6617-6826,
 * which the reference runs after every H1 fit, codeml.c:5215-5216).
 * post[4*npatt]: class-major posteriors per site pattern (classes 0, 1, 2a,
 * 2b; P(w2 > 1) = post[2] + post[3]); ml3 (may be NULL) = {log f(X) as the
 * reference prints it, the per-site scale S1, log of the grid sum};
 * postw20 (may be NULL) = posterior of w0 on its 10 grid points, then of w2;
 * postp0p1 (may be NULL) = [100] posterior on the p0-p1 triangles;
 * lnf_sets (may be NULL) = [121*npatt] log f(x_h | w set) in the
 * reference's set order.  The engine's last evaluation is left at x. */
int  bsa_beb(bsa_engine *eng, const double *x, int np, double *post, double *ml3,
             double *postw20, double *postp0p1, double *lnf_sets);

/* the parameters of the last evaluation in natural form:
 * out[0..4] = kappa, p0, p1, w0, w2; out[5..6] = Qfactor background,
 * foreground; out[7..10] = class proportions */
int  bsa_last_params(const bsa_engine *eng, double *out11);

/* enable != 0: run each evaluation's device work as one HIP graph, captured
 * at the first evaluation of each kind (lnL only, lnL + gradient) and
 * replayed afterwards (the launch sequence depends only on the tree and the
 * data); cuts launch overhead for small genes.  bsa_timing then reports the
 * whole device time in out[0] and 0 for the phases. */
int  bsa_set_graph(bsa_engine *eng, int enable);

/* timing in ms, from HIP events on the engine's stream: out[0] P(t) and tip
 * tables, out[1] pruning (all levels), out[2] root reduction, out[3] child
 * adjoints (all depths; 0 without gradient), out[4] branch matrices (one
 * launch), out[5] adjoint finish (eigen-basis kernels, set sums), out[6]
 * host wall time of the whole evaluation, all for the last evaluation;
 * out[7..13] the same summed since creation or the last reset; out[14]
 * evaluations, out[15] pruning launches and out[16] adjoint launches (child
 * adjoints + branch matrices) in that sum.  reset != 0 clears the sums
 * after reading them. */
int  bsa_timing(bsa_engine *eng, double *out17, int reset);

/* host only (no GPU): model A at natural parameters, as SetParametersNSsites
 * (codeml.c:2470-2520) with eigenQcodon/eigenQREV (codeml.c:3103-3272,
 * tools.c:4850-4937): freqK[4] class proportions, Qfactor[2] (background,
 * foreground), and when Root, U, V are non-NULL the eigen systems for
 * omega = w0, 1, w2: Root[3][61], U[3][61*61], V[3][61*61], Q = U diag(Root) V */
int  bsa_model_a_setup(double kappa, double p0, double p1, double w0, double w2,
                       const double *pi61, double *freqK4, double *Qfactor2, double *Root,
                       double *U, double *V);

/* work of one evaluation, counted on 61 states:
 * out[0] pruning flops executed (site classes 2/3 share partials with 0/1
 * off the foreground path), out[1] pruning algorithmic HBM bytes executed
 * (partials read and written, tip codes, scale exponents), out[2] flops and
 * out[3] bytes of the reference algorithm's pruning (4 classes over the
 * whole tree), out[4] pruning launches (tree levels), out[5] pruning
 * (node, class) jobs; per gradient: out[6] flops and out[7] algorithmic
 * bytes of the child adjoints, out[8] flops and out[9] algorithmic bytes of
 * the branch matrices */
int  bsa_work(bsa_engine *eng, double *out10);

/* launch shapes: pruning tile (64 or 128 patterns per block), child-adjoint
 * tiles per block (0: chosen by each launch's size, else 1..8), branch-matrix
 * tiles per chunk (0: chosen by the gene's size, at most 32, else 1..64).  Results do not
 * depend on the first two; the chunking sets the order in which a branch
 * matrix is summed (gradients move in the 16th digit).  No counterpart in
 * the reference. */
int  bsa_set_shapes(bsa_engine *eng, int tile, int tpb, int tpc);

/* pruning schedule: one launch per tree level (BSA_PRUNE_LEVELS, the
 * default), or every level in one persistent launch (BSA_PRUNE_WALK: blocks
 * take (gene, 128-pattern tile) items in turn and walk the gene's pruning
 * jobs in level order for their tile; results bitwise those of the level
 * launches; BSA_KERNEL_MFMA or BSA_KERNEL_MFMA32 only).  Replaces the
 * reference's launch and host synchronisation per edge
 * (cuda-codeml.cu:1306-1321). */
#define BSA_PRUNE_LEVELS 0
#define BSA_PRUNE_WALK   1
int  bsa_set_prune(bsa_engine *eng, int mode);

/* autotuner: times each candidate of each launch parameter in turn
 * (pruning mode, pruning tile, child-adjoint tiles per block, branch-matrix
 * tiles per chunk) by the device time of `reps` lnL + gradient evaluations
 * at x, and keeps the fastest.  out7 (may be NULL) = {tile, tpb, tpc,
 * chunks, best ms, ms of the shapes it started from, pruning mode}.  The
 * engine's last evaluation is at x.  No counterpart in the reference. */
int  bsa_tune(bsa_engine *eng, const double *x, int np, int reps, double *out7);

/* codeml's parameter bounds for model A in the engine's x layout
 * (SetxBound codeml.c:1786-1828, SetxBoundTimes treesub.c:3872-3884):
 * branch lengths [4e-6, 50], kappa [1e-4, 999], transformed p's [-99, 99]
 * (untransformed: [1e-6, 1 - 1e-6]), w0 [1e-6, 1], w2 [1, 999] */
int  bsa_default_bounds(const bsa_engine *eng, double *lo, double *hi);

/* gradient used by bsa_fit */
#define BSA_GRAD_ADJOINT  0   /* bsa_lnl_grad                               */
#define BSA_GRAD_FD       1   /* bsa_lnl_grad_fd, forward differences       */

typedef struct {
    int    max_iter;          /* quasi-Newton iterations (0: 2000; < 0: none,
                                 only the Newton iterations below)           */
    double e;                 /* convergence tolerance on relative changes of
                                 -lnL and x, as ming2's e (<= 0: 1e-8)        */
    double max_step;          /* largest change of one variable per line
                                 search start (<= 0: 5)                      */
    int    gradient;          /* BSA_GRAD_ADJOINT or BSA_GRAD_FD             */
    int    newton;            /* Newton iterations after the quasi-Newton
                                 ones (0: none): Hessian of the free
                                 variables by forward differences of the
                                 gradient (one evaluation per variable),
                                 damped until positive definite; lands on the
                                 maximum also along flat directions, such as
                                 w2 on a ridge, where BFGS's curvature
                                 estimate is poor (no counterpart in ming2)  */
} bsa_fit_opts;

/* why a fit stopped (info[2] of bsa_fit); BSA_STOP_HEND, BSA_STOP_PGRAD
 * and BSA_STOP_NEWTON are convergence */
#define BSA_STOP_LIMIT     0  /* iteration limit                              */
#define BSA_STOP_HEND      1  /* two successive steps changed -lnL and x by
                                 less than e relative, confirmed after a
                                 restart of the Hessian approximation (the
                                 reference's H_end rule, tools.c:5724-5740) */
#define BSA_STOP_PGRAD     2  /* projected gradient below
                                 1e-8 + 1e-10 max(1, -lnL)                   */
#define BSA_STOP_STALL     3  /* -lnL fell by less than 1e-3 e |lnL| over the
                                 last 50 iterations (flat ridge)              */
#define BSA_STOP_NODESCENT 4  /* no descent step along steepest descent       */
#define BSA_STOP_NEWTON    5  /* the last Newton step (bsa_fit_opts.newton)
                                 was below 1e-6 max(1, |x_i|)                */

/* Maximum-likelihood fit from the start x (np values, moved into [lo, hi]):
 * bounded BFGS (active set, projected line search).  Replaces
 * ming2(fout, &lnL, CUDA_lfundG, NULL, x, xb, space, e, np) at codeml.c:688
 * (ming2: tools.c:6361-6545).  On return x holds the estimates, *neg_lnl
 * = -lnL there, and info (may be NULL) = {iterations, evaluations, stop
 * reason BSA_STOP_*}; the engine's last evaluation is at x, so
 * bsa_class_lnl / bsa_last_params describe the fit.  A variable with
 * lo == hi stays fixed at that value. */
int  bsa_fit(bsa_engine *eng, double *x, int np, const double *lo, const double *hi,
             const bsa_fit_opts *opts, double *neg_lnl, int *info3);

/* diagnostic: the same optimiser on a caller-supplied objective (returns
 * nonzero where it is undefined), for testing it without a GPU; info as in
 * bsa_fit.  nblock (0..n): the last nblock variables, whose Hessian block
 * Newton iterations on more than 160 free variables form exactly, as the
 * preconditioner of their conjugate-gradient step (bsa_fit names codeml's
 * model parameters, after the branch lengths).  No counterpart in the
 * reference. */
typedef int (*bsa_objective)(const double *x, double *f, double *grad, void *ctx);
int  bsa_debug_minimize(int n, bsa_objective fn, void *ctx, double *x, const double *lo,
                        const double *hi, const bsa_fit_opts *opts, int nblock, double *fmin,
                        int *info3);

/* ------------------------------------------------------------------------
 * Batch engine: many genes, each with its own tree, alignment and model
 * options, evaluated together on one GPU (the genome scan: the reference
 * runs codeml once per gene, one process each, codeml.c:357; each gene's
 * likelihood is the one bsa_lnl computes).  The genes' jobs share one set
 * of device buffers, so one evaluation of all of them costs one launch per
 * tree level of the deepest tree; patterns are padded to the largest gene's
 * count, so genes of similar size batch best.  Arrays of per-gene pointers
 * are indexed by gene (creation order).  No counterpart in the reference.
 * ------------------------------------------------------------------------ */
typedef struct bsa_batch bsa_batch;

/* all problems share device, precision, kernel and tile */
int  bsa_batch_create(const bsa_problem *const *probs, int ngene, bsa_batch **out);
int  bsa_batch_destroy(bsa_batch *batch);
int  bsa_batch_size(const bsa_batch *batch);                    /* genes, or -1 */
/* model options of one gene, or of every gene with gene < 0 */
int  bsa_batch_set_model(bsa_batch *batch, int gene, const bsa_model_opts *opts);
int  bsa_batch_num_params(const bsa_batch *batch, int gene);     /* np, or -1 */
int  bsa_batch_default_bounds(const bsa_batch *batch, int gene, double *lo, double *hi);
/* -lnL of every gene at x[gene] (np of that gene values each) */
int  bsa_batch_lnl(bsa_batch *batch, const double *const *x, double *neg_lnl);
/* -lnL and d(-lnL)/dx of every gene (adjoint sweep, as bsa_lnl_grad) */
int  bsa_batch_lnl_grad(bsa_batch *batch, const double *const *x, double *neg_lnl,
                        double *const *grad);
/* the same for the listed genes only (n ascending gene indices; x, grad and
 * neg_lnl indexed by list entry; grad may be NULL for -lnL alone): only
 * their jobs are launched; the other genes' last results are kept */
int  bsa_batch_eval_genes(bsa_batch *batch, const int *genes, int n, const double *const *x,
                          double *neg_lnl, double *const *grad);
/* ML fits of every gene from x[gene] (bsa_fit's optimiser, adjoint
 * gradient only), one optimiser thread per gene being fitted (at most 256
 * at once, the next gene starting when a fit ends); their requests are
 * served in rounds, each evaluating the genes whose points are ready
 * (continuous batching), so a gene costs device time only for its own
 * points.  lo/hi (either, or any entry, may be NULL:
 * codeml's bounds); info[3*gene .. 3*gene+2] as bsa_fit's info (may be
 * NULL); every gene's last evaluation is left at its estimates. */
int  bsa_batch_fit(bsa_batch *batch, double *const *x, const double *const *lo,
                   const double *const *hi, const bsa_fit_opts *opts, double *neg_lnl,
                   int *info);
/* The per-gene likelihood-ratio fits of a genome scan on the batch engine:
 * for every gene, H1 (model A, w2 free) from x1[gene], then H0 (w2 fixed
 * at 1, codeml.c:1648-1649) from the H1 estimates, then, where the H0
 * maximum lies above the H1 one by more than 1e-6 in lnL, H1 again from
 * the H0 estimates with w2 = 1, keeping the higher H1 maximum.  This is the
 * sequence codeml's users run as two control files per gene (the reference
 * has no single entry for it).  Each gene's chain runs in its own optimiser
 * thread and does not wait for other genes' fits; requests are served in
 * rounds as in bsa_batch_fit, with codeml's bounds.  Every gene's model
 * options are set to H1 (fix_omega = 0) first and are H1 on return.
 * x1[gene] (np of H1): start in, H1 estimates out; x0[gene] (np - 1): H0
 * estimates out.  neg_lnl[2*gene] = -lnL at the H1 estimates,
 * neg_lnl[2*gene+1] at the H0 ones.  info[7*gene ..] (may be NULL): H1
 * iterations and evaluations (both H1 fits summed), the stop reason of the
 * H1 fit whose maximum was kept, H0 iterations, evaluations, stop reason,
 * and 1 where H1 was fitted again.  Every gene's
 * last evaluation is left at its H1 estimates. */
int  bsa_batch_fit_lrt(bsa_batch *batch, double *const *x1, double *const *x0,
                       const bsa_fit_opts *opts, double *neg_lnl, int *info);
/* bsa_tune for a batch, at x[gene] for every gene */
int  bsa_batch_tune(bsa_batch *batch, const double *const *x, int reps, double *out7);
/* per-gene outputs of the last evaluation, as bsa_class_lnl / bsa_last_params */
int  bsa_batch_class_lnl(bsa_batch *batch, int gene, double *fhK, double *site_lnf);
int  bsa_batch_last_params(const bsa_batch *batch, int gene, double *out11);
/* as bsa_timing / bsa_work / bsa_set_graph / bsa_set_prune, for the whole
 * batch */
int  bsa_batch_timing(bsa_batch *batch, double *out17, int reset);
int  bsa_batch_work(bsa_batch *batch, double *out10);
int  bsa_batch_set_graph(bsa_batch *batch, int enable);
int  bsa_batch_set_prune(bsa_batch *batch, int mode);
/* diagnostic, host only: plan a batch of the given genes without a device
 * and check every index of its schedules against the buffers they index;
 * out8 = {genes, padded patterns, partial slots, pruning launches, pruning
 * jobs, adjoint jobs, eigen-basis jobs, branch-matrix chunks} */
int  bsa_debug_forest_plan(const bsa_problem *const *probs, int ngene, double *out8);

/* ------------------------------------------------------------------------
 * Likelihood of one gene that allocates no device memory.  No counterpart in
 * the reference: it is for callers that cannot call hipMalloc (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * The same computation as bsa_create + bsa_lnl + bsa_class_lnl (bitwise the
 * same results), with every device buffer in one array of BSA_NOALLOC_BYTES
 * that the library holds on each GPU it runs on until the process ends; a
 * gene whose buffers do not fit (beyond about 64 taxa x 1,000 patterns in
 * FP64, twice the patterns in FP32) is refused with BSA_EINVAL, and the error text
 * gives the bytes it needs.  All device work runs on `stream` (a hipStream_t;
 * NULL: the null stream) and has finished when the call returns.  Calls on
 * one device are serialized by the library (they share that device's array,
 * which keeps the last call's contents).  fhK [4*npatt] and site_lnf
 * [npatt] (may be NULL) are as bsa_class_lnl's; opts NULL means H1 with
 * transformed p's.
 * ------------------------------------------------------------------------ */
#define BSA_NOALLOC_BYTES ((size_t)128 << 20)
int  bsa_lnl_noalloc(const bsa_problem *prob, const bsa_model_opts *opts, const double *x, int np,
                     double *neg_lnl, double *fhK, double *site_lnf, void *stream);

const char *bsa_last_error(void);

/* diagnostic: one 16x16x4 matrix-core product in FP64 and in FP32 of
 * A (16x4, row-major) and B (4x16) on `device`; D64/D32 get each lane's four
 * accumulator values, lane-major [64][4] (no counterpart in the reference;
 * used by the tests to pin the lane layout the pruning kernel assumes) */
int  bsa_debug_mfma_probe(int device, const double *A, const double *B, double *D64,
                          double *D32);
const char *bsa_version(void);

#ifdef __cplusplus
}
#endif
#endif
