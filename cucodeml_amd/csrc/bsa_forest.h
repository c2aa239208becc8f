// Internal header: the host description of one gene (tree, compressed
// alignment, its pruning and adjoint schedules in its own node numbering,
// model cache) and the device state that evaluates one or more genes with
// one launch sequence (a "forest": the genes' jobs renumbered into shared
// buffers, one launch per tree level across all genes).  The single-gene
// engine is a forest of one gene; the batch engine a forest of many.
#pragma once
#include "bsa.h"
#include "bsa_kernels.h"
#include "bsa_model.h"

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <functional>
#include <shared_mutex>
#include <vector>

namespace bsa {

// sets the calling thread's error text (bsa_last_error) and returns code
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error_text();

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::bsa::fail(BSA_EDEVICE, "%s: %s (%s:%d)", #call,                  \
                               hipGetErrorString(e_), __FILE__, __LINE__);            \
    } while (0)

// A graph capture must not see HIP calls of other threads that would
// invalidate it (allocations, synchronous copies): every entry point that
// calls HIP holds this lock shared, a capture holds it exclusively.
extern std::shared_mutex g_hip_mu;
#define HIP_SHARED std::shared_lock<std::shared_mutex> hip_shared_lock_(::bsa::g_hip_mu)

struct Gene {
    int ns = 0, nnode = 0, root = 0, nbranch = 0, npatt = 0, ncodes = 0, namb = 0;
    std::vector<int> father, label, ibranch, height;
    std::vector<std::vector<int>> sons;
    std::vector<char> onpath;                 // proper ancestors of a foreground branch
    std::vector<double> pi, fpatt;
    std::vector<unsigned char> z;             // ns x npatt codes
    std::vector<unsigned char> code_mask;     // ncodes x 61
    bsa_model_opts opts{0, 1.0, 1};

    // forward schedule, own numbering: P(t) jobs, pruning jobs by level
    std::vector<PmatJob> pjobs;
    std::vector<PruneJob> prune;
    std::vector<int> level_off;               // level l: prune[level_off[l] .. level_off[l+1])
    int nslot = 0;
    std::vector<std::array<int, 4>> slot;     // inside slot per node and class
    int root_slot[4] = {-1, -1, -1, -1};
    // pruning work (out[0..5] of bsa_work) and gradient work (out[6..9])
    double work[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

    // adjoint plan, own numbering; adj[i].m_index = i
    bool adj_planned = false;
    std::vector<AdjJob> adj;                  // grouped by depth of the child node
    std::vector<int> adj_off;                 // depth d: adj[adj_off[d-1] .. adj_off[d])
    std::vector<int> adj_nint;                // depth d: the first adj_nint[d-1] have an interior child
    std::vector<EigJob> ejobs;                // set = eigen set 0..2 of the gene
    std::vector<int> ejob_node;               // node of each EigJob
    int root_aslot[4] = {-1, -1, -1, -1};

    // place in its forest
    int gi = 0;                               // gene index
    int tip_base = 0, int_base = 0, slot_base = 0, pj_base = 0, ej_base = 0;
    int amb_base = 0, mask_base = 0;
    // its jobs in the forest's arrays, per launch: pruning level l, adjoint
    // depth d (interior children, then tips), [first, end)
    std::vector<std::array<int, 2>> prune_rng, adj_int_rng, adj_tip_rng;

    // model cache and last evaluation
    ModelA model{};
    bool have_eigen = false;
    double cache_kappa = NAN, cache_w0 = NAN, cache_w2 = NAN;
    double last_lnl = NAN;
    bool evaluated = false;
};

// copy and check one gene's tree and data, build its forward schedule;
// `wbytes` = bytes per stored partial (work accounting)
int gene_init(Gene &g, const bsa_problem *p, int wbytes);
// plan the adjoint sweep of a gene (host only)
void gene_plan_adjoint(Gene &g, int wbytes);
int gene_num_params(const Gene &g);
// codeml's variables -> natural parameters (f_and_x, tools.c:1280-1305)
void natural_params(const Gene &g, const double *x, double *kappa, double *p0, double *p1,
                    double *w0, double *w2);
void default_bounds(const Gene &g, double *lo, double *hi);

struct Forest {
    int dev = 0, precision = BSA_FP64, kernel = BSA_KERNEL_MFMA, tile = 64;
    int G = 0, npad = 0, ncodes = kS, ns_total = 0, nint_total = 0, nslot = 0, namb_rows = 0;
    int nmask_rows = 0;
    std::vector<Gene *> genes;

    // combined forward schedule
    std::vector<PmatJob> pjobs;
    std::vector<PruneJob> prune;
    std::vector<int> level_off;
    std::vector<int> level_maxint;            // most interior children of a job, per level
    // persistent pruning (one launch for every level, k_prune_walk): each
    // gene's jobs in level order (wjobs[woff[g] .. woff[g + 1]), indices into
    // prune), the genes by falling job count (walk_order), the items of the
    // whole forest (gene * ntile + tile, genes in walk_order)
    std::vector<int> wjobs, woff, walk_order, witems;
    int prune_walk = 0;                       // 1: pruning by k_prune_walk
    int walk_nblk = 0;                        // its grid: 2 blocks per CU

    // device
    hipStream_t stream = nullptr;
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    bool use_graph = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};   // lnL only, with gradient
    // model blocks of the genes (kGeneBlock words each)
    double *d_uvr = nullptr;
    PmatJob *d_pjobs = nullptr;
    PruneJob *d_prune = nullptr;
    int *d_wjobs = nullptr, *d_woff = nullptr, *d_witems = nullptr;
    RootArgs *d_rargs = nullptr;                   // [2][G]: without, with adjoint seeds
    void *d_PT = nullptr, *d_AT = nullptr, *d_L = nullptr;
    int *d_E = nullptr;
    unsigned char *d_z = nullptr, *d_mask_amb = nullptr;
    double *d_pi = nullptr, *d_fpatt = nullptr, *d_fh = nullptr, *d_site = nullptr;
    double *d_partial = nullptr, *d_out = nullptr;
    int nroot_blk = 0;

    // host staging (pinned)
    double *h_uvr = nullptr;                       // same layout as d_uvr
    double *h_out = nullptr;                       // [G][kRootSums]
    PmatJob *h_pjobs = nullptr;
    // job lists of a round (forest_run_round): pinned host and device copies
    int *h_sel = nullptr, *d_sel = nullptr;
    size_t sel_cap = 0;

    // pmat, prune, root, child adjoints, branch matrices, adjoint finish, wall
    double timing[7] = {0, 0, 0, 0, 0, 0, 0};
    double timing_sum[7] = {0, 0, 0, 0, 0, 0, 0};
    double n_eval = 0, n_launch = 0, n_adj_launch = 0;
    double work[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

    // adjoint (set up on the first gradient)
    bool adj_ready = false;
    std::vector<AdjJob> adj;
    std::vector<int> adj_off, adj_nint;
    std::vector<EigJob> ejobs;
    std::vector<int> set_off, set_list;
    int nchunk = 0, tpc = 0, nmrow = 0;
    // launch shapes the tuner may set: child-adjoint tiles per block (0: by
    // launch size), branch-matrix tiles per chunk (0: by gene size); chunks
    // the branch-matrix partials have room for
    int tpb_set = 0, tpc_set = 0, mpart_chunks = 0;
    AdjJob *d_adj = nullptr;
    EigJob *d_ejobs = nullptr, *h_ejobs = nullptr;
    int *d_set_off = nullptr, *d_set_list = nullptr;
    void *d_A = nullptr, *d_C = nullptr;
    int *d_EA = nullptr;
    double *d_Mpart = nullptr, *d_W = nullptr, *d_G = nullptr, *d_dtau = nullptr, *d_Z = nullptr;
    double *d_mask64 = nullptr;
    // per tip row (ns_total x npad): the tip's patterns ordered by its code
    // (the order the branch-matrix kernel visits a tip's patterns in)
    int *d_tperm = nullptr;
    double *h_adjout = nullptr;                    // dtau[nej], then Z[3 G][4096]

    size_t model_block() const { return (size_t)G * kGeneBlock; }
};

// renumber the genes' jobs into one forest (host only)
int forest_plan(Forest &f, const std::vector<Gene *> &genes, int precision, int kernel, int tile);
// plan the adjoint sweep of every gene and renumber it (host only)
void forest_plan_adjoint(Forest &f);
// every index of the planned schedules against the buffer it indexes
int forest_check(const Forest &f);
// allocate and fill the device buffers of a planned forest
int forest_upload(Forest &f, int device);
// plan + check + upload; genes must outlive the forest
int forest_build(Forest &f, const std::vector<Gene *> &genes, int device, int precision,
                 int kernel, int tile);
void forest_free(Forest &f);
// plan (per gene), check, renumber and allocate the adjoint sweep
int forest_setup_adjoint(Forest &f);
// in graph mode, capture the evaluation's graph before its first use
int forest_prepare_graph(Forest &f, bool adjoint);
// host part of gene g's evaluation at natural parameters (branch lengths
// from x): model set-up, staged into the forest's per-evaluation buffers.
// qf_fixed, when given, replaces the two Qfactors the model would give.
int gene_stage(Gene &g, Forest &f, const double *x, double kappa, double p0, double p1, double w0,
               double w2, const double *qf_fixed, bool adjoint);
int gene_stage_x(Gene &g, Forest &f, const double *x, bool adjoint);
// device work of one evaluation of every gene (inputs staged), results
// copied back; then gene_collect reads gene g's lnL
int forest_run(Forest &f, bool adjoint);
// an evaluation with gradient in two parts (inputs staged with adjoint):
// part 1 the forward pass, keeping what the adjoint sweep needs (gene_collect
// then reads lnL); part 2 the sweep of the last part 1 (then
// gene_chain_rule).  Bitwise the results of forest_run(f, true); eager
// launches only.  A fit asks for part 2 only at the points it accepts.
int forest_run_part(Forest &f, int part);
// the same for the listed genes only (ascending gene indices, inputs
// staged): their jobs alone are launched and only their inputs and outputs
// cross PCIe; other genes' staging buffers may be written meanwhile.  Eager
// launches (no graph).
int forest_run_round(Forest &f, const std::vector<int> &genes, bool adjoint);
// a round of forward passes for the genes fwd and adjoint sweeps for the
// genes adj (ascending gene indices each; either list may be empty).  With
// keep the forward passes store what a sweep needs; a gene in adj gets the
// sweep of its last forward pass run with keep: this round's where it is in
// fwd too, else an earlier round's (gene regions of the device buffers are
// disjoint, so other genes' rounds leave it intact).
int forest_run_round(Forest &f, const std::vector<int> &fwd, const std::vector<int> &adj,
                     bool keep);
double gene_collect(Gene &g, const Forest &f);
// d(-lnL)/dx of gene g from the adjoint outputs of the last forest_run
void gene_chain_rule(const Gene &g, const Forest &f, const double *x, double *grad);

// Bytes of device memory a planned forest's likelihood pass takes when its
// buffers are laid out one after another (256-byte aligned) from `base`; with
// base != nullptr the forest's device pointers are set to that layout.
size_t forest_place_lnl(Forest &f, char *base);
// One likelihood evaluation of a planned, checked one-gene forest at codeml's
// x, with every device buffer inside [base, base + cap) (device memory the
// caller owns) and all work on stream s; nothing is allocated on the device.
// fhK [4 npatt] / site [npatt] may be null.  Refuses a gene that does not
// fit.  Afterwards f holds no buffers (nothing to free).
int forest_lnl_in_place(Forest &f, char *base, size_t cap, hipStream_t s, const double *x,
                        double *lnl, double *fhK, double *site);

// Launch shapes: pruning tile (64 or 128 patterns), child-adjoint tiles per
// block (0 = by launch size, else 1..8), branch-matrix tiles per chunk (0 = by
// gene size, up to kAdjTilesPerChunkAuto, else 1..kAdjTilesPerChunk).
// Captured graphs are dropped.
int forest_set_shapes(Forest &f, int tile, int tpb, int tpc);
// Pruning by one launch per level (BSA_PRUNE_LEVELS) or by the persistent
// walk (BSA_PRUNE_WALK, FP64-product kernels only); captured graphs are
// dropped.
int forest_set_prune(Forest &f, int mode);
// Autotuner: for each parameter in turn (pruning mode, pruning tile,
// child-adjoint tiles per block, branch-matrix tiles per chunk) every
// candidate is timed over `reps` calls of eval (one evaluation with
// gradient each) by the device time of the evaluation, and the fastest is
// kept.  out[7] = {tile, tpb, tpc, chunks, best ms, ms of the shapes it
// started from, pruning mode}.
int forest_tune(Forest &f, const std::function<int()> &eval, int reps, double *out7);

}  // namespace bsa
