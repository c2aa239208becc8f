// Device-side data layout and kernel launchers (internal header).
#pragma once
#include <hip/hip_runtime.h>

namespace bsa {

constexpr int kPadP = 128;      // patterns are padded to a multiple of this
constexpr int kMaxCh = 8;       // children per node handled by one job

// Model block of one gene in device memory: its three eigen sets (U and V
// 64 x 64 each, then the 64 roots; sets for omega = w0, 1, w2), then 8 words
// whose first four are the class proportions.  Gene g's block starts at
// g * kGeneBlock; eigen set s of gene g is forest set 3 g + s.
constexpr int kUVRSet = 64 * 64 * 2 + 64;
constexpr int kGeneBlock = 3 * kUVRSet + 8;
__host__ __device__ inline size_t uvr_offset(int set)
{
    return (size_t)(set / 3) * kGeneBlock + (size_t)(set % 3) * kUVRSet;
}

// Launchers take an optional list `sel` of the jobs (or genes, or eigen
// sets) to run: block i runs job sel[i]; with sel = nullptr, block i runs
// job i.

// One (node, site class) job of the pruning pass: multiply the child
// contributions together, rescale per pattern by a power of two, store.
struct PruneJob {
    int out_slot;               // partial slot written
    int nch;
    int ch_tip[kMaxCh];         // tip index, or -1 for an interior child
    int ch_slot[kMaxCh];        // partial slot of an interior child
    int ch_pmat[kMaxCh];        // P(t) matrix index (node * 3 + eigen set)
    int ch_cslot[kMaxCh];       // slot of the child's contribution P I in the
                                // contribution buffer (interior children), kept
                                // for the adjoint sweep when that buffer is given
};

// One P(t) matrix to build: branch above `node`, eigen set `set`.
struct PmatJob {
    int node, set, pm;          // pm = node * 3 + set
    int tip;                    // tip index if the node is a tip, else -1
    int amb;                    // first row of this gene's ambiguity masks
    double t;                   // effective branch length t * Qfactor
};

// root pass of one gene (blockIdx.y of the root kernel); per-gene inputs and
// outputs sit at gene-indexed offsets of the buffers the launch gets
struct RootArgs {
    int npatt, npad;
    int root_slot[4];           // root partial slot per class
    int seed_aslot[4];          // adjoint slot of the root per class, or -1:
                                // no adjoint seeds are written
};

// quantities reduced over patterns by the root pass: lnL, then per class
// Sf_c = sum_h fpatt_h f_hc / f_h (= d lnL / d p_c)
constexpr int kRootSums = 5;

// One job of the adjoint (preorder) sweep: the branch above node v for one
// site class, or for a merged pair of classes (m, m+2) where v's partials and
// P(t) are shared by both.  Each term t (one or two) gives
// X_t = (adjoint of the parent, class of term t) x (contributions of v's
// siblings for that class); X = sum_t X_t is the derivative of lnL with
// respect to v's contribution to its parent.  From it the job writes v's
// adjoint (if v is interior) and the branch matrix M = sum_h X_h^T I_h over
// its chunk of patterns.
struct AdjJob {
    int v_tip;                  // tip index of v, or -1
    int v_mask;                 // first row of this gene's code-indicator table
    int v_slot;                 // inside partial slot of v (interior)
    int v_pmat;                 // P(t) matrix index of v's branch
    int v_aslot;                // adjoint slot written for v (interior) or -1
    int m_index;                // row of the branch-matrix buffer
    int nterm;                  // 1 or 2
    int a_aslot[2];             // adjoint slot of the parent, per term
    int nsib;
    int sib_tip[kMaxCh];
    int sib_slot[2][kMaxCh];    // inside slot of an interior sibling (its exponent), per term
    int sib_cslot[2][kMaxCh];   // contribution slot of an interior sibling, per term
    int sib_pmat[2][kMaxCh];    // P(t) of a tip sibling, per term
};

// per (branch, class) eigen-basis job: M -> (d lnL / d tau, A o X)
struct EigJob {
    int m_index;                // branch-matrix row
    int set;                    // eigen set (0: w0, 1: omega = 1, 2: w2)
    double tau;                 // t * Qfactor
};

constexpr int kAdjTilesPerChunk = 64;   // most 64-pattern tiles per branch-matrix block
constexpr int kAdjTilesPerChunkAuto = 32;   // most of them when chosen by the gene's size
constexpr int kWParts = 8;              // partial sums per eigen set in the W reduction

// the same on the FP32 matrix cores (float storage, BSA_KERNEL_MFMA32)
void launch_adjoint_level32(hipStream_t s, const AdjJob *jobs, const int *sel, int njobs, int npad,
                            const unsigned char *z, const double *PT, const double *TT, int ncodes,
                            const int *E, const float *Cb, float *A, int *EA, int tpb);

// adjoints of one depth's interior children (jobs with v_tip < 0 only);
// tpb: 64-pattern tiles per block, 0 for a choice by the launch's size
template <typename T>
void launch_adjoint_level(hipStream_t s, const AdjJob *jobs, const int *sel, int njobs, int npad,
                          const unsigned char *z, const double *PT, const double *TT, int ncodes,
                          const int *E, const T *Cb, T *A, int *EA, int tpb);

// branch-matrix chunk partials of all jobs (after the sweep); m32 (float
// storage only): interior children's products on the FP32 matrix cores;
// tperm: per tip row, the order in which the tip's patterns are visited (a
// permutation of 0..npad-1)
template <typename T>
void launch_adjoint_m(hipStream_t s, const AdjJob *jobs, const int *sel, int njobs, int npad,
                      int nchunk, int tpc, bool m32, const unsigned char *z, const double *PT,
                      const double *TT, int ncodes, const double *mask64, const int *tperm,
                      const T *L, const int *E, const T *Cb, const T *A, const int *EA,
                      double *Mpart);

// for every EigJob sums the chunk partials of its branch matrix M, forms
// A = U^T M V^T, d lnL/d tau and G = A o X(tau); then per eigen set
// W = sum G (kWParts partials in Wpart[kWParts][nsets][64*64]) and
// Z = V^T W U^T.  Outputs dtau[job] and Z[set][64*64] (3 sets per gene) for
// the selected jobs (esel) and sets (ssel).
void launch_adjoint_finish(hipStream_t s, const double *Mpart, int nchunk, const EigJob *ejobs,
                           const int *esel, int nej, const int *ssel, int nsets,
                           const int *set_off, const int *set_list, const double *uvr, double *G,
                           double *dtau, double *Wpart, double *Z);

// launchers; T = storage type of the conditional likelihoods (double or float)
void launch_pmat(hipStream_t s, const double *uvr, const PmatJob *jobs, const int *sel, int njobs,
                 double *PT, const unsigned char *mask_amb, int ncodes, double *TT);

// kernel 0 = MFMA (tile 64 or 128 patterns per block), 1 = VALU (tile 64),
// 2 = FP32 matrix cores (float storage; 128-pattern blocks, tile ignored)
// Cb: contribution buffer (interior children's P I per class) or nullptr;
// maxint: the most interior children of any job in the launch
template <typename T>
void launch_prune(hipStream_t s, int kernel, int tile, const PruneJob *jobs, const int *sel,
                  int njobs, int maxint, int npad, const unsigned char *z, const double *PT,
                  const double *TT, int ncodes, T *L, int *E, T *Cb);

// every pruning level of a forest in one persistent launch: nblk blocks take
// the items (gene * ntile + tile, 128-pattern tiles, ntile = npad / 128) in
// turn, each walking its gene's jobs wjobs[woff[g] .. woff[g + 1]) (indices
// into jobs, in level order) for its tile; FP64 matrix-core products, the
// results of the level launches bitwise
template <typename T>
void launch_prune_walk(hipStream_t s, const PruneJob *jobs, const int *wjobs, const int *woff,
                       const int *items, int nitems, int nblk, int npad,
                       const unsigned char *z, const double *PT, const double *TT, int ncodes,
                       T *L, int *E, T *Cb);

// per-pattern class values, mixture, reduced sums (out[kRootSums]); with
// seed_aslot[0] >= 0 also the adjoint seeds of the root (A, EA).  args:
// device array of RootArgs per gene; gene g (gsel[i] for block row i, or i)
// reads pi[g*64], its class proportions in its model block (device memory,
// so that a captured graph reads the current ones), fpatt[g*npad], writes
// fh[g*4*npad], site[g*npad], partial[g*nblk*2*kRootSums] and
// out[g*kRootSums]
template <typename T>
void launch_root(hipStream_t s, const RootArgs *args, const int *gsel, int ngene, const T *L,
                 const int *E, const double *pi, const double *uvr, const double *fpatt, double *fh,
                 double *site, double *partial, int nblk, double *out, T *A, int *EA);

// raw MFMA lane-layout probe: one wave, writes each lane's accumulator
void launch_mfma_probe(hipStream_t s, const double *A, const double *B, double *D64,
                       const float *Af, const float *Bf, float *D32);

// Bayes empirical Bayes: per-pattern normalisation of the w-set table
// F[nset][npad] (log f in, exp(log f - max) out, max in mx); log f(X | g)
// for every grid point; per-pattern class posteriors over the grid points in
// glist with weights W.  pc/iw: [ngrid][4] class proportions and set indices.
void launch_beb_norm(hipStream_t s, double *F, int nset, int npatt, int npad, double *mx);
void launch_beb_lnfx(hipStream_t s, const double *F, int npad, int npatt, const double *fpatt,
                     const double *pc, const int *iw, int ngrid, double *lnfx);
void launch_beb_post(hipStream_t s, const double *F, int npad, int npatt, const double *pc,
                     const int *iw, const int *glist, const double *W, int ng, double *post);

}  // namespace bsa
