// Device-side data layout and kernel launchers (internal header).
#pragma once
#include <hip/hip_runtime.h>

namespace bsa {

constexpr int kPadP = 128;      // patterns are padded to a multiple of this
constexpr int kMaxCh = 8;       // children per node handled by one job

// One (node, site class) job of the pruning pass: multiply the child
// contributions together, rescale per pattern by a power of two, store.
struct PruneJob {
    int out_slot;               // partial slot written
    int nch;
    int ch_tip[kMaxCh];         // tip index, or -1 for an interior child
    int ch_slot[kMaxCh];        // partial slot of an interior child
    int ch_pmat[kMaxCh];        // P(t) matrix index (node * 3 + eigen set)
};

// One P(t) matrix to build: branch above `node`, eigen set `set`.
struct PmatJob {
    int node, set, pm;          // pm = node * 3 + set
    int tip;                    // tip index if the node is a tip, else -1
    double t;                   // effective branch length t * Qfactor
};

struct RootArgs {
    int npatt, npad;
    int root_slot[4];           // root partial slot per class
    double freqK[4];
};

// launchers; T = storage type of the conditional likelihoods (double or float)
void launch_pmat(hipStream_t s, const double *uvr, const PmatJob *jobs, int njobs, double *PT,
                 const unsigned char *mask_amb, int namb, double *AT);

// kernel 0 = MFMA (tile 64 or 128 patterns per block), 1 = VALU (tile 64)
template <typename T>
void launch_prune(hipStream_t s, int kernel, int tile, const PruneJob *jobs, int njobs, int npad,
                  const unsigned char *z, const double *PT, const double *AT, int namb, T *L,
                  int *E);

template <typename T>
void launch_root(hipStream_t s, const RootArgs &a, const T *L, const int *E, const double *pi,
                 const double *fpatt, double *fh, double *site, double *partial, int nblk,
                 double *out);

// raw MFMA lane-layout probe: one wave, writes each lane's accumulator
void launch_mfma_probe(hipStream_t s, const double *A, const double *B, double *D64,
                       const float *Af, const float *Bf, float *D32);

}  // namespace bsa
