// Device-side data layout and kernel launchers (internal header).
#pragma once
#include <hip/hip_runtime.h>

namespace bsa {

constexpr int kTileP = 64;      // patterns per pruning tile
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

// launchers; T = double or float
template <typename T>
void launch_pmat(hipStream_t s, const double *uvr, const PmatJob *jobs, int njobs, T *PT,
                 const unsigned char *mask_amb, int namb, T *AT);

template <typename T>
void launch_prune(hipStream_t s, int kernel, const PruneJob *jobs, int njobs, int npad,
                  const unsigned char *z, const T *PT, const T *AT, int namb, T *L, int *E);

template <typename T>
void launch_root(hipStream_t s, const RootArgs &a, const T *L, const int *E, const double *pi,
                 const double *fpatt, double *fh, double *site, double *partial, int nblk,
                 double *out);

// raw MFMA lane-layout probe: one wave, writes each lane's accumulator
void launch_mfma_probe(hipStream_t s, const double *A, const double *B, double *D64,
                       const float *Af, const float *Bf, float *D32);

}  // namespace bsa
