// HIP kernels for gfx950 (MI355X) of the branch-site model A likelihood:
//   k_pmat        batched P(t) - I = U diag(exp(root t)) V - I, every (branch, eigen set)
//   k_tiptab      per-tip rows of every code (sense codons and ambiguity codes)
//   k_prune_mfma  one tree level of Felsenstein pruning on the FP64 matrix cores
//   k_prune_res   the same with the children's P resident in LDS (big levels)
//   k_prune_valu  the same on plain vector FMA
//   k_prune32     the same on the FP32 matrix cores, each pattern row's two largest
//                 child terms and the identity term in FP64 (BSA_KERNEL_FP32MC)
//   k_root        per-pattern class log-likelihoods, mixture and weighted sum
//   k_adj_a       adjoints of one depth's interior children (preorder sweep)
//   k_adj_a32     the same on the FP32 matrix cores (BSA_KERNEL_MFMA32)
//   k_adj_m       branch matrices of every (branch, class) job (interior children on
//                 the FP32 matrix cores with BSA_KERNEL_MFMA32)
//   k_eigbranch, k_wreduce, k_zform   eigen-basis derivatives and their sums
//   k_beb_*       Bayes empirical Bayes over the 10^4-point grid
//
// Arithmetic is FP64 (k_adj_a32, k_prune32 and the FP32 branch-matrix path
// aside); T (double or float) is only the storage type of the conditional
// likelihoods.  Layouts:
//   PT[pm][k][j]    = P_{j,k}(t) - delta_{jk} (double) for matrix pm = node*3 + set,
//                     64 x 64, rows and columns >= 61 zero (k: child state,
//                     j: parent state).  P - I is computed without the
//                     cancellation that P itself would suffer for short
//                     branches.  The FP64 matrix-core kernels add the identity
//                     when they stage P in LDS and multiply the full P;
//                     k_prune_valu and k_adj_a32 add it after the product, so
//                     that the small off-diagonal terms are summed on their
//                     own scale (in FP32 the near-1 diagonal would swamp them).
//   L[slot][h][j]   (T) conditional likelihood of a node / site class, pattern
//                   h, parent-side state j (64, lanes >= 61 stay exactly zero)
//   E[slot][h]      accumulated power-of-two exponent: true value = L * 2^E
// FP32 storage rounds stochastically with a hash of (slot, pattern, state) as
// the random bits: unbiased, uncorrelated across patterns that share
// sub-trees, and bitwise repeatable.
#include <algorithm>

#include "bsa_kernels.h"

#include <math.h>
#include <stdint.h>

namespace bsa {

typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr double kLn2 = 0.69314718055994530942;
constexpr double kLogFloor = -690.77552789821368;   // log(1e-300)
constexpr int kRootPerWave = 4;    // patterns per wave in k_root (16 per block)

// ------------------------------------------------------------------ P(t)

// One block per (branch, set).  U is staged transposed so that the 64 lanes
// of a wave (64 consecutive parent states j) read consecutive LDS words; the
// wave's child-state index is uniform, so V reads are broadcasts.
__global__ __launch_bounds__(256) void k_pmat(const double *__restrict__ uvr,
                                              const PmatJob *__restrict__ jobs,
                                              const int *__restrict__ sel,
                                              double *__restrict__ PT)
{
    __shared__ double Ut[64 * 64];   // Ut[m][j] = U[j][m]
    __shared__ double Ve[64 * 64];   // Ve[m][k] = exp(root_m t) V[m][k]
    __shared__ double eR[64];        // exp(root_m t)
    const PmatJob jb = jobs[sel ? sel[blockIdx.x] : blockIdx.x];
    const double *U = uvr + uvr_offset(jb.set);
    const double *V = U + 4096;
    const double *R = U + 8192;
    double *out = PT + (size_t)jb.pm * 4096;
    const int tid = threadIdx.x;
    if (jb.t < 1e-100) {   // PMatUVRoot returns the identity (tools.c:527-530)
        for (int q = tid; q < 4096; q += 256) out[q] = 0.0;
        return;
    }
    if (tid < 64) eR[tid] = exp(jb.t * R[tid]);   // 64 exponentials, not one per element
    __syncthreads();
    for (int q = tid; q < 4096; q += 256) {
        int r = q >> 6, c = q & 63;
        Ut[c * 64 + r] = U[q];
        Ve[q] = V[q] * eR[r];
    }
    __syncthreads();
    const int j = tid & 63, kq = tid >> 6;
    double acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = 0;
    for (int m = 0; m < 64; m++) {
        double u = Ut[m * 64 + j];
#pragma unroll
        for (int q = 0; q < 16; q++) acc[q] = fma(u, Ve[m * 64 + kq + 4 * q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int k = kq + 4 * q;
        double p = acc[q];
        if (p < 0) p = 0;   // negatives clamped (tools.c:536-538)
        if (k == j && k < 61) p -= 1.0;
        out[k * 64 + j] = p;
    }
}

// Tip table of every tip branch: TT[pm][c][j] = sum over the sense codons k
// code c allows of P_{j,k} (row c of P itself for a sense codon), so that a
// tip's contribution to parent state j is one load, whatever the code.
__global__ __launch_bounds__(256) void k_tiptab(const PmatJob *__restrict__ jobs,
                                                const int *__restrict__ sel,
                                                const double *__restrict__ PT,
                                                const unsigned char *__restrict__ mask_amb,
                                                int ncodes, double *__restrict__ TT)
{
    const PmatJob jb = jobs[sel ? sel[blockIdx.x] : blockIdx.x];
    if (jb.tip < 0) return;
    const double *P = PT + (size_t)jb.pm * 4096;
    double *out = TT + (size_t)jb.pm * ncodes * 64;
    const int j = threadIdx.x & 63;
    for (int c = threadIdx.x >> 6; c < ncodes; c += 4) {
        double s = 0;
        if (c < 61) {
            s = P[c * 64 + j] + (c == j ? 1.0 : 0.0);
        } else {
            const unsigned char *m = mask_amb + (size_t)(jb.amb + c - 61) * 64;
            for (int k = 0; k < 61; k++)
                if (m[k]) s += P[k * 64 + j] + (k == j ? 1.0 : 0.0);
        }
        out[c * 64 + j] = s;
    }
}

// ------------------------------------------------------------------ helpers

// contribution of a tip with code `code` to parent state j (tip table)
__device__ inline double tip_value(const double *TT, int pm, int ncodes, int code, int j)
{
    return TT[((size_t)pm * ncodes + code) * 64 + j];
}

__device__ inline uint32_t mix32(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t x = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u) * 0x85EBCA77u ^ c * 0xC2B2AE3Du;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

// one xorshift32 step (Marsaglia, J. Stat. Softw. 8(14), 2003): further
// random bits for the next value a lane stores after a hashed first one
__device__ inline uint32_t xorshift32(uint32_t x)
{
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

// random bits for the values a lane stores in one pattern row: a hash of
// (buffer, pattern, lane), never 0 (a fixed point of xorshift32)
__device__ inline uint32_t row_seed(uint32_t buf, uint32_t h, uint32_t lane)
{
    const uint32_t x = mix32(buf, h, lane);
    return x ? x : 0x6B43A9B5u;
}

// storage of conditional likelihoods
template <typename T> struct Store;
template <> struct Store<double> {
    static __device__ inline double get(const double *p) { return *p; }
    static __device__ inline void put(double *p, double v, uint32_t) { *p = v; }
    // N (8 or 16) consecutive values from an N*8-byte aligned address
    template <int N = 16>
    static __device__ inline void load16(const double *p, double *a)
    {
        const d2 *q = reinterpret_cast<const d2 *>(p);
#pragma unroll
        for (int i = 0; i < N / 2; i++) {
            d2 v = q[i];
            a[2 * i] = v.x;
            a[2 * i + 1] = v.y;
        }
    }
};
template <> struct Store<float> {
    static __device__ inline double get(const float *p) { return (double)*p; }
    // stochastic rounding of v >= 0: add 29 random bits below the float
    // mantissa of the double, then truncate (exact conversion afterwards)
    static __device__ inline void put(float *p, double v, uint32_t r)
    {
        unsigned long long u = (unsigned long long)__double_as_longlong(v);
        u += (unsigned long long)(r & 0x1FFFFFFFu);
        u &= ~0x1FFFFFFFull;
        *p = (float)__longlong_as_double((long long)u);
    }
    template <int N = 16>
    static __device__ inline void load16(const float *p, double *a)
    {
        const f4 *q = reinterpret_cast<const f4 *>(p);
#pragma unroll
        for (int i = 0; i < N / 4; i++) {
            f4 v = q[i];
            a[4 * i] = v.x;
            a[4 * i + 1] = v.y;
            a[4 * i + 2] = v.z;
            a[4 * i + 3] = v.w;
        }
    }
};

__device__ inline d4 mfma64(double a, double b, d4 c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ------------------------------------------------------------------ pruning (MFMA)

// Block = 4 waves, one (node, class) job, a tile of 64*RG patterns x 64 parent
// states; wave w owns 16-pattern row groups 64 rg + 16 w (rg < RG), so the P
// tile staged in LDS for a child serves 64*RG patterns.  Interior children
// are (16 x 64) x (64 x 64) products per row group: 16 k-steps x 4 column
// blocks of the FP64 16x16x4 MFMA with B = P from LDS (P - I as computed,
// identity added on staging).  The k order
// is free, so lane group g = lane/16 takes child states 16g..16g+15 at steps
// 0..15 and each lane reads 16 consecutive values of its pattern row straight
// from HBM.  Accumulator element i of column block b of lane l is pattern
// (row group base) + l/16 + 4i, parent state 16b + l%16.
template <typename T, int RG>
__global__ __launch_bounds__(256, RG == 1 ? 4 : 2) void k_prune_mfma(const PruneJob *__restrict__ jobs,
                                                    const int *__restrict__ sel, int npad,
                                                    const unsigned char *__restrict__ z,
                                                    const double *__restrict__ PT,
                                                    const double *__restrict__ TT, int ncodes,
                                                    T *__restrict__ L, int *__restrict__ E,
                                                    T *__restrict__ Cb)
{
    __shared__ double sP[64 * 64];
    const PruneJob &jb = jobs[sel ? sel[blockIdx.y] : blockIdx.y];
    const int h0 = blockIdx.x * (64 * RG);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    d4 acc[RG][4];
    int esum[RG][4];
#pragma unroll
    for (int q = 0; q < RG; q++)
#pragma unroll
        for (int b = 0; b < 4; b++) {
            acc[q][b] = d4{1, 1, 1, 1};
            esum[q][b] = 0;
        }
    const int nch = jb.nch;
    for (int c = 0; c < nch; c++) {
        const int pm = jb.ch_pmat[c];
        const double *P = PT + (size_t)pm * 4096;
        const int tip = jb.ch_tip[c];
        if (tip >= 0) {
            const unsigned char *zt = z + (size_t)tip * npad;
#pragma unroll
            for (int q = 0; q < RG; q++) {
                const int hb = h0 + 64 * q + 16 * w + g;
                int code[4];
#pragma unroll
                for (int i = 0; i < 4; i++) code[i] = zt[hb + 4 * i];
#pragma unroll
                for (int b = 0; b < 4; b++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        acc[q][b][i] *= tip_value(TT, pm, ncodes, code[i], 16 * b + r);
            }
        } else {
            const int slot = jb.ch_slot[c];
            const T *Lc = L + (size_t)slot * npad * 64;
            const int *Ec = E + (size_t)slot * npad;
            __syncthreads();
            for (int q = tid; q < 4096; q += 256) sP[q] = P[q] + ((q >> 6) == (q & 63) ? 1.0 : 0.0);
            __syncthreads();
#pragma unroll
            for (int q = 0; q < RG; q++) {
                const int hw = h0 + 64 * q + 16 * w;
                const int hb = hw + g;
                double a[16];
                Store<T>::load16(Lc + (size_t)(hw + r) * 64 + 16 * g, a);
                d4 tmp[4];
#pragma unroll
                for (int b = 0; b < 4; b++) tmp[b] = d4{0, 0, 0, 0};
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const double *prow = sP + (16 * g + kk) * 64 + r;
#pragma unroll
                    for (int b = 0; b < 4; b++) tmp[b] = mfma64(a[kk], prow[16 * b], tmp[b]);
                }
                if (Cb) {
                    const int cs = jb.ch_cslot[c];
                    T *Co = Cb + (size_t)cs * npad * 64;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        uint32_t rb = row_seed((uint32_t)cs ^ 0x3C6EF372u, (uint32_t)(hb + 4 * i),
                                               (uint32_t)r);
#pragma unroll
                        for (int b = 0; b < 4; b++) {
                            Store<T>::put(Co + (size_t)(hb + 4 * i) * 64 + 16 * b + r,
                                          tmp[b][i], rb);
                            rb = xorshift32(rb);
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < 4; b++) acc[q][b] *= tmp[b];
#pragma unroll
                for (int i = 0; i < 4; i++) esum[q][i] += Ec[hb + 4 * i];
            }
        }
    }
    // per-pattern power-of-two rescaling: the 64 states of a row live in the
    // 4 column blocks of the 16 lanes of one lane group
    T *Lo = L + (size_t)jb.out_slot * npad * 64;
#pragma unroll
    for (int q = 0; q < RG; q++) {
        const int hb = h0 + 64 * q + 16 * w + g;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double m = acc[q][0][i];
#pragma unroll
            for (int b = 1; b < 4; b++) m = acc[q][b][i] > m ? acc[q][b][i] : m;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                double o = __shfl_xor(m, off, 64);
                m = o > m ? o : m;
            }
            int e = 0;
            if (m > 0) frexp(m, &e);
            const int h = hb + 4 * i;
            uint32_t rb = row_seed((uint32_t)jb.out_slot, (uint32_t)h, (uint32_t)r);
#pragma unroll
            for (int b = 0; b < 4; b++) {
                Store<T>::put(Lo + (size_t)h * 64 + 16 * b + r, ldexp(acc[q][b][i], -e), rb);
                rb = xorshift32(rb);
            }
            if (r == 0) E[(size_t)jb.out_slot * npad + h] = esum[q][i] + e;
        }
    }
}

// The same job with the first two interior children's P resident in LDS: a
// block of 8 waves takes one (node, class) job and tpb tiles of 128
// patterns.  The P tiles are staged once; then each wave runs its 16-pattern
// rows of every tile with no further barrier, so the waves overlap one
// another's loads and products freely.  A job with more interior children
// (the unrooted root's three, multifurcations; launched with tpb = 1)
// stages the later ones in pairs over the first, between barriers.  At 128
// VGPRs and 64 KB of LDS two such blocks share a CU, so one block's staging
// overlaps the other's products.  Per pattern the arithmetic and the order
// of the children are those of k_prune_mfma, so results are the same
// bitwise.
constexpr int kResWaves = 8;

template <typename T>
__global__ __launch_bounds__(64 * kResWaves, 4) void k_prune_res(const PruneJob *__restrict__ jobs,
                                                                const int *__restrict__ sel,
                                                                int npad,
                                                                const unsigned char *__restrict__ z,
                                                                const double *__restrict__ PT,
                                                                const double *__restrict__ TT,
                                                                int ncodes, T *__restrict__ L,
                                                                int *__restrict__ E,
                                                                T *__restrict__ Cb, int tpb)
{
    __shared__ double sP[2][64 * 64];
    const PruneJob &jb = jobs[sel ? sel[blockIdx.y] : blockIdx.y];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    const int nch = jb.nch;
    {
        int k = 0;
        for (int c = 0; c < nch && k < 2; c++) {
            if (jb.ch_tip[c] >= 0) continue;
            const double *P = PT + (size_t)jb.ch_pmat[c] * 4096;
            for (int q = tid; q < 4096; q += 64 * kResWaves)
                sP[k][q] = P[q] + ((q >> 6) == (q & 63) ? 1.0 : 0.0);
            k++;
        }
    }
    __syncthreads();
    T *Lo = L + (size_t)jb.out_slot * npad * 64;
    for (int t = 0; t < tpb; t++) {
        const int hw = (blockIdx.x * tpb + t) * (16 * kResWaves) + 16 * w;
        if (hw >= npad) break;   // the same for the whole wave
        const int hb = hw + g;
        d4 acc[4];
        int esum[4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            acc[b] = d4{1, 1, 1, 1};
            esum[b] = 0;
        }
        int k = 0;
        for (int c = 0; c < nch; c++) {
            const int pm = jb.ch_pmat[c];
            const int tip = jb.ch_tip[c];
            if (tip >= 0) {
                const unsigned char *zt = z + (size_t)tip * npad;
                int code[4];
#pragma unroll
                for (int i = 0; i < 4; i++) code[i] = zt[hb + 4 * i];
#pragma unroll
                for (int b = 0; b < 4; b++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        acc[b][i] *= tip_value(TT, pm, ncodes, code[i], 16 * b + r);
            } else {
                if (k >= 2 && (k & 1) == 0) {
                    // a third or later interior child (the unrooted root,
                    // multifurcations; tpb is 1 for such jobs): its P and the
                    // next one's are staged over the first pair
                    __syncthreads();
                    int k2 = 0;
                    for (int c2 = c; c2 < nch && k2 < 2; c2++) {
                        if (jb.ch_tip[c2] >= 0) continue;
                        const double *P = PT + (size_t)jb.ch_pmat[c2] * 4096;
                        for (int e = tid; e < 4096; e += 64 * kResWaves)
                            sP[k2][e] = P[e] + ((e >> 6) == (e & 63) ? 1.0 : 0.0);
                        k2++;
                    }
                    __syncthreads();
                }
                const int slot = jb.ch_slot[c];
                const T *Lc = L + (size_t)slot * npad * 64;
                const int *Ec = E + (size_t)slot * npad;
                double a[16];
                Store<T>::load16(Lc + (size_t)(hw + r) * 64 + 16 * g, a);
                d4 tmp[4];
#pragma unroll
                for (int b = 0; b < 4; b++) tmp[b] = d4{0, 0, 0, 0};
                const double *Pk = sP[k & 1];
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const double *prow = Pk + (16 * g + kk) * 64 + r;
#pragma unroll
                    for (int b = 0; b < 4; b++) tmp[b] = mfma64(a[kk], prow[16 * b], tmp[b]);
                }
                if (Cb) {
                    const int cs = jb.ch_cslot[c];
                    T *Co = Cb + (size_t)cs * npad * 64;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        uint32_t rb = row_seed((uint32_t)cs ^ 0x3C6EF372u, (uint32_t)(hb + 4 * i),
                                               (uint32_t)r);
#pragma unroll
                        for (int b = 0; b < 4; b++) {
                            Store<T>::put(Co + (size_t)(hb + 4 * i) * 64 + 16 * b + r,
                                          tmp[b][i], rb);
                            rb = xorshift32(rb);
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < 4; b++) acc[b] *= tmp[b];
#pragma unroll
                for (int i = 0; i < 4; i++) esum[i] += Ec[hb + 4 * i];
                k++;
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double m = acc[0][i];
#pragma unroll
            for (int b = 1; b < 4; b++) m = acc[b][i] > m ? acc[b][i] : m;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                double o = __shfl_xor(m, off, 64);
                m = o > m ? o : m;
            }
            int e = 0;
            if (m > 0) frexp(m, &e);
            const int h = hb + 4 * i;
            uint32_t rb = row_seed((uint32_t)jb.out_slot, (uint32_t)h, (uint32_t)r);
#pragma unroll
            for (int b = 0; b < 4; b++) {
                Store<T>::put(Lo + (size_t)h * 64 + 16 * b + r, ldexp(acc[b][i], -e), rb);
                rb = xorshift32(rb);
            }
            if (r == 0) E[(size_t)jb.out_slot * npad + h] = esum[i] + e;
        }
    }
}

// Persistent pruning of a forest of genes: one launch for every level.  An
// item is (gene, 128-pattern tile); block b takes items b, b + gridDim.x,
// ... of `items` (ordered by the host, genes with the most jobs first) and
// walks the gene's jobs in level order (wjobs[woff[g] .. woff[g + 1]),
// indices into jobs) for its tile.  Patterns are independent, so a tile's
// jobs depend only on earlier jobs of the same tile, which the same block
// ran: no flag or wait between blocks.  Interior children's P are staged
// two at a time (the first two of every job as in k_prune_res; a third or
// more after them), each pair after a __syncthreads() (a workgroup-scope
// fence with the barrier): every job that reads another job's partials has
// an interior child, so that barrier orders the loads after the child job's
// stores, and the staging after the last reads of sP.  Per pattern
// the arithmetic and the order of the children are those of k_prune_res
// and k_prune_mfma, so the results are the same bitwise.
template <typename T>
__global__ __launch_bounds__(64 * kResWaves, 4) void k_prune_walk(const PruneJob *__restrict__ jobs,
                                                                 const int *__restrict__ wjobs,
                                                                 const int *__restrict__ woff,
                                                                 const int *__restrict__ items,
                                                                 int nitems, int ntile, int npad,
                                                                 const unsigned char *__restrict__ z,
                                                                 const double *__restrict__ PT,
                                                                 const double *__restrict__ TT,
                                                                 int ncodes, T *__restrict__ L,
                                                                 int *__restrict__ E,
                                                                 T *__restrict__ Cb)
{
    __shared__ double sP[2][64 * 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
        const int item = items[it];
        const int gene = item / ntile, tile = item - gene * ntile;
        const int hw = tile * (16 * kResWaves) + 16 * w;
        const int hb = hw + g;
        const bool live = hw < npad;   // the same for the whole wave
        for (int q = woff[gene]; q < woff[gene + 1]; q++) {
            const PruneJob &jb = jobs[wjobs[q]];
            const int nch = jb.nch;
            d4 acc[4];
            int esum[4];
#pragma unroll
            for (int b = 0; b < 4; b++) {
                acc[b] = d4{1, 1, 1, 1};
                esum[b] = 0;
            }
            int k = 0;   // interior children seen
            for (int c = 0; c < nch; c++) {
                const int pm = jb.ch_pmat[c];
                const int tip = jb.ch_tip[c];
                if (tip >= 0) {
                    if (!live) continue;
                    const unsigned char *zt = z + (size_t)tip * npad;
                    int code[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) code[i] = zt[hb + 4 * i];
#pragma unroll
                    for (int b = 0; b < 4; b++)
#pragma unroll
                        for (int i = 0; i < 4; i++)
                            acc[b][i] *= tip_value(TT, pm, ncodes, code[i], 16 * b + r);
                    continue;
                }
                if ((k & 1) == 0) {
                    // stage this interior child's P and the next one's: the
                    // previous job's (or pair's) reads of sP and the previous
                    // job's stores come first
                    __syncthreads();
                    int kk2 = 0;
                    for (int c2 = c; c2 < nch && kk2 < 2; c2++) {
                        if (jb.ch_tip[c2] >= 0) continue;
                        const double *P = PT + (size_t)jb.ch_pmat[c2] * 4096;
                        for (int e = tid; e < 4096; e += 64 * kResWaves)
                            sP[kk2][e] = P[e] + ((e >> 6) == (e & 63) ? 1.0 : 0.0);
                        kk2++;
                    }
                    __syncthreads();
                }
                const double *Pk = sP[k & 1];
                k++;
                if (!live) continue;
                const int slot = jb.ch_slot[c];
                const T *Lc = L + (size_t)slot * npad * 64;
                const int *Ec = E + (size_t)slot * npad;
                double a[16];
                Store<T>::load16(Lc + (size_t)(hw + r) * 64 + 16 * g, a);
                d4 tmp[4];
#pragma unroll
                for (int b = 0; b < 4; b++) tmp[b] = d4{0, 0, 0, 0};
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const double *prow = Pk + (16 * g + kk) * 64 + r;
#pragma unroll
                    for (int b = 0; b < 4; b++) tmp[b] = mfma64(a[kk], prow[16 * b], tmp[b]);
                }
                if (Cb) {
                    const int cs = jb.ch_cslot[c];
                    T *Co = Cb + (size_t)cs * npad * 64;
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        uint32_t rb = row_seed((uint32_t)cs ^ 0x3C6EF372u, (uint32_t)(hb + 4 * i),
                                               (uint32_t)r);
#pragma unroll
                        for (int b = 0; b < 4; b++) {
                            Store<T>::put(Co + (size_t)(hb + 4 * i) * 64 + 16 * b + r,
                                          tmp[b][i], rb);
                            rb = xorshift32(rb);
                        }
                    }
                }
#pragma unroll
                for (int b = 0; b < 4; b++) acc[b] *= tmp[b];
#pragma unroll
                for (int i = 0; i < 4; i++) esum[i] += Ec[hb + 4 * i];
            }
            if (!live) continue;
            T *Lo = L + (size_t)jb.out_slot * npad * 64;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                double m = acc[0][i];
#pragma unroll
                for (int b = 1; b < 4; b++) m = acc[b][i] > m ? acc[b][i] : m;
#pragma unroll
                for (int off = 1; off < 16; off <<= 1) {
                    double o = __shfl_xor(m, off, 64);
                    m = o > m ? o : m;
                }
                int e = 0;
                if (m > 0) frexp(m, &e);
                const int h = hb + 4 * i;
                uint32_t rb = row_seed((uint32_t)jb.out_slot, (uint32_t)h, (uint32_t)r);
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    Store<T>::put(Lo + (size_t)h * 64 + 16 * b + r, ldexp(acc[b][i], -e), rb);
                    rb = xorshift32(rb);
                }
                if (r == 0) E[(size_t)jb.out_slot * npad + h] = esum[i] + e;
            }
        }
    }
}

// ------------------------------------------------------------------ pruning (FP32 MFMA)

__device__ inline f4 mfma32(float a, float b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int kT32 = 65;   // floats per row of a wave's transpose tile in LDS

// Top two (value, state) of a pattern row whose 64 states are spread over
// the 4 lanes 16g + r (g = 0..3) of a wave, 16 states per lane (a[kk] is
// state 16g + kk): order by value, then by the lower state; afterwards every
// one of the 4 lanes holds the row's pair.
struct Top2 {
    float v1, v2;
    int c1, c2;
};
__device__ inline bool top_before(float va, int ca, float vb, int cb)
{
    return va > vb || (va == vb && ca < cb);
}
__device__ inline Top2 row_top2(const float *a, int g)
{
    // branch-free (selects), so that the 16 updates and the 2 merges stay
    // straight-line code
    Top2 t{-1.f, -1.f, 64, 64};
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
        const float v = a[kk];
        const int c = 16 * g + kk;
        const bool b1 = top_before(v, c, t.v1, t.c1), b2 = top_before(v, c, t.v2, t.c2);
        t.v2 = b1 ? t.v1 : (b2 ? v : t.v2);
        t.c2 = b1 ? t.c1 : (b2 ? c : t.c2);
        t.v1 = b1 ? v : t.v1;
        t.c1 = b1 ? c : t.c1;
    }
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
        const float pv1 = __shfl_xor(t.v1, off, 64), pv2 = __shfl_xor(t.v2, off, 64);
        const int pc1 = __shfl_xor(t.c1, off, 64), pc2 = __shfl_xor(t.c2, off, 64);
        // merged pair: the first of the two firsts, then the better of the
        // other first and the winner's second
        const bool pw = top_before(pv1, pc1, t.v1, t.c1);
        const float ov = pw ? t.v1 : pv1, sv = pw ? pv2 : t.v2;
        const int oc = pw ? t.c1 : pc1, sc = pw ? pc2 : t.c2;
        const bool ob = top_before(ov, oc, sv, sc);
        t.v1 = pw ? pv1 : t.v1;
        t.c1 = pw ? pc1 : t.c1;
        t.v2 = ob ? ov : sv;
        t.c2 = ob ? oc : sc;
    }
    return t;
}

// One tree level of pruning with the products on the FP32 matrix cores
// (BSA_KERNEL_FP32MC, float storage), for jobs with at most two interior
// children; blocks as k_prune_res (8 waves, 128 patterns per tile, both
// children's fl32(P - I) resident in LDS).  Per interior child and pattern h
//   C(h, j) = L(h, j) + sum_{k in top2(h)} L(h, k) (P - I)_{j,k}
//                     + sum_{k not in top2(h)} L(h, k) fl32(P - I)_{j,k},
// the first two terms in FP64, the last on the matrix cores (16x16x4, 16
// k-steps x 4 column blocks).  The matrix core adds each product to its FP32
// accumulator in turn with round-to-nearest, so a term below half an ulp of
// the running sum is lost outright: with every off-diagonal term >= 0, the
// plain contraction loses the smaller terms systematically (lnL 1.5e-3 too
// low at 128 taxa x 10,000 codons).  The two largest L(h, k) of a row carry
// almost all of a row's magnitude; without them the accumulator holds only
// terms of comparable size.  L(h, j) comes from a transpose tile of the
// wave's rows in LDS; accumulator element i of lane (g, r) is pattern
// 4g + i of the wave's 16, parent state 16b + r.  Children, scaling and
// stores as in k_prune_res.  At 156 VGPRs (three waves per SIMD, one block
// per CU) it spills nothing; held to 128 for two blocks per CU it spilled
// 24 VGPRs and ran 6 % slower.
__global__ __launch_bounds__(64 * kResWaves, 2) void k_prune32(const PruneJob *__restrict__ jobs,
                                                              const int *__restrict__ sel, int npad,
                                                              const unsigned char *__restrict__ z,
                                                              const double *__restrict__ PT,
                                                              const double *__restrict__ TT,
                                                              int ncodes, float *__restrict__ L,
                                                              int *__restrict__ E,
                                                              float *__restrict__ Cb, int tpb)
{
    __shared__ float sP[2][64 * kT32];          // fl32(P - I)[k][j] per interior child
    __shared__ float sX[kResWaves][16 * kT32];  // per wave: its 16 child rows
    __shared__ Top2 sT[kResWaves][16];           // per wave: each row's top two
    const PruneJob &jb = jobs[sel ? sel[blockIdx.y] : blockIdx.y];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    const int nch = jb.nch;
    {
        int k = 0;
        for (int c = 0; c < nch && k < 2; c++) {
            if (jb.ch_tip[c] >= 0) continue;
            const double *P = PT + (size_t)jb.ch_pmat[c] * 4096;
            for (int q = tid; q < 4096; q += 64 * kResWaves)
                sP[k][(q >> 6) * kT32 + (q & 63)] = (float)P[q];
            k++;
        }
    }
    __syncthreads();
    float *xw = sX[w];
    float *Lo = L + (size_t)jb.out_slot * npad * 64;
    for (int t = 0; t < tpb; t++) {
        const int hw = (blockIdx.x * tpb + t) * (16 * kResWaves) + 16 * w;
        if (hw >= npad) break;   // the same for the whole wave
        d4 acc[4];
        int esum[4];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            acc[b] = d4{1, 1, 1, 1};
            esum[b] = 0;
        }
        int k = 0;
        for (int c = 0; c < nch; c++) {
            const int pm = jb.ch_pmat[c];
            const int tip = jb.ch_tip[c];
            if (tip >= 0) {
                const unsigned char *zt = z + (size_t)tip * npad + hw + 4 * g;
                int code[4];
#pragma unroll
                for (int i = 0; i < 4; i++) code[i] = zt[i];
#pragma unroll
                for (int b = 0; b < 4; b++)
#pragma unroll
                    for (int i = 0; i < 4; i++)
                        acc[b][i] *= tip_value(TT, pm, ncodes, code[i], 16 * b + r);
                continue;
            }
            const int slot = jb.ch_slot[c];
            const float *Lc = L + (size_t)slot * npad * 64;
            const int *Ec = E + (size_t)slot * npad;
            float a[16];
            {
                const f4 *q = reinterpret_cast<const f4 *>(Lc + (size_t)(hw + r) * 64 + 16 * g);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const f4 v = q[i];
                    a[4 * i] = v.x;
                    a[4 * i + 1] = v.y;
                    a[4 * i + 2] = v.z;
                    a[4 * i + 3] = v.w;
                }
            }
            // the previous child's reads of xw (the wave's own rows) come first
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int kk = 0; kk < 16; kk++) xw[r * kT32 + 16 * g + kk] = a[kk];
            const Top2 tp = row_top2(a, g);
            if (g == 0) sT[w][r] = tp;
#pragma unroll
            for (int kk = 0; kk < 16; kk++)
                if (16 * g + kk == tp.c1 || 16 * g + kk == tp.c2) a[kk] = 0.f;
            __builtin_amdgcn_wave_barrier();
            const float *Pk = sP[k];
            f4 Y[4];
#pragma unroll
            for (int b = 0; b < 4; b++) Y[b] = f4{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                const float *prow = Pk + (16 * g + kk) * kT32 + r;
#pragma unroll
                for (int b = 0; b < 4; b++) Y[b] = mfma32(a[kk], prow[16 * b], Y[b]);
            }
            float *Co = Cb ? Cb + (size_t)jb.ch_cslot[c] * npad * 64 : nullptr;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                // row 4g + i of the wave (its top two from LDS, one address
                // per lane group: a broadcast)
                const int src = 4 * g + i;
                const Top2 tq = sT[w][src];
                const float v1 = tq.v1, v2 = tq.v2;
                const int c1 = tq.c1 & 63, c2 = tq.c2 & 63;
                const int h = hw + src;
                uint32_t rb = Co ? row_seed((uint32_t)jb.ch_cslot[c] ^ 0x3C6EF372u, (uint32_t)h,
                                            (uint32_t)r)
                                 : 0u;
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int j = 16 * b + r;
                    const double cv = (double)xw[src * kT32 + j] +
                                      (double)v1 * (double)Pk[c1 * kT32 + j] +
                                      (double)v2 * (double)Pk[c2 * kT32 + j] + (double)Y[b][i];
                    if (Co) {
                        Store<float>::put(Co + (size_t)h * 64 + j, cv, rb);
                        rb = xorshift32(rb);
                    }
                    acc[b][i] *= cv;
                }
                esum[i] += Ec[h];
            }
            k++;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            double m = acc[0][i];
#pragma unroll
            for (int b = 1; b < 4; b++) m = acc[b][i] > m ? acc[b][i] : m;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                double o = __shfl_xor(m, off, 64);
                m = o > m ? o : m;
            }
            int e = 0;
            if (m > 0) frexp(m, &e);
            const int h = hw + 4 * g + i;
            uint32_t rb = row_seed((uint32_t)jb.out_slot, (uint32_t)h, (uint32_t)r);
#pragma unroll
            for (int b = 0; b < 4; b++) {
                Store<float>::put(Lo + (size_t)h * 64 + 16 * b + r, ldexp(acc[b][i], -e), rb);
                rb = xorshift32(rb);
            }
            if (r == 0) E[(size_t)jb.out_slot * npad + h] = esum[i] + e;
        }
    }
}

// ------------------------------------------------------------------ pruning (VALU)

// Same job on vector FMA, tile of 64 patterns: thread = parent state j (lane)
// x 16 patterns (wave); identity part added after the product as above.
template <typename T>
__global__ __launch_bounds__(256) void k_prune_valu(const PruneJob *__restrict__ jobs,
                                                    const int *__restrict__ sel, int npad,
                                                    const unsigned char *__restrict__ z,
                                                    const double *__restrict__ PT,
                                                    const double *__restrict__ TT, int ncodes,
                                                    T *__restrict__ L, int *__restrict__ E,
                                                    T *__restrict__ Cb)
{
    __shared__ double sP[64 * 64];
    __shared__ double sA[64 * 64];
    const PruneJob &jb = jobs[sel ? sel[blockIdx.y] : blockIdx.y];
    const int h0 = blockIdx.x * 64;
    const int tid = threadIdx.x, j = tid & 63, pg = tid >> 6;
    const int hb = h0 + 16 * pg;
    double acc[16];
    int esum[16];
#pragma unroll
    for (int i = 0; i < 16; i++) { acc[i] = 1; esum[i] = 0; }
    const int nch = jb.nch;
    for (int c = 0; c < nch; c++) {
        const int pm = jb.ch_pmat[c];
        const double *P = PT + (size_t)pm * 4096;
        const int tip = jb.ch_tip[c];
        if (tip >= 0) {
            const unsigned char *zt = z + (size_t)tip * npad + hb;
#pragma unroll
            for (int i = 0; i < 16; i++) acc[i] *= tip_value(TT, pm, ncodes, (int)zt[i], j);
        } else {
            const int slot = jb.ch_slot[c];
            const T *A = L + ((size_t)slot * npad + h0) * 64;
            __syncthreads();
            for (int q = tid; q < 4096; q += 256) { sP[q] = P[q]; sA[q] = Store<T>::get(A + q); }
            __syncthreads();
            double tmp[16];
#pragma unroll
            for (int i = 0; i < 16; i++) tmp[i] = 0;
            for (int k = 0; k < 64; k++) {
                double p = sP[k * 64 + j];
#pragma unroll
                for (int i = 0; i < 16; i++) tmp[i] = fma(sA[(16 * pg + i) * 64 + k], p, tmp[i]);
            }
            const int *Ec = E + (size_t)slot * npad + hb;
            T *Co = Cb ? Cb + (size_t)jb.ch_cslot[c] * npad * 64 : nullptr;
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const double cv = tmp[i] + sA[(16 * pg + i) * 64 + j];
                if (Co)
                    Store<T>::put(Co + (size_t)(hb + i) * 64 + j, cv,
                                  mix32((uint32_t)jb.ch_cslot[c] ^ 0x3C6EF372u, (uint32_t)(hb + i),
                                        (uint32_t)j));
                acc[i] *= cv;
                esum[i] += Ec[i];
            }
        }
    }
    T *Lo = L + (size_t)jb.out_slot * npad * 64;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        double m = acc[i];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            double o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        int e = 0;
        if (m > 0) frexp(m, &e);
        const int h = hb + i;
        Store<T>::put(Lo + (size_t)h * 64 + j, ldexp(acc[i], -e),
                      mix32((uint32_t)jb.out_slot, (uint32_t)h, (uint32_t)j));
        if (j == 0) E[(size_t)jb.out_slot * npad + h] = esum[i] + e;
    }
}

// ------------------------------------------------------------------ root

// One wave per pattern at a time, lane = state.  Per class: f = sum_i pi_i L_i
// (<= 0 -> 1e-300, treesub.c:7522-7533; a floored class gets log(1e-300)
// itself: the reference would add its own node-scaling logs here, which
// depend on where it scales), l_c = log f + E ln 2; site value
// m + log sum_c p_c exp(l_c - m) (treesub.c:7364-7393); the fpatt-weighted
// site sum is compensated (Neumaier) per wave, per block and across blocks.
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline void neumaier(double &s, double &c, double x)
{
    double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

template <typename T>
__global__ __launch_bounds__(256) void k_root(const RootArgs *__restrict__ args,
                                              const int *__restrict__ gsel,
                                              const T *__restrict__ L,
                                              const int *__restrict__ E,
                                              const double *__restrict__ pi,
                                              const double *__restrict__ uvr,
                                              const double *__restrict__ fpatt,
                                              double *__restrict__ fh, double *__restrict__ site,
                                              double *__restrict__ partial, T *__restrict__ A,
                                              int *__restrict__ EA)
{
    __shared__ double ss[4][kRootSums], sc[4][kRootSums];
    const int gi = gsel ? gsel[blockIdx.y] : blockIdx.y;
    const RootArgs a = args[gi];
    pi += (size_t)gi * 64;
    const double *fk = uvr + (size_t)gi * kGeneBlock + 3 * kUVRSet;
    fpatt += (size_t)gi * a.npad;
    fh += (size_t)gi * 4 * a.npad;
    site += (size_t)gi * a.npad;
    partial += (size_t)gi * gridDim.x * 2 * kRootSums;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double pil = lane < 61 ? pi[lane] : 0.0;
    const bool seed = a.seed_aslot[0] >= 0;
    double s[kRootSums], comp[kRootSums];
#pragma unroll
    for (int q = 0; q < kRootSums; q++) s[q] = comp[q] = 0;
    for (int q = 0; q < kRootPerWave; q++) {
        const int h = blockIdx.x * (4 * kRootPerWave) + w * kRootPerWave + q;
        if (h >= a.npad) break;
        if (h >= a.npatt) {   // padding: zero adjoint seeds
            if (seed)
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const size_t row = (size_t)a.seed_aslot[c] * a.npad + h;
                    Store<T>::put(A + row * 64 + lane, 0.0, 0);
                    if (lane == 0) EA[row] = 0;
                }
            continue;
        }
        double l[4], v[4];
        int e[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const size_t row = (size_t)a.root_slot[c] * a.npad + h;
            v[c] = wave_sum(pil * Store<T>::get(L + row * 64 + lane));
            e[c] = E[row];
            // a class that cannot produce the pattern gets f = 1e-300
            l[c] = v[c] > 0 ? log(v[c]) + (double)e[c] * kLn2 : kLogFloor;
        }
        double m = l[0];
#pragma unroll
        for (int c = 1; c < 4; c++) m = l[c] > m ? l[c] : m;
        double f = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) f += fk[c] * exp(l[c] - m);
        const double st = m + log(f);
        const double wgt = fpatt[h] > 0 ? fpatt[h] : 0.0;
        if (seed) {
            // d lnL / d L_root^c(h, j) = w pi_j p_c / f_h; stored scaled by
            // 2^E_root: w pi_j post_c / ftilde_c, exponent -E_root
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const size_t row = (size_t)a.seed_aslot[c] * a.npad + h;
                const double post = fk[c] * exp(l[c] - st);
                const double val = v[c] > 0 ? wgt * pil * post / v[c] : 0.0;
                Store<T>::put(A + row * 64 + lane, val, (uint32_t)(h * 64 + lane) ^ 0x5bd1e995u);
                if (lane == 0) EA[row] = -e[c];
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 4; c++) fh[(size_t)c * a.npad + h] = l[c];
            site[h] = st;
            if (wgt > 0) {
                neumaier(s[0], comp[0], wgt * st);
#pragma unroll
                for (int c = 0; c < 4; c++) neumaier(s[1 + c], comp[1 + c], wgt * exp(l[c] - st));
            }
        }
    }
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < kRootSums; q++) { ss[w][q] = s[q]; sc[w][q] = comp[q]; }
    __syncthreads();
    if (threadIdx.x < kRootSums) {
        const int q = threadIdx.x;
        double S = 0, Cc = 0;
        for (int i = 0; i < 4; i++) { neumaier(S, Cc, ss[i][q]); Cc += sc[i][q]; }
        partial[(size_t)blockIdx.x * 2 * kRootSums + 2 * q] = S;
        partial[(size_t)blockIdx.x * 2 * kRootSums + 2 * q + 1] = Cc;
    }
}

// One wave per reduced quantity: lane l sums blocks l, l+64, ...
// (compensated), then a fixed butterfly combines the 64 (sum, compensation)
// pairs; the order never depends on timing, so results are bitwise repeatable.
__global__ __launch_bounds__(64) void k_final(const int *__restrict__ gsel,
                                              const double *__restrict__ partial, int nblk,
                                              double *__restrict__ out)
{
    const int lane = threadIdx.x, q = blockIdx.x;
    const int gi = gsel ? gsel[blockIdx.y] : blockIdx.y;
    partial += (size_t)gi * nblk * 2 * kRootSums;
    out += (size_t)gi * kRootSums;
    double S = 0, Cc = 0;
    for (int i = lane; i < nblk; i += 64) {
        neumaier(S, Cc, partial[(size_t)i * 2 * kRootSums + 2 * q]);
        Cc += partial[(size_t)i * 2 * kRootSums + 2 * q + 1];
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double So = __shfl_xor(S, off, 64), Co = __shfl_xor(Cc, off, 64);
        // combine in lane-index order so both partners compute the same value
        double lo = (lane & off) ? So : S, hi = (lane & off) ? S : So;
        double lc = (lane & off) ? Co : Cc, hc = (lane & off) ? Cc : Co;
        double s2 = lo, c2 = lc;
        neumaier(s2, c2, hi);
        S = s2;
        Cc = c2 + hc;
    }
    if (lane == 0) out[q] = S + Cc;
}

// ------------------------------------------------------------------ adjoint

constexpr int kLdsStride = 65;   // doubles per LDS row: row and column reads both conflict-free

__device__ inline double pow2i(int d)
{
    d = d < -1000 ? -1000 : (d > 1000 ? 1000 : d);
    return ldexp(1.0, d);
}

// Value of X_t(h, j) = Abar_parent,t(h, j) * prod_{siblings} C_s,t(h, j) for N
// (16 or 8) consecutive states j0..j0+N-1 of pattern h (one lane), and its
// exponent.
template <typename T, int N = 16>
__device__ inline int x_row16(const AdjJob &jb, int term, int h, int j0, int npad,
                              const unsigned char *z, const double *TT, int ncodes, const int *E,
                              const T *Cb, const T *A, const int *EA, double *x)
{
    const int asl = jb.a_aslot[term];
    Store<T>::template load16<N>(A + ((size_t)asl * npad + h) * 64 + j0, x);
    int ex = EA[(size_t)asl * npad + h];
    for (int q = 0; q < jb.nsib; q++) {
        const int tip = jb.sib_tip[q];
        if (tip >= 0) {
            const int pm = jb.sib_pmat[term][q];
            const int code = z[(size_t)tip * npad + h];
#pragma unroll
            for (int kk = 0; kk < N; kk++) x[kk] *= tip_value(TT, pm, ncodes, code, j0 + kk);
        } else {
            double c[N];
            Store<T>::template load16<N>(Cb + ((size_t)jb.sib_cslot[term][q] * npad + h) * 64 + j0,
                                         c);
#pragma unroll
            for (int kk = 0; kk < N; kk++) x[kk] *= c[kk];
            ex += E[(size_t)jb.sib_slot[term][q] * npad + h];
        }
    }
    return ex;
}

// X(h, j0..j0+N-1) = sum over the job's terms of X_t, exponents aligned to
// the largest; returns the exponent (the same for every j0 of a pattern).
template <typename T, int N = 16>
__device__ inline int x_sum16(const AdjJob &jb, int h, int j0, int npad, const unsigned char *z,
                              const double *TT, int ncodes, const int *E, const T *Cb,
                              const T *A, const int *EA, double *x)
{
    int ex = x_row16<T, N>(jb, 0, h, j0, npad, z, TT, ncodes, E, Cb, A, EA, x);
    for (int term = 1; term < jb.nterm; term++) {
        double t[N];
        const int et = x_row16<T, N>(jb, term, h, j0, npad, z, TT, ncodes, E, Cb, A, EA, t);
        const int em = ex > et ? ex : et;
        const double f0 = pow2i(ex - em), f1 = pow2i(et - em);
#pragma unroll
        for (int kk = 0; kk < N; kk++) x[kk] = x[kk] * f0 + t[kk] * f1;
        ex = em;
    }
    return ex;
}

// Adjoint of the interior children of one tree depth, per (job, run of tpb
// 64-pattern tiles): Abar_v(h, k) = sum_j X(h, j) P_{j,k}.  Shaped like the
// pruning kernel: wave w owns patterns 16w..16w+15 of each tile, each lane
// forms the 16 X values it feeds to the MFMA (pattern 16w + r, states
// 16g..16g+15) straight from HBM/L2, and P_v (identity included) is staged in
// LDS once per block.
template <typename T>
__global__ __launch_bounds__(256, 3) void k_adj_a(const AdjJob *__restrict__ jobs,
                                                  const int *__restrict__ sel, int npad,
                                                  const unsigned char *__restrict__ z,
                                                  const double *__restrict__ PT,
                                                  const double *__restrict__ TT, int ncodes,
                                                  const int *__restrict__ E,
                                                  const T *__restrict__ Cb, T *__restrict__ A,
                                                  int *__restrict__ EA, int tpb)
{
    __shared__ double sP[64 * kLdsStride];
    const AdjJob &jb = jobs[sel ? sel[blockIdx.y] : blockIdx.y];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    {
        const double *P = PT + (size_t)jb.v_pmat * 4096;
        for (int q = tid; q < 4096; q += 256) {
            const int k = q >> 6, j = q & 63;
            sP[k * kLdsStride + j] = P[q] + (k == j ? 1.0 : 0.0);
        }
    }
    const int ntile = npad / 64;
    for (int t = 0; t < tpb; t++) {
        const int tile = blockIdx.x * tpb + t;
        if (tile >= ntile) break;
        const int hw = tile * 64 + 16 * w;
        double x[16];
        const int ex = x_sum16<T>(jb, hw + r, 16 * g, npad, z, TT, ncodes, E, Cb, A, EA, x);
        if (t == 0) __syncthreads();
        d4 Y[4];
#pragma unroll
        for (int b = 0; b < 4; b++) Y[b] = d4{0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < 16; kk++)
#pragma unroll
            for (int b = 0; b < 4; b++)
                Y[b] = mfma64(x[kk], sP[(16 * b + r) * kLdsStride + 16 * g + kk], Y[b]);
        T *Av = A + (size_t)jb.v_aslot * npad * 64;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            // output row g + 4i of the wave; its X exponent sits in lane g + 4i
            const int exr = __shfl(ex, g + 4 * i, 64);
            double m = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) m = Y[b][i] > m ? Y[b][i] : m;
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                double o = __shfl_xor(m, off, 64);
                m = o > m ? o : m;
            }
            int e = 0;
            if (m > 0) frexp(m, &e);
            const int h = hw + g + 4 * i;
            uint32_t rb = row_seed((uint32_t)jb.v_aslot ^ 0xA5A5A5A5u, (uint32_t)h, (uint32_t)r);
#pragma unroll
            for (int b = 0; b < 4; b++) {
                Store<T>::put(Av + (size_t)h * 64 + 16 * b + r, ldexp(Y[b][i], -e), rb);
                rb = xorshift32(rb);
            }
            if (r == 0) EA[(size_t)jb.v_aslot * npad + h] = exr + e;
        }
    }
}

// Adjoints of one depth's interior children with FP32 matrix-core products
// (float storage, BSA_KERNEL_MFMA32):
//   Abar_v(h, k) = X(h, k) + sum_j X(h, j) fl32(P_v - I)_{j,k}
// X formed in FP64 as in k_adj_a, rounded to FP32 (to nearest) as the
// matrix-core operand and for the identity term, which is added in FP64 from
// a transpose tile of the wave's X rows in LDS; the off-diagonal terms are
// summed on their own scale by the FP32 accumulators (16x16x4, 16 k-steps x
// 4 column blocks).  The accumulator element i of lane (g, r) is pattern
// 4g + i of the wave's 16, state 16b + r.  An adjoint feeds only the
// gradient, where this rounding stays below 1e-6 relative; the likelihood
// itself keeps FP64 products.
__global__ __launch_bounds__(256, 4) void k_adj_a32(const AdjJob *__restrict__ jobs,
                                                    const int *__restrict__ sel, int npad,
                                                    const unsigned char *__restrict__ z,
                                                    const double *__restrict__ PT,
                                                    const double *__restrict__ TT, int ncodes,
                                                    const int *__restrict__ E,
                                                    const float *__restrict__ Cb,
                                                    float *__restrict__ A, int *__restrict__ EA,
                                                    int tpb)
{
    __shared__ float sP[64 * kT32];            // fl32(P_v - I), [k][j]
    __shared__ float sX[4][16 * kT32];         // per wave: X of its 16 patterns
    const AdjJob &jb = jobs[sel ? sel[blockIdx.y] : blockIdx.y];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    {
        const double *P = PT + (size_t)jb.v_pmat * 4096;
        for (int q = tid; q < 4096; q += 256) sP[(q >> 6) * kT32 + (q & 63)] = (float)P[q];
    }
    float *xw = sX[w];
    const int ntile = npad / 64;
    for (int t = 0; t < tpb; t++) {
        const int tile = blockIdx.x * tpb + t;
        if (tile >= ntile) break;
        const int hw = tile * 64 + 16 * w;
        double x[16];
        const int ex = x_sum16<float>(jb, hw + r, 16 * g, npad, z, TT, ncodes, E, Cb, A, EA, x);
        float xf[16];
        // the previous tile's reads of xw (the wave's own rows) come first
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            xf[kk] = (float)x[kk];
            xw[r * kT32 + 16 * g + kk] = xf[kk];
        }
        if (t == 0) __syncthreads();
        else __builtin_amdgcn_wave_barrier();
        f4 Y[4];
#pragma unroll
        for (int b = 0; b < 4; b++) Y[b] = f4{0, 0, 0, 0};
#pragma unroll
        for (int kk = 0; kk < 16; kk++)
#pragma unroll
            for (int b = 0; b < 4; b++)
                Y[b] = mfma32(xf[kk], sP[(16 * b + r) * kT32 + 16 * g + kk], Y[b]);
        float *Av = A + (size_t)jb.v_aslot * npad * 64;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            // output row 4g + i of the wave; its X exponent sits in lane 4g + i
            const int exr = __shfl(ex, 4 * g + i, 64);
            double yv[4];
            double m = 0;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                yv[b] = (double)xw[(4 * g + i) * kT32 + 16 * b + r] + (double)Y[b][i];
                m = yv[b] > m ? yv[b] : m;
            }
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) {
                double o = __shfl_xor(m, off, 64);
                m = o > m ? o : m;
            }
            int e = 0;
            if (m > 0) frexp(m, &e);
            const int h = hw + 4 * g + i;
            uint32_t rb = row_seed((uint32_t)jb.v_aslot ^ 0xA5A5A5A5u, (uint32_t)h, (uint32_t)r);
#pragma unroll
            for (int b = 0; b < 4; b++) {
                Store<float>::put(Av + (size_t)h * 64 + 16 * b + r, ldexp(yv[b], -e), rb);
                rb = xorshift32(rb);
            }
            if (r == 0) EA[(size_t)jb.v_aslot * npad + h] = exr + e;
        }
    }
}

// Branch matrices on the FP32 matrix cores (BSA_KERNEL_MFMA32): the product
//   M_v[j][k] += sum_h X(h, j) 2^{s(h)} L_v(h, k),   s(h) = ex(h) + E_v(h),
// over one 64-pattern tile, per 16-pattern group q (the rows one wave forms):
// the group's X rows are stored in LDS as float(X(h, j) 2^{s(h) - s_q}), s_q
// the largest s(h) of the group (so the values are <= 1 and keep their
// relative sizes), its four FP32 16x16x4 steps are summed in an FP32 tile,
// and the tile times 2^{s_q} is added to the FP64 accumulators.  sXm: the
// tile's scaled rows (row stride kT32), sGrp[q] = 2^{s_q}, bv[kk] =
// L_v(h0 + 4kk + g, 16w + r).  Wave w adds columns 16w..16w+15; Macc[a]
// element i is row 16a + 4g + i (the FP32 accumulator layout).
__device__ inline void m_tile32(const float *sXm, const double *sGrp, const float *bv, d4 Macc[4])
{
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        f4 Tq[4];
#pragma unroll
        for (int a = 0; a < 4; a++) Tq[a] = f4{0, 0, 0, 0};
#pragma unroll
        for (int st = 0; st < 4; st++) {
            const int kk = 4 * q + st;
#pragma unroll
            for (int a = 0; a < 4; a++)
                Tq[a] = mfma32(sXm[(4 * kk + g) * kT32 + 16 * a + r], bv[kk], Tq[a]);
        }
        const double sc = sGrp[q];
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
            for (int i = 0; i < 4; i++) Macc[a][i] += (double)Tq[a][i] * sc;
    }
}

// the largest s of the wave's 16 patterns (lane 16g + r holds pattern r's s)
__device__ inline int group_max16(int s)
{
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
        const int o = __shfl_xor(s, off, 64);
        s = o > s ? o : s;
    }
    return s;
}

// Branch matrices of every adjoint job, one launch after the sweep:
// M_v[j][k] += sum_h X(h, j) 2^{ex(h) + E_v(h)} I_v(h, k) over a chunk of tiles.
// Per tile the block forms the scaled X tile in LDS (each lane one pattern and
// 16 states, as in the child-adjoint kernel); wave w then owns columns
// 16w..16w+15 of M and loads its B operand (I_v) straight from HBM/L2.  The
// waves' column strips are disjoint, so each chunk writes its partial with no
// further reduction (chunk partials are summed later in a fixed order).
// For a tip v, I_v(h, .) is the indicator of the codons the tip's code allows
// (one column for a resolved codon).  Its patterns are visited in the order of
// tperm, sorted by that code, so a tile holds one or a few codons: there wave
// w owns rows 16w..16w+15 of M instead, and multiplies only the 16-column
// strips that the tile's codes reach (usually one of four).  tperm entries
// hold the pattern in the low 24 bits and its code in the top 8.
// M32 (float storage, BSA_KERNEL_MFMA32): an interior child's product runs on
// the FP32 matrix cores with m_tile32; tips keep the FP64 path.
template <typename T, bool M32>
__global__ __launch_bounds__(256, sizeof(T) == 4 && !M32 ? 4 : 3) void k_adj_m(const AdjJob *__restrict__ jobs,
                                                  const int *__restrict__ sel, int npad,
                                                  const unsigned char *__restrict__ z,
                                                  const double *__restrict__ TT, int ncodes,
                                                  const double *__restrict__ mask64,
                                                  const int *__restrict__ tperm,
                                                  const T *__restrict__ L,
                                                  const int *__restrict__ E,
                                                  const T *__restrict__ Cb,
                                                  const T *__restrict__ A,
                                                  const int *__restrict__ EA,
                                                  double *__restrict__ Mpart, int nchunk, int tpc)
{
    static_assert(!M32 || sizeof(T) == 4, "FP32 branch-matrix products need float storage");
    __shared__ double sX[64 * kLdsStride];
    __shared__ double sGrp[4];
    const AdjJob &jb = jobs[sel ? sel[blockIdx.y] : blockIdx.y];
    const int chunk = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    const int ntile = npad / 64;
    const bool vtip = jb.v_tip >= 0;
    const int col = 16 * w + r;
    const int *perm = vtip ? tperm + (size_t)jb.v_tip * npad : nullptr;
    // a tip's pattern for this lane's X row, loaded one tile ahead
    int pnext = 0;
    if (vtip && chunk * tpc < ntile) pnext = perm[chunk * tpc * 64 + 16 * w + r];
    d4 Macc[4];
#pragma unroll
    for (int a = 0; a < 4; a++) Macc[a] = d4{0, 0, 0, 0};
    for (int t = 0; t < tpc; t++) {
        const int tile = chunk * tpc + t;
        if (tile >= ntile) break;
        // row i of the tile is pattern h0 + i, or for a tip the pattern in
        // place h0 + i of the tip's order
        const int h0 = tile * 64;
        int hx = h0 + 16 * w + r;
        if (vtip) {
            hx = pnext & 0xFFFFFF;
            if (t + 1 < tpc && tile + 1 < ntile) pnext = perm[h0 + 64 + 16 * w + r];
        }
        // X in two halves of 8 states, so that at most 8 values of each
        // operand are live at a time (VGPRs for four blocks per CU)
        const int exv = vtip ? 0 : E[(size_t)jb.v_slot * npad + hx];
        if constexpr (M32) {
            if (!vtip) {
                float *sXm = reinterpret_cast<float *>(sX);
                float bv[16];
                const T *Lv = L + ((size_t)jb.v_slot * npad + h0 + g) * 64 + col;
#pragma unroll
                for (int kk = 0; kk < 16; kk++) bv[kk] = Lv[(size_t)(4 * kk) * 64];
                double sc;
                int sg;
                {
                    double x[8];
                    const int ex =
                        x_sum16<T, 8>(jb, hx, 16 * g, npad, z, TT, ncodes, E, Cb, A, EA, x);
                    const int sh = ex + exv;
                    sg = group_max16(sh);
                    sc = pow2i(sh - sg);
                    __syncthreads();   // the previous tile's reads of sX, sGrp are done
#pragma unroll
                    for (int kk = 0; kk < 8; kk++)
                        sXm[(16 * w + r) * kT32 + 16 * g + kk] = (float)(x[kk] * sc);
                }
                {
                    double x[8];
                    x_sum16<T, 8>(jb, hx, 16 * g + 8, npad, z, TT, ncodes, E, Cb, A, EA, x);
#pragma unroll
                    for (int kk = 0; kk < 8; kk++)
                        sXm[(16 * w + r) * kT32 + 16 * g + 8 + kk] = (float)(x[kk] * sc);
                }
                if (lane == 0) sGrp[w] = pow2i(sg);
                __syncthreads();
                m_tile32(sXm, sGrp, bv, Macc);
                continue;
            }
        }
        {
            double x[8];
            const int ex = x_sum16<T, 8>(jb, hx, 16 * g, npad, z, TT, ncodes, E, Cb, A, EA, x);
            const double sc = pow2i(ex + exv);
            __syncthreads();   // the previous tile's reads of sX are done
#pragma unroll
            for (int kk = 0; kk < 8; kk++) sX[(16 * w + r) * kLdsStride + 16 * g + kk] = x[kk] * sc;
        }
        {
            double x[8];
            const int ex =
                x_sum16<T, 8>(jb, hx, 16 * g + 8, npad, z, TT, ncodes, E, Cb, A, EA, x);
            const double sc = pow2i(ex + exv);
#pragma unroll
            for (int kk = 0; kk < 8; kk++)
                sX[(16 * w + r) * kLdsStride + 16 * g + 8 + kk] = x[kk] * sc;
        }
        if (vtip) {
            // the tip's codes in rows 4kk + g of the tile, and the column
            // strips they reach (every strip for an ambiguous code)
            int code[16];
            unsigned reach = 0;
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                code[kk] = (unsigned)perm[h0 + 4 * kk + g] >> 24;
                reach |= code[kk] < 61 ? 1u << (code[kk] >> 4) : 0xFu;
            }
            __syncthreads();
            // wave w: rows 16w..16w+15 of M, one product per strip the tile reaches
#pragma unroll
            for (int sp = 0; sp < 4; sp++) {
                if (!__any((reach >> sp) & 1u)) continue;   // the same for the whole wave
                const int c = 16 * sp + r;
#pragma unroll
                for (int kk = 0; kk < 16; kk++) {
                    const double b = code[kk] < 61
                                         ? (code[kk] == c ? 1.0 : 0.0)
                                         : mask64[(size_t)(jb.v_mask + code[kk]) * 64 + c];
                    Macc[sp] = mfma64(sX[(4 * kk + g) * kLdsStride + 16 * w + r], b, Macc[sp]);
                }
            }
        } else {
            double bv[16];
            const T *Lv = L + ((size_t)jb.v_slot * npad + h0 + g) * 64 + col;
#pragma unroll
            for (int kk = 0; kk < 16; kk++) bv[kk] = Store<T>::get(Lv + (size_t)(4 * kk) * 64);
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < 16; kk++)
#pragma unroll
                for (int a = 0; a < 4; a++)
                    Macc[a] = mfma64(sX[(4 * kk + g) * kLdsStride + 16 * a + r], bv[kk], Macc[a]);
        }
    }
    // interior child: Macc[a] holds rows 16a.., columns 16w..; tip: rows
    // 16w.., columns 16a..
    // (FP32 products of an interior child: rows 16a + 4g + i)
    double *Mo = Mpart + ((size_t)jb.m_index * nchunk + chunk) * 4096;
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int row = vtip ? 16 * w + g + 4 * i
                                 : 16 * a + (M32 ? 4 * g + i : g + 4 * i);
            const int cl = vtip ? 16 * a + r : col;
            Mo[row * 64 + cl] = Macc[a][i];
        }
}

// R = S1^T X S2^T for 64 x 64 matrices on the matrix cores (16x16x4 FP64
// steps, operands from LDS).  On entry sS holds S1 and sX holds X (row
// stride kLdsStride); S2 (row-major, 64 x 64) is read from global memory
// into sS once S1 is no longer needed, and S1^T X is kept in sX.  Wave w
// owns result rows 16w..16w+15: on return acc[b][i] = R[16w + g + 4i][16b + r]
// for lane 16g + r.  The contraction index runs over 16g + s at step s.
__device__ inline void two_products(double *sS, double *sX, const double *__restrict__ S2,
                                    d4 acc[4])
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
#pragma unroll
    for (int b = 0; b < 4; b++) acc[b] = d4{0, 0, 0, 0};
#pragma unroll
    for (int st = 0; st < 16; st++) {
        const int jj = 16 * g + st;
        const double ua = sS[jj * kLdsStride + 16 * w + r];
#pragma unroll
        for (int b = 0; b < 4; b++) acc[b] = mfma64(ua, sX[jj * kLdsStride + 16 * b + r], acc[b]);
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
        for (int i = 0; i < 4; i++) sX[(16 * w + g + 4 * i) * kLdsStride + 16 * b + r] = acc[b][i];
    for (int q = tid; q < 4096; q += 256) sS[(q >> 6) * kLdsStride + (q & 63)] = S2[q];
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 4; b++) acc[b] = d4{0, 0, 0, 0};
#pragma unroll
    for (int st = 0; st < 16; st++) {
        const int kk = 16 * g + st;
        const double ya = sX[(16 * w + r) * kLdsStride + kk];
#pragma unroll
        for (int b = 0; b < 4; b++) acc[b] = mfma64(ya, sS[(16 * b + r) * kLdsStride + kk], acc[b]);
    }
}

// Per (branch, class): M = sum of its chunk partials (fixed order), then
// A = U^T M V^T in the eigenbasis of its set;
// d lnL / d tau = sum_a A_aa lambda_a e^{lambda_a tau} and
// G = A o X(tau), X_ab = (e^{lambda_a tau} - e^{lambda_b tau}) / (lambda_a - lambda_b)
// (tau e^{lambda tau} on ties), so that for any rate-matrix parameter theta
// d lnL / d theta = <G, V dQ/dtheta U> (derivative of U e^{Lambda tau} V).
__global__ __launch_bounds__(256) void k_eigbranch(const EigJob *__restrict__ jobs,
                                                   const int *__restrict__ sel,
                                                   const double *__restrict__ Mpart, int nchunk,
                                                   const double *__restrict__ uvr,
                                                   double *__restrict__ G, double *__restrict__ dtau)
{
    __shared__ double sM[64 * kLdsStride];
    __shared__ double sU[64 * kLdsStride];
    __shared__ double sE[64];
    __shared__ double sred[4];
    const int ji = sel ? sel[blockIdx.x] : blockIdx.x;
    const EigJob jb = jobs[ji];
    const double *U = uvr + uvr_offset(jb.set);
    const double *V = U + 4096;
    const double *R = U + 8192;
    const int tid = threadIdx.x;
    const double *Mp = Mpart + (size_t)jb.m_index * nchunk * 4096;
    for (int q = tid; q < 4096; q += 256) {
        double m = 0;
        for (int c = 0; c < nchunk; c++) m += Mp[(size_t)c * 4096 + q];
        sM[(q >> 6) * kLdsStride + (q & 63)] = m;
        sU[(q >> 6) * kLdsStride + (q & 63)] = U[q];
    }
    if (tid < 64) sE[tid] = exp(R[tid] * jb.tau);
    __syncthreads();
    // A = U^T M V^T; lane 16g + r of wave w holds A[16w + g + 4i][16b + r]
    const int lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    d4 acc[4];
    two_products(sU, sM, V, acc);
    double dsum = 0;
    double *Go = G + (size_t)ji * 4096;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int c = 16 * b + r;
        const double lb = R[c], eb = sE[c];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int a = 16 * w + g + 4 * i;
            const double la = R[a], ea = sE[a];
            double x;
            const double dl = la - lb;
            if (fabs(dl * jb.tau) < 1e-6) {
                // e^{lb tau} (e^{dl tau} - 1)/dl = e^{lb tau} tau (1 + dl tau/2 + ...)
                const double u = dl * jb.tau;
                x = eb * jb.tau * (1.0 + u * (0.5 + u / 6.0));
            } else {
                x = (ea - eb) / dl;
            }
            Go[a * 64 + c] = acc[b][i] * x;
            if (a == c) dsum += acc[b][i] * la * ea;
        }
    }
    // reduce the diagonal terms
    for (int off = 32; off > 0; off >>= 1) dsum += __shfl_xor(dsum, off, 64);
    if ((tid & 63) == 0) sred[tid >> 6] = dsum;
    __syncthreads();
    if (tid == 0) dtau[ji] = sred[0] + sred[1] + sred[2] + sred[3];
}

// Per eigen set: partial sums of G over its (branch, class) jobs; part p
// takes list entries p, p + kWParts, ... in order (one thread per element).
__global__ __launch_bounds__(256) void k_wreduce(const double *__restrict__ G,
                                                 const int *__restrict__ set_off,
                                                 const int *__restrict__ set_list,
                                                 const int *__restrict__ ssel,
                                                 double *__restrict__ Wpart)
{
    const int set = ssel ? ssel[blockIdx.x] : blockIdx.x;
    const int q = blockIdx.y * 256 + threadIdx.x, p = blockIdx.z;
    double s = 0;
    for (int i = set_off[set] + p; i < set_off[set + 1]; i += kWParts)
        s += G[(size_t)set_list[i] * 4096 + q];
    Wpart[((size_t)p * gridDim.x + blockIdx.x) * 4096 + q] = s;
}

// Per eigen set: W = sum of its parts (fixed order), Z = V^T W U^T (so that
// <W, V dQ U> = <Z, dQ>).
__global__ __launch_bounds__(256) void k_zform(const double *__restrict__ Wpart,
                                               const int *__restrict__ ssel,
                                               const double *__restrict__ uvr,
                                               double *__restrict__ Z)
{
    __shared__ double sW[64 * kLdsStride];
    __shared__ double sB[64 * kLdsStride];
    const int set = ssel ? ssel[blockIdx.x] : blockIdx.x, pos = blockIdx.x, tid = threadIdx.x;
    const double *U = uvr + uvr_offset(set);
    const double *V = U + 4096;
    for (int q = tid; q < 4096; q += 256) {
        double w = 0;
        for (int p = 0; p < kWParts; p++) w += Wpart[((size_t)p * gridDim.x + pos) * 4096 + q];
        sW[(q >> 6) * kLdsStride + (q & 63)] = w;
        sB[(q >> 6) * kLdsStride + (q & 63)] = V[q];
    }
    __syncthreads();
    // Z = V^T W U^T; lane 16g + r of wave w holds Z[16w + g + 4i][16b + r]
    const int lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    d4 acc[4];
    two_products(sB, sW, U, acc);
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            Z[(size_t)set * 4096 + (16 * w + g + 4 * i) * 64 + 16 * b + r] = acc[b][i];
}

// ------------------------------------------------------------------ probe

__global__ void k_mfma_probe(const double *A, const double *B, double *D, const float *Af,
                             const float *Bf, float *Df)
{
    const int l = threadIdx.x;
    if (l >= 64) return;
    d4 c = {0, 0, 0, 0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(l & 15) * 4 + (l >> 4)], B[(l >> 4) * 16 + (l & 15)],
                                             c, 0, 0, 0);
    f4 cf = {0, 0, 0, 0};
    cf = __builtin_amdgcn_mfma_f32_16x16x4f32(Af[(l & 15) * 4 + (l >> 4)],
                                              Bf[(l >> 4) * 16 + (l & 15)], cf, 0, 0, 0);
    for (int i = 0; i < 4; i++) {
        D[l * 4 + i] = c[i];
        Df[l * 4 + i] = cf[i];
    }
}

// ------------------------------------------------------------------ launchers

void launch_pmat(hipStream_t s, const double *uvr, const PmatJob *jobs, const int *sel, int njobs,
                 double *PT, const unsigned char *mask_amb, int ncodes, double *TT)
{
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_pmat, dim3(njobs), dim3(256), 0, s, uvr, jobs, sel, PT);
    hipLaunchKernelGGL(k_tiptab, dim3(njobs), dim3(256), 0, s, jobs, sel, (const double *)PT,
                       mask_amb, ncodes, TT);
}

template <typename T>
void launch_prune(hipStream_t s, int kernel, int tile, const PruneJob *jobs, const int *sel,
                  int njobs, int maxint, int npad, const unsigned char *z, const double *PT,
                  const double *TT, int ncodes, T *L, int *E, T *Cb)
{
    if (njobs <= 0) return;
    // P resident in LDS where the launch has at least half as many blocks of
    // 8 waves as CUs (below that the 4-wave blocks spread the work better);
    // one tile per block, so that jobs with more than two interior children
    // can stage their later children's P over the first pair
    const int nres = (npad + 16 * kResWaves - 1) / (16 * kResWaves);
    if constexpr (sizeof(T) == 4) {
        // FP32 matrix-core products (jobs with more than two interior
        // children keep the FP64 kernel)
        if (kernel == 2 && maxint <= 2) {
            hipLaunchKernelGGL(k_prune32, dim3(nres, njobs), dim3(64 * kResWaves), 0, s, jobs, sel,
                               npad, z, PT, TT, ncodes, L, E, Cb, 1);
            return;
        }
    }
    if (kernel == 2) kernel = 0;
    if (kernel == 0 && tile == 64 && (long)nres * njobs >= 128) {
        hipLaunchKernelGGL(k_prune_res<T>, dim3(nres, njobs), dim3(64 * kResWaves), 0, s, jobs,
                           sel, npad, z, PT, TT, ncodes, L, E, Cb, 1);
        return;
    }
    if (kernel == 0 && tile == 128)
        hipLaunchKernelGGL((k_prune_mfma<T, 2>), dim3(npad / 128, njobs), dim3(256), 0, s, jobs,
                           sel, npad, z, PT, TT, ncodes, L, E, Cb);
    else if (kernel == 0)
        hipLaunchKernelGGL((k_prune_mfma<T, 1>), dim3(npad / 64, njobs), dim3(256), 0, s, jobs,
                           sel, npad, z, PT, TT, ncodes, L, E, Cb);
    else
        hipLaunchKernelGGL(k_prune_valu<T>, dim3(npad / 64, njobs), dim3(256), 0, s, jobs, sel,
                           npad, z, PT, TT, ncodes, L, E, Cb);
}

template <typename T>
void launch_prune_walk(hipStream_t s, const PruneJob *jobs, const int *wjobs, const int *woff,
                       const int *items, int nitems, int nblk, int npad,
                       const unsigned char *z, const double *PT, const double *TT, int ncodes,
                       T *L, int *E, T *Cb)
{
    if (nitems <= 0) return;
    const int ntile = (npad + 16 * kResWaves - 1) / (16 * kResWaves);
    hipLaunchKernelGGL(k_prune_walk<T>, dim3(std::max(1, std::min(nblk, nitems))),
                       dim3(64 * kResWaves), 0, s, jobs, wjobs, woff, items, nitems, ntile, npad,
                       z, PT, TT, ncodes, L, E, Cb);
}

void launch_adjoint_level32(hipStream_t s, const AdjJob *jobs, const int *sel, int njobs, int npad,
                            const unsigned char *z, const double *PT, const double *TT, int ncodes,
                            const int *E, const float *Cb, float *A, int *EA, int tpb)
{
    if (njobs <= 0) return;
    const int ntile = npad / 64;
    if (tpb <= 0) tpb = std::max(1, std::min(4, njobs * ntile / 512));
    hipLaunchKernelGGL(k_adj_a32, dim3((ntile + tpb - 1) / tpb, njobs), dim3(256), 0, s, jobs, sel,
                       npad, z, PT, TT, ncodes, E, Cb, A, EA, tpb);
}

template <typename T>
void launch_root(hipStream_t s, const RootArgs *args, const int *gsel, int ngene, const T *L,
                 const int *E, const double *pi, const double *uvr, const double *fpatt, double *fh,
                 double *site, double *partial, int nblk, double *out, T *A, int *EA)
{
    if (ngene <= 0) return;
    hipLaunchKernelGGL(k_root<T>, dim3(nblk, ngene), dim3(256), 0, s, args, gsel, L, E, pi, uvr,
                       fpatt, fh, site, partial, A, EA);
    hipLaunchKernelGGL(k_final, dim3(kRootSums, ngene), dim3(64), 0, s, gsel,
                       (const double *)partial, nblk, out);
}

template <typename T>
void launch_adjoint_level(hipStream_t s, const AdjJob *jobs, const int *sel, int njobs, int npad,
                          const unsigned char *z, const double *PT, const double *TT, int ncodes,
                          const int *E, const T *Cb, T *A, int *EA, int tpb)
{
    if (njobs <= 0) return;
    // several tiles per block (P_v staged once) when the level has blocks to
    // spare: up to 4, while the launch keeps at least 512 tiles' worth of
    // blocks per tile of a block
    const int ntile = npad / 64;
    if (tpb <= 0) tpb = std::max(1, std::min(4, njobs * ntile / 512));
    hipLaunchKernelGGL(k_adj_a<T>, dim3((ntile + tpb - 1) / tpb, njobs), dim3(256), 0, s, jobs,
                       sel, npad, z, PT, TT, ncodes, E, Cb, A, EA, tpb);
}

template <typename T>
void launch_adjoint_m(hipStream_t s, const AdjJob *jobs, const int *sel, int njobs, int npad,
                      int nchunk, int tpc, bool m32, const unsigned char *z, const double *PT,
                      const double *TT, int ncodes, const double *mask64, const int *tperm,
                      const T *L, const int *E, const T *Cb, const T *A, const int *EA,
                      double *Mpart)
{
    if (njobs <= 0) return;
    (void)PT;   // tip rows come from TT, interior rows from L
    if constexpr (sizeof(T) == 4) {
        if (m32) {
            hipLaunchKernelGGL((k_adj_m<T, true>), dim3(nchunk, njobs), dim3(256), 0, s, jobs, sel,
                               npad, z, TT, ncodes, mask64, tperm, L, E, Cb, A, EA, Mpart, nchunk,
                               tpc);
            return;
        }
    }
    hipLaunchKernelGGL((k_adj_m<T, false>), dim3(nchunk, njobs), dim3(256), 0, s, jobs, sel, npad,
                       z, TT, ncodes, mask64, tperm, L, E, Cb, A, EA, Mpart, nchunk, tpc);
}

void launch_adjoint_finish(hipStream_t s, const double *Mpart, int nchunk, const EigJob *ejobs,
                           const int *esel, int nej, const int *ssel, int nsets,
                           const int *set_off, const int *set_list, const double *uvr, double *G,
                           double *dtau, double *Wpart, double *Z)
{
    if (nej <= 0 || nsets <= 0) return;
    hipLaunchKernelGGL(k_eigbranch, dim3(nej), dim3(256), 0, s, ejobs, esel, Mpart, nchunk, uvr, G,
                       dtau);
    hipLaunchKernelGGL(k_wreduce, dim3(nsets, 16, kWParts), dim3(256), 0, s, (const double *)G,
                       set_off, set_list, ssel, Wpart);
    hipLaunchKernelGGL(k_zform, dim3(nsets), dim3(256), 0, s, (const double *)Wpart, ssel, uvr, Z);
}

template void launch_adjoint_level<double>(hipStream_t, const AdjJob *, const int *, int, int,
                                           const unsigned char *, const double *, const double *,
                                           int, const int *, const double *, double *, int *, int);
template void launch_adjoint_level<float>(hipStream_t, const AdjJob *, const int *, int, int,
                                          const unsigned char *, const double *, const double *,
                                          int, const int *, const float *, float *, int *, int);
template void launch_adjoint_m<double>(hipStream_t, const AdjJob *, const int *, int, int, int,
                                       int, bool, const unsigned char *, const double *,
                                       const double *,
                                       int, const double *, const int *, const double *,
                                       const int *, const double *, const double *, const int *,
                                       double *);
template void launch_adjoint_m<float>(hipStream_t, const AdjJob *, const int *, int, int, int, int,
                                      bool, const unsigned char *, const double *, const double *, int,
                                      const double *, const int *, const float *, const int *,
                                      const float *, const float *, const int *, double *);

void launch_mfma_probe(hipStream_t s, const double *A, const double *B, double *D64,
                       const float *Af, const float *Bf, float *D32)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D64, Af, Bf, D32);
}

template void launch_prune<double>(hipStream_t, int, int, const PruneJob *, const int *, int, int,
                                   int, const unsigned char *, const double *, const double *, int,
                                   double *, int *, double *);
template void launch_prune<float>(hipStream_t, int, int, const PruneJob *, const int *, int, int,
                                  int, const unsigned char *, const double *, const double *, int,
                                  float *, int *, float *);
template void launch_prune_walk<double>(hipStream_t, const PruneJob *, const int *, const int *,
                                        const int *, int, int, int, const unsigned char *,
                                        const double *, const double *, int, double *, int *,
                                        double *);
template void launch_prune_walk<float>(hipStream_t, const PruneJob *, const int *, const int *,
                                       const int *, int, int, int, const unsigned char *,
                                       const double *, const double *, int, float *, int *, float *);
template void launch_root<double>(hipStream_t, const RootArgs *, const int *, int, const double *,
                                  const int *, const double *, const double *, const double *,
                                  double *, double *, double *, int, double *, double *, int *);
template void launch_root<float>(hipStream_t, const RootArgs *, const int *, int, const float *,
                                 const int *, const double *, const double *, const double *,
                                 double *, double *, double *, int, double *, float *, int *);

// ------------------------------------------------------------------ Bayes empirical Bayes

// Per pattern: m_h = max over the w sets of log f(x_h | w), F[s][h] =
// exp(log f - m_h) in place (get_grid_para_like_ACD, codeml.c:6600-6606).
__global__ __launch_bounds__(256) void k_beb_norm(double *__restrict__ F, int nset, int npatt,
                                                  int npad, double *__restrict__ mx)
{
    const int h = blockIdx.x * 256 + threadIdx.x;
    if (h >= npatt) return;
    double m = F[h];
    for (int s = 1; s < nset; s++) m = fmax(m, F[(size_t)s * npad + h]);
    for (int s = 0; s < nset; s++) F[(size_t)s * npad + h] = exp(F[(size_t)s * npad + h] - m);
    mx[h] = m;
}

// One block per grid point g: lnfX[g] = sum_h fpatt_h log(sum_k pc[g][k]
// F[iw[g][k]][h]); patterns with a sum below 1e-300 are left out, as in the
// reference (lfunNSsites_ACD, codeml.c:6713-6722).  Fixed summation order.
__global__ __launch_bounds__(256) void k_beb_lnfx(const double *__restrict__ F, int npad,
                                                  int npatt, const double *__restrict__ fpatt,
                                                  const double *__restrict__ pc,
                                                  const int *__restrict__ iw,
                                                  double *__restrict__ lnfx)
{
    __shared__ double sred[4];
    const int g = blockIdx.x, tid = threadIdx.x;
    double p[4];
    const double *Fk[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        p[k] = pc[g * 4 + k];
        Fk[k] = F + (size_t)iw[g * 4 + k] * npad;
    }
    double acc = 0;
    for (int h = tid; h < npatt; h += 256) {
        double fh = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) fh += p[k] * Fk[k][h];
        if (fh >= 1e-300) acc += log(fh) * fpatt[h];
    }
    acc = wave_sum(acc);
    if ((tid & 63) == 0) sred[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) lnfx[g] = ((sred[0] + sred[1]) + sred[2]) + sred[3];
}

// Per pattern: post[k][h] = sum over listed grid points of
// (pc[g][k] F_k / sum_k' pc[g][k'] F_k') W_g, W_g = exp(lnfX[g] - log f(X))
// (codeml.c:6759-6769), grid points in ascending order.
__global__ __launch_bounds__(256) void k_beb_post(const double *__restrict__ F, int npad,
                                                  int npatt, const double *__restrict__ pc,
                                                  const int *__restrict__ iw,
                                                  const int *__restrict__ glist,
                                                  const double *__restrict__ W, int ng,
                                                  double *__restrict__ post)
{
    const int h = blockIdx.x * 256 + threadIdx.x;
    if (h >= npatt) return;
    double a[4] = {0, 0, 0, 0};
    for (int i = 0; i < ng; i++) {
        const int g = glist[i];
        double f[4], fh = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            f[k] = pc[g * 4 + k] * F[(size_t)iw[g * 4 + k] * npad + h];
            fh += f[k];
        }
        if (!(fh > 0)) continue;
        const double w = W[i] / fh;
#pragma unroll
        for (int k = 0; k < 4; k++) a[k] += f[k] * w;
    }
#pragma unroll
    for (int k = 0; k < 4; k++) post[(size_t)k * npatt + h] = a[k];
}

void launch_beb_norm(hipStream_t s, double *F, int nset, int npatt, int npad, double *mx)
{
    hipLaunchKernelGGL(k_beb_norm, dim3((npatt + 255) / 256), dim3(256), 0, s, F, nset, npatt, npad,
                       mx);
}

void launch_beb_lnfx(hipStream_t s, const double *F, int npad, int npatt, const double *fpatt,
                     const double *pc, const int *iw, int ngrid, double *lnfx)
{
    hipLaunchKernelGGL(k_beb_lnfx, dim3(ngrid), dim3(256), 0, s, F, npad, npatt, fpatt, pc, iw,
                       lnfx);
}

void launch_beb_post(hipStream_t s, const double *F, int npad, int npatt, const double *pc,
                     const int *iw, const int *glist, const double *W, int ng, double *post)
{
    hipLaunchKernelGGL(k_beb_post, dim3((npatt + 255) / 256), dim3(256), 0, s, F, npad, npatt, pc,
                       iw, glist, W, ng, post);
}

}  // namespace bsa
