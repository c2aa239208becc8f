// HIP kernels for gfx950 (MI355X) of the branch-site model A likelihood:
//   k_pmat        batched P(t) = U diag(exp(root t)) V for every (branch, eigen set)
//   k_amb         per-tip rows for ambiguous codons (sums of P columns)
//   k_prune_mfma  one tree level of Felsenstein pruning on the matrix cores
//   k_prune_valu  the same on plain vector FMA
//   k_root        per-pattern class log-likelihoods, mixture and weighted sum
//
// Layouts (T = double or float):
//   PT[pm][k][j]    = P_{j,k}(t) for matrix pm = node*3 + set, 64 x 64, rows and
//                     columns >= 61 are zero (k: child state, j: parent state)
//   L[slot][h][j]   conditional likelihood of a node / site class, pattern h,
//                   parent-side state j (64, lanes >= 61 stay exactly zero)
//   E[slot][h]      accumulated power-of-two exponent: true value = L * 2^E
#include "bsa_kernels.h"

#include <math.h>

namespace bsa {

typedef double d2 __attribute__((ext_vector_type(2)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kUVR = 64 * 64 * 2 + 64;   // per eigen set: U 64x64, V 64x64, root 64
constexpr double kLn2 = 0.69314718055994530942;

// ------------------------------------------------------------------ P(t)

// One block per (branch, set).  U is staged transposed so that the 64 lanes
// of a wave (64 consecutive parent states j) read consecutive LDS words; the
// wave's child-state index is uniform, so V reads are broadcasts.
template <typename T>
__global__ __launch_bounds__(256) void k_pmat(const double *__restrict__ uvr,
                                              const PmatJob *__restrict__ jobs,
                                              T *__restrict__ PT)
{
    __shared__ double Ut[64 * 64];   // Ut[m][j] = U[j][m]
    __shared__ double Ve[64 * 64];   // Ve[m][k] = exp(root_m t) V[m][k]
    const PmatJob jb = jobs[blockIdx.x];
    const double *U = uvr + (size_t)jb.set * kUVR;
    const double *V = U + 4096;
    const double *R = U + 8192;
    T *out = PT + (size_t)jb.pm * 4096;
    const int tid = threadIdx.x;
    if (jb.t < 1e-100) {   // PMatUVRoot returns the identity (tools.c:527-530)
        for (int q = tid; q < 4096; q += 256) {
            int k = q >> 6, j = q & 63;
            out[q] = (T)((k == j && k < 61) ? 1.0 : 0.0);
        }
        return;
    }
    for (int q = tid; q < 4096; q += 256) {
        int r = q >> 6, c = q & 63;
        Ut[c * 64 + r] = U[q];
        Ve[q] = V[q] * exp(jb.t * R[r]);
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
        double p = acc[q];
        if (p < 0) p = 0;   // negatives clamped (tools.c:536-538)
        out[(kq + 4 * q) * 64 + j] = (T)p;
    }
}

// Rows of ambiguous codes for tip branches: AT[pm][a][j] = sum_{k in a} P_{j,k}.
template <typename T>
__global__ __launch_bounds__(256) void k_amb(const PmatJob *__restrict__ jobs,
                                             const T *__restrict__ PT,
                                             const unsigned char *__restrict__ mask_amb, int namb,
                                             T *__restrict__ AT)
{
    const PmatJob jb = jobs[blockIdx.x];
    if (jb.tip < 0) return;
    const T *P = PT + (size_t)jb.pm * 4096;
    T *out = AT + (size_t)jb.pm * namb * 64;
    const int j = threadIdx.x & 63;
    for (int a = threadIdx.x >> 6; a < namb; a += 4) {
        double s = 0;
        for (int k = 0; k < 61; k++)
            if (mask_amb[a * 64 + k]) s += (double)P[k * 64 + j];
        out[a * 64 + j] = (T)s;
    }
}

// ------------------------------------------------------------------ helpers

__device__ inline double dev_frexp(double x, int *e) { return frexp(x, e); }
__device__ inline float dev_frexp(float x, int *e) { return frexpf(x, e); }
__device__ inline double dev_ldexp(double x, int e) { return ldexp(x, e); }
__device__ inline float dev_ldexp(float x, int e) { return ldexpf(x, e); }

template <typename T>
__device__ inline T tip_value(const T *P, const T *AT, int pm, int namb, int code, int j)
{
    return code < 61 ? P[code * 64 + j] : AT[((size_t)pm * namb + (code - 61)) * 64 + j];
}

template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef d4 V;
    static __device__ inline V op(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // 16 consecutive doubles from a 128-byte aligned address
    static __device__ inline void load16(const double *p, double *a)
    {
        const d2 *q = reinterpret_cast<const d2 *>(p);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            d2 v = q[i];
            a[2 * i] = v.x;
            a[2 * i + 1] = v.y;
        }
    }
};
template <> struct Mfma<float> {
    typedef f4 V;
    static __device__ inline V op(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline void load16(const float *p, float *a)
    {
        const f4 *q = reinterpret_cast<const f4 *>(p);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            f4 v = q[i];
            a[4 * i] = v.x;
            a[4 * i + 1] = v.y;
            a[4 * i + 2] = v.z;
            a[4 * i + 3] = v.w;
        }
    }
};

// ------------------------------------------------------------------ pruning (MFMA)

// Block = 4 waves = one tile of 64 patterns x 64 parent states of one
// (node, class) job; wave w owns patterns 16w..16w+15.  Interior children are
// a (16 x 64) x (64 x 64) product per wave as 16 k-steps x 4 column blocks of
// the 16x16x4 MFMA.  The k order inside the product is free, so lane group
// g = lane/16 takes child states 16g..16g+15 at steps 0..15: each lane then
// reads 16 consecutive values of its pattern row straight from HBM.
// Accumulator layout: lane l, element i of column block jb holds
// pattern 16w + 4(l/16) + i, parent state 16 jb + l%16.
template <typename T>
__global__ __launch_bounds__(256) void k_prune_mfma(const PruneJob *__restrict__ jobs, int npad,
                                                    const unsigned char *__restrict__ z,
                                                    const T *__restrict__ PT,
                                                    const T *__restrict__ AT, int namb,
                                                    T *__restrict__ L, int *__restrict__ E)
{
    typedef typename Mfma<T>::V V;
    __shared__ T sP[64 * 64];
    const PruneJob &jb = jobs[blockIdx.y];
    const int h0 = blockIdx.x * kTileP;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, g = lane >> 4, r = lane & 15;
    const int hrow = h0 + 16 * w + 4 * g;        // first of this lane's 4 output patterns
    V acc[4];
    int esum[4] = {0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < 4; b++) acc[b] = V{1, 1, 1, 1};
    const int nch = jb.nch;
    for (int c = 0; c < nch; c++) {
        const int pm = jb.ch_pmat[c];
        const T *P = PT + (size_t)pm * 4096;
        const int tip = jb.ch_tip[c];
        if (tip >= 0) {
            const unsigned char *zt = z + (size_t)tip * npad + hrow;
            int code[4];
#pragma unroll
            for (int i = 0; i < 4; i++) code[i] = zt[i];
#pragma unroll
            for (int b = 0; b < 4; b++)
#pragma unroll
                for (int i = 0; i < 4; i++)
                    acc[b][i] *= tip_value(P, AT, pm, namb, code[i], 16 * b + r);
        } else {
            const int slot = jb.ch_slot[c];
            __syncthreads();
            for (int q = tid; q < 4096; q += 256) sP[q] = P[q];
            __syncthreads();
            T a[16];
            Mfma<T>::load16(L + ((size_t)slot * npad + h0 + 16 * w + r) * 64 + 16 * g, a);
            V tmp[4];
#pragma unroll
            for (int b = 0; b < 4; b++) tmp[b] = V{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < 16; kk++) {
                const T *prow = sP + (16 * g + kk) * 64 + r;
#pragma unroll
                for (int b = 0; b < 4; b++) tmp[b] = Mfma<T>::op(a[kk], prow[16 * b], tmp[b]);
            }
#pragma unroll
            for (int b = 0; b < 4; b++) acc[b] *= tmp[b];
            const int *Ec = E + (size_t)slot * npad + hrow;
#pragma unroll
            for (int i = 0; i < 4; i++) esum[i] += Ec[i];
        }
    }
    // per-pattern power-of-two rescaling: max over the 64 states of a row lives
    // in the 4 column blocks of the 16 lanes of one lane group
    T *Lo = L + ((size_t)jb.out_slot * npad + hrow) * 64;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        T m = acc[0][i];
#pragma unroll
        for (int b = 1; b < 4; b++) m = acc[b][i] > m ? acc[b][i] : m;
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
            T o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        int e = 0;
        if (m > 0) dev_frexp(m, &e);
#pragma unroll
        for (int b = 0; b < 4; b++) Lo[i * 64 + 16 * b + r] = dev_ldexp(acc[b][i], -e);
        if (r == 0) E[(size_t)jb.out_slot * npad + hrow + i] = esum[i] + e;
    }
}

// ------------------------------------------------------------------ pruning (VALU)

// Same job on vector FMA: thread = parent state j (lane) x 16 patterns (wave).
template <typename T>
__global__ __launch_bounds__(256) void k_prune_valu(const PruneJob *__restrict__ jobs, int npad,
                                                    const unsigned char *__restrict__ z,
                                                    const T *__restrict__ PT,
                                                    const T *__restrict__ AT, int namb,
                                                    T *__restrict__ L, int *__restrict__ E)
{
    __shared__ T sP[64 * 64];
    __shared__ T sA[64 * 64];
    const PruneJob &jb = jobs[blockIdx.y];
    const int h0 = blockIdx.x * kTileP;
    const int tid = threadIdx.x, j = tid & 63, pg = tid >> 6;
    const int hb = h0 + 16 * pg;
    T acc[16];
    int esum[16];
#pragma unroll
    for (int i = 0; i < 16; i++) { acc[i] = 1; esum[i] = 0; }
    const int nch = jb.nch;
    for (int c = 0; c < nch; c++) {
        const int pm = jb.ch_pmat[c];
        const T *P = PT + (size_t)pm * 4096;
        const int tip = jb.ch_tip[c];
        if (tip >= 0) {
            const unsigned char *zt = z + (size_t)tip * npad + hb;
#pragma unroll
            for (int i = 0; i < 16; i++) acc[i] *= tip_value(P, AT, pm, namb, (int)zt[i], j);
        } else {
            const int slot = jb.ch_slot[c];
            const T *A = L + ((size_t)slot * npad + h0) * 64;
            __syncthreads();
            for (int q = tid; q < 4096; q += 256) { sP[q] = P[q]; sA[q] = A[q]; }
            __syncthreads();
            T tmp[16];
#pragma unroll
            for (int i = 0; i < 16; i++) tmp[i] = 0;
            for (int k = 0; k < 64; k++) {
                T p = sP[k * 64 + j];
#pragma unroll
                for (int i = 0; i < 16; i++) tmp[i] = fma(sA[(16 * pg + i) * 64 + k], p, tmp[i]);
            }
            const int *Ec = E + (size_t)slot * npad + hb;
#pragma unroll
            for (int i = 0; i < 16; i++) { acc[i] *= tmp[i]; esum[i] += Ec[i]; }
        }
    }
    T *Lo = L + ((size_t)jb.out_slot * npad + hb) * 64;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        T m = acc[i];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            T o = __shfl_xor(m, off, 64);
            m = o > m ? o : m;
        }
        int e = 0;
        if (m > 0) dev_frexp(m, &e);
        Lo[i * 64 + j] = dev_ldexp(acc[i], -e);
        if (j == 0) E[(size_t)jb.out_slot * npad + hb + i] = esum[i] + e;
    }
}

// ------------------------------------------------------------------ root

// One wave per pattern at a time, lane = state.  Per class: f = sum_i pi_i L_i
// (<= 0 -> 1e-300, treesub.c:7522-7533), l_c = log f + E ln 2; site value
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
__global__ __launch_bounds__(256) void k_root(RootArgs a, const T *__restrict__ L,
                                              const int *__restrict__ E,
                                              const double *__restrict__ pi,
                                              const double *__restrict__ fpatt,
                                              double *__restrict__ fh, double *__restrict__ site,
                                              double *__restrict__ partial)
{
    __shared__ double ss[4], sc[4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double pil = lane < 61 ? pi[lane] : 0.0;
    double s = 0, comp = 0;
    for (int q = 0; q < 16; q++) {
        const int h = blockIdx.x * 64 + w * 16 + q;
        if (h >= a.npatt) break;
        double l[4];
#pragma unroll
        for (int c = 0; c < 4; c++) {
            const size_t row = (size_t)a.root_slot[c] * a.npad + h;
            double v = wave_sum(pil * (double)L[row * 64 + lane]);
            if (v <= 0) v = 1e-300;
            l[c] = log(v) + (double)E[row] * kLn2;
        }
        double m = l[0];
#pragma unroll
        for (int c = 1; c < 4; c++) m = l[c] > m ? l[c] : m;
        double f = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) f += a.freqK[c] * exp(l[c] - m);
        const double st = m + log(f);
        if (lane == 0) {
#pragma unroll
            for (int c = 0; c < 4; c++) fh[(size_t)c * a.npad + h] = l[c];
            site[h] = st;
            if (fpatt[h] > 0) neumaier(s, comp, fpatt[h] * st);
        }
    }
    if (lane == 0) { ss[w] = s; sc[w] = comp; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double S = 0, Cc = 0;
        for (int i = 0; i < 4; i++) { neumaier(S, Cc, ss[i]); Cc += sc[i]; }
        partial[2 * blockIdx.x] = S;
        partial[2 * blockIdx.x + 1] = Cc;
    }
}

__global__ void k_final(const double *__restrict__ partial, int nblk, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double S = 0, Cc = 0;
    for (int i = 0; i < nblk; i++) { neumaier(S, Cc, partial[2 * i]); Cc += partial[2 * i + 1]; }
    out[0] = S + Cc;
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

template <typename T>
void launch_pmat(hipStream_t s, const double *uvr, const PmatJob *jobs, int njobs, T *PT,
                 const unsigned char *mask_amb, int namb, T *AT)
{
    if (njobs <= 0) return;
    hipLaunchKernelGGL(k_pmat<T>, dim3(njobs), dim3(256), 0, s, uvr, jobs, PT);
    if (namb > 0)
        hipLaunchKernelGGL(k_amb<T>, dim3(njobs), dim3(256), 0, s, jobs, (const T *)PT, mask_amb,
                           namb, AT);
}

template <typename T>
void launch_prune(hipStream_t s, int kernel, const PruneJob *jobs, int njobs, int npad,
                  const unsigned char *z, const T *PT, const T *AT, int namb, T *L, int *E)
{
    if (njobs <= 0) return;
    dim3 grid(npad / kTileP, njobs);
    if (kernel == 0)
        hipLaunchKernelGGL(k_prune_mfma<T>, grid, dim3(256), 0, s, jobs, npad, z, PT, AT, namb, L, E);
    else
        hipLaunchKernelGGL(k_prune_valu<T>, grid, dim3(256), 0, s, jobs, npad, z, PT, AT, namb, L, E);
}

template <typename T>
void launch_root(hipStream_t s, const RootArgs &a, const T *L, const int *E, const double *pi,
                 const double *fpatt, double *fh, double *site, double *partial, int nblk,
                 double *out)
{
    hipLaunchKernelGGL(k_root<T>, dim3(nblk), dim3(256), 0, s, a, L, E, pi, fpatt, fh, site, partial);
    hipLaunchKernelGGL(k_final, dim3(1), dim3(64), 0, s, (const double *)partial, nblk, out);
}

void launch_mfma_probe(hipStream_t s, const double *A, const double *B, double *D64,
                       const float *Af, const float *Bf, float *D32)
{
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, s, A, B, D64, Af, Bf, D32);
}

template void launch_pmat<double>(hipStream_t, const double *, const PmatJob *, int, double *,
                                  const unsigned char *, int, double *);
template void launch_pmat<float>(hipStream_t, const double *, const PmatJob *, int, float *,
                                 const unsigned char *, int, float *);
template void launch_prune<double>(hipStream_t, int, const PruneJob *, int, int,
                                   const unsigned char *, const double *, const double *, int,
                                   double *, int *);
template void launch_prune<float>(hipStream_t, int, const PruneJob *, int, int,
                                  const unsigned char *, const float *, const float *, int,
                                  float *, int *);
template void launch_root<double>(hipStream_t, const RootArgs &, const double *, const int *,
                                  const double *, const double *, double *, double *, double *,
                                  int, double *);
template void launch_root<float>(hipStream_t, const RootArgs &, const float *, const int *,
                                 const double *, const double *, double *, double *, double *,
                                 int, double *);

}  // namespace bsa
