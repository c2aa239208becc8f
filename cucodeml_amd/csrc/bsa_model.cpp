// Host FP64 model set-up for branch-site model A: GY94 rate matrices, the
// symmetric eigen decomposition they are exponentiated through, Qfactor
// branch scaling and the site-class table.  Everything here is O(61^3) per
// likelihood evaluation and runs on the host in double precision.
#include "bsa_model.h"

#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <utility>

#include <immintrin.h>
#include <unistd.h>

namespace bsa {

static const char kAA64[] =
    "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";

CodonTables::CodonTables()
{
    int k = 0;
    for (int c = 0; c < 64; c++) {
        from64[c] = -1;
        if (kAA64[c] != '*') { from61[k] = c; from64[c] = k; k++; }
    }
    // kind[i][j]: 0 = not a single-nucleotide change; else bit0 transition,
    // bit1 nonsynonymous, +4 marks a one-step neighbour
    for (int i = 0; i < kS; i++)
        for (int j = 0; j < kS; j++) {
            int a = from61[i], b = from61[j];
            int fa[3] = {a / 16, (a / 4) % 4, a % 4}, fb[3] = {b / 16, (b / 4) % 4, b % 4};
            int nd = 0, pos = 0;
            for (int p = 0; p < 3; p++)
                if (fa[p] != fb[p]) { nd++; pos = p; }
            unsigned char kd = 0;
            if (i != j && nd == 1) {
                kd = 4;
                int s = fa[pos] + fb[pos];
                if (s == 1 || s == 5) kd |= 1;          // T<->C or A<->G
                if (kAA64[a] != kAA64[b]) kd |= 2;
            }
            kind[i][j] = kd;
        }
}

const CodonTables &tables()
{
    static const CodonTables t;
    return t;
}

// Q_ij = q_ij pi_j with q = kappa^[transition] * omega^[nonsynonymous] for
// one-step neighbours (codeml.c:3147-3186); returns the mean rate
// -sum_i pi_i Q_ii.
double build_q(double kappa, double omega, const double *pi, double *Q)
{
    const CodonTables &T = tables();
    for (int i = 0; i < kS; i++) {
        double rs = 0;
        for (int j = 0; j < kS; j++) {
            unsigned char kd = T.kind[i][j];
            double q = 0;
            if (kd) {
                q = 1;
                if (kd & 1) q = kappa;
                if (kd & 2) q *= omega;
                q *= pi[j];
            }
            Q[i * kS + j] = q;
            rs += q;
        }
        Q[i * kS + i] = -rs;
    }
    double mr = 0;
    for (int i = 0; i < kS; i++) mr -= pi[i] * Q[i * kS + i];
    return mr;
}

double mean_rate(double kappa, double omega, const double *pi)
{
    const CodonTables &T = tables();
    double mr = 0;
    for (int i = 0; i < kS; i++) {
        double rs = 0;
        for (int j = 0; j < kS; j++) {
            unsigned char kd = T.kind[i][j];
            if (!kd) continue;
            double q = (kd & 1) ? kappa : 1.0;
            if (kd & 2) q *= omega;
            rs += q * pi[j];
        }
        mr += pi[i] * rs;
    }
    return mr;
}

// Householder reduction of the symmetric matrix in z (n x n, row-major) to
// tridiagonal form (diagonal d, sub-diagonal e), accumulating the orthogonal
// transform in z; then the implicit-shift QL iteration on the tridiagonal
// matrix, rotating z into the eigenvectors (columns).  This is the classic
// Householder + QL pair of algorithms (Wilkinson & Reinsch, Handbook for
// Automatic Computation vol. II, procedures tred2 and tql2).  A QL sweep's
// rotations depend only on d and e, so they are applied to the vectors while
// the next sweep's are computed.
static bool sym_eigen(int n, double *z, double *d, double *e)
{
    // The reduction reads and updates columns of the (symmetric) input and
    // of the accumulated transform; it runs on the transposed view so that
    // those are contiguous rows.  Afterwards row v of the buffer holds
    // column v of the transform, which is the layout the QL rotations want.
#define Z(i, j) z[(j) * n + (i)]
    for (int j = 0; j < n; j++) d[j] = Z(n - 1, j);
    for (int i = n - 1; i > 0; i--) {
        double scale = 0, h = 0;
        for (int k = 0; k < i; k++) scale += std::fabs(d[k]);
        if (scale == 0) {
            e[i] = d[i - 1];
            for (int j = 0; j < i; j++) {
                d[j] = Z(i - 1, j);
                Z(i, j) = 0;
                Z(j, i) = 0;
            }
        } else {
            for (int k = 0; k < i; k++) {
                d[k] /= scale;
                h += d[k] * d[k];
            }
            double f = d[i - 1];
            double g = std::sqrt(h);
            if (f > 0) g = -g;
            e[i] = scale * g;
            h -= f * g;
            d[i - 1] = f - g;
            for (int j = 0; j < i; j++) e[j] = 0;
            for (int j = 0; j < i; j++) {
                f = d[j];
                Z(j, i) = f;
                g = e[j] + Z(j, j) * f;
                for (int k = j + 1; k <= i - 1; k++) {
                    g += Z(k, j) * d[k];
                    e[k] += Z(k, j) * f;
                }
                e[j] = g;
            }
            f = 0;
            for (int j = 0; j < i; j++) {
                e[j] /= h;
                f += e[j] * d[j];
            }
            double hh = f / (h + h);
            for (int j = 0; j < i; j++) e[j] -= hh * d[j];
            for (int j = 0; j < i; j++) {
                f = d[j];
                g = e[j];
                for (int k = j; k <= i - 1; k++) Z(k, j) -= (f * e[k] + g * d[k]);
                d[j] = Z(i - 1, j);
                Z(i, j) = 0;
            }
        }
        d[i] = h;
    }
    for (int i = 0; i < n - 1; i++) {
        Z(n - 1, i) = Z(i, i);
        Z(i, i) = 1;
        double h = d[i + 1];
        if (h != 0) {
            for (int k = 0; k <= i; k++) d[k] = Z(k, i + 1) / h;
            for (int j = 0; j <= i; j++) {
                double g = 0;
                for (int k = 0; k <= i; k++) g += Z(k, i + 1) * Z(k, j);
                for (int k = 0; k <= i; k++) Z(k, j) -= g * d[k];
            }
        }
        for (int k = 0; k <= i; k++) Z(k, i + 1) = 0;
    }
    for (int j = 0; j < n; j++) {
        d[j] = Z(n - 1, j);
        Z(n - 1, j) = 0;
    }
    Z(n - 1, n - 1) = 1;
    e[0] = 0;
#undef Z
    // ZT(v, k) = component k of vector v
#define ZT(v, k) z[(v) * n + (k)]

    // QL with implicit shifts.  The rotations of one sweep are applied to the
    // vectors while the next sweep's rotations are computed: that chain is
    // bound by the latency of its square roots and divisions, the vector
    // updates by throughput, so the two overlap.  Every vector component
    // still sees the same rotations in the same order.
    for (int i = 1; i < n; i++) e[i - 1] = e[i];
    e[n - 1] = 0;
    if (n > 64) return false;   // rotation buffers below hold 64 (n <= 61 here)
    double cs[2][64][2];
    int cur = 0;           // the sweep being computed writes cs[cur]
    int pj = -1, pl = 0;   // rotations pj, pj - 1, .., pl of cs[cur ^ 1] still to apply
    auto apply_one = [&]() {
        const double c = cs[cur ^ 1][pj][0], s = cs[cur ^ 1][pj][1];
        double *zi = &ZT(pj, 0), *zi1 = &ZT(pj + 1, 0);
        for (int k = 0; k < n; k++) {
            const double a = zi[k], b = zi1[k];
            zi1[k] = s * a + c * b;
            zi[k] = c * a - s * b;
        }
        pj--;
    };
    double f = 0, tst1 = 0;
    const double eps = std::ldexp(1.0, -52);
    for (int l = 0; l < n; l++) {
        tst1 = std::fmax(tst1, std::fabs(d[l]) + std::fabs(e[l]));
        int m = l;
        while (m < n) {
            if (std::fabs(e[m]) <= eps * tst1) break;
            m++;
        }
        if (m > l) {
            int iter = 0;
            do {
                if (++iter > 60) return false;
                double g = d[l];
                double p = (d[l + 1] - g) / (2 * e[l]);
                double r = std::hypot(p, 1.0);
                if (p < 0) r = -r;
                d[l] = e[l] / (p + r);
                d[l + 1] = e[l] * (p + r);
                double dl1 = d[l + 1];
                double h = g - d[l];
                for (int i = l + 2; i < n; i++) d[i] -= h;
                f += h;
                p = d[m];
                double c = 1, c2 = 1, c3 = 1, el1 = e[l + 1], s = 0, s2 = 0;
                for (int i = m - 1; i >= l; i--) {
                    c3 = c2;
                    c2 = c;
                    s2 = s;
                    g = c * e[i];
                    h = c * p;
                    // p and e[i] are of the order of the matrix norm here, so
                    // the plain square root cannot overflow (hypot is ~3x slower)
                    r = std::sqrt(p * p + e[i] * e[i]);
                    e[i + 1] = s * r;
                    s = e[i] / r;
                    c = p / r;
                    p = c * d[i] - s * g;
                    d[i + 1] = h + s * (c * g + s * d[i]);
                    cs[cur][i][0] = c;
                    cs[cur][i][1] = s;
                    if (pj >= pl) apply_one();
                }
                while (pj >= pl) apply_one();   // the rest of the previous sweep
                pj = m - 1;                     // this sweep's rotations are next
                pl = l;
                cur ^= 1;
                p = -s * s2 * c3 * el1 * e[l] / dl1;
                e[l] = s * p;
                d[l] = c * p;
            } while (std::fabs(e[l]) > eps * tst1);
        }
        d[l] = d[l] + f;
        e[l] = 0;
    }
    while (pj >= pl) apply_one();
    // sort descending (eigenvalues and vector columns)
    for (int i = 0; i < n - 1; i++) {
        int k = i;
        double p = d[i];
        for (int j = i + 1; j < n; j++)
            if (d[j] > p) { k = j; p = d[j]; }
        if (k != i) {
            d[k] = d[i];
            d[i] = p;
            for (int j = 0; j < n; j++) std::swap(ZT(i, j), ZT(k, j));
        }
    }
    // back to eigenvectors in columns
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) std::swap(ZT(i, j), ZT(j, i));
    return true;
#undef ZT
}

// Q = U diag(root) V with UV = I, via the symmetric S = Pi^1/2 Q Pi^-1/2;
// codons with pi = 0 are split off with root 0 and identity rows
// (eigenQREV, tools.c:4850-4937).  U, V: kS x kS row-major.
bool eigen_rev(const double *Q, const double *pi, double *root, double *U, double *V)
{
    int idx[kS], m = 0;
    double sq[kS];
    for (int i = 0; i < kS; i++)
        if (pi[i] > 1e-100) { idx[m] = i; sq[m] = std::sqrt(pi[i]); m++; }
    double A[kS * kS], d[kS], e[kS];
    for (int i = 0; i < m; i++) {
        A[i * m + i] = Q[idx[i] * kS + idx[i]];
        for (int j = 0; j < i; j++) {
            double s = Q[idx[i] * kS + idx[j]] * sq[i] / sq[j];
            A[i * m + j] = s;
            A[j * m + i] = s;
        }
    }
    if (!sym_eigen(m, A, d, e)) return false;
    std::memset(U, 0, sizeof(double) * kS * kS);
    std::memset(V, 0, sizeof(double) * kS * kS);
    for (int i = 0; i < kS; i++) root[i] = 0;
    for (int i = 0; i < kS; i++)
        if (pi[i] <= 1e-100) { U[i * kS + i] = 1; V[i * kS + i] = 1; }
    for (int k = 0; k < m; k++) root[idx[k]] = d[k];
    for (int i = 0; i < m; i++)
        for (int k = 0; k < m; k++) {
            U[idx[i] * kS + idx[k]] = A[i * m + k] / sq[i];
            V[idx[k] * kS + idx[i]] = A[i * m + k] * sq[i];
        }
    for (int i = 0; i < kS; i++)
        if (!std::isfinite(root[i])) return false;
    return true;
}

// Two helper threads for the eigen systems of one evaluation, shared by the
// process.  After a task a helper spins for kSpin waiting for the next one
// (an optimiser asks every few milliseconds, and handing a task to a
// spinning thread takes about a microsecond, where starting a thread takes
// tens), then sleeps on a condition variable until it is woken.  One caller
// at a time: try_run returns false at once while another caller holds them.
// The object is never destroyed, so helpers parked at process exit are
// harmless.  A child made by fork() has no helper threads, so there
// try_run always declines.
class EigenHelpers {
public:
    static EigenHelpers &get()
    {
        static EigenHelpers *h = new EigenHelpers();
        return *h;
    }
    // fn(0) on the caller, fn(1) .. fn(n - 1) (n <= 3) on the helpers
    bool try_run(int n, const std::function<void(int)> &fn)
    {
        if (n < 1 || n > 3 || getpid() != pid_ || busy_.exchange(true, std::memory_order_acquire))
            return false;
        for (int k = 1; k < n; k++) {
            Helper &h = helper_[k - 1];
            h.fn = &fn;
            h.arg = k;
            {
                std::lock_guard<std::mutex> lk(h.mu);
                h.posted.store(h.posted.load(std::memory_order_relaxed) + 1,
                               std::memory_order_release);
            }
            h.cv.notify_one();
        }
        fn(0);
        for (int k = 1; k < n; k++) {
            Helper &h = helper_[k - 1];
            const unsigned long want = h.posted.load(std::memory_order_relaxed);
            while (h.done.load(std::memory_order_acquire) != want) _mm_pause();
        }
        busy_.store(false, std::memory_order_release);
        return true;
    }

private:
    static constexpr std::chrono::microseconds kSpin{10000};
    struct Helper {
        std::atomic<unsigned long> posted{0}, done{0};
        const std::function<void(int)> *fn = nullptr;
        int arg = 0;
        std::mutex mu;
        std::condition_variable cv;
    };
    EigenHelpers() : pid_(getpid())
    {
        for (Helper &h : helper_) std::thread([&h] { serve(h); }).detach();
    }
    static void serve(Helper &h)
    {
        unsigned long seen = 0;
        for (;;) {
            const auto t0 = std::chrono::steady_clock::now();
            int polls = 0;
            while (h.posted.load(std::memory_order_acquire) == seen) {
                _mm_pause();
                if (++polls % 256 == 0 && std::chrono::steady_clock::now() - t0 > kSpin) {
                    std::unique_lock<std::mutex> lk(h.mu);
                    h.cv.wait(lk, [&] { return h.posted.load(std::memory_order_acquire) != seen; });
                }
            }
            seen = h.posted.load(std::memory_order_acquire);
            (*h.fn)(h.arg);
            h.done.store(seen, std::memory_order_release);
        }
    }
    const pid_t pid_;
    std::atomic<bool> busy_{false};
    Helper helper_[2];
};

// Model A at natural parameters (SetParametersNSsites, codeml.c:2470-2520):
// proportions (p0, p1, (1-p0-p1)p0/(p0+p1), (1-p0-p1)p1/(p0+p1)); Qfactor of
// background branches from the mean rate at the background-average omega
// (p0 w0 + p1)/(p0+p1), of foreground branches at p0 w0 + p1 + (1-p0-p1) w2;
// eigen systems for omega = w0, 1, w2.
bool setup_model_a(double kappa, double p0, double p1, double w0, double w2, const double *pi,
                   ModelA *m, bool need_eigen, bool threads)
{
    double t = p0 + p1;
    if (!(t > 1e-100)) return false;
    m->kappa = kappa; m->p0 = p0; m->p1 = p1; m->w0 = w0; m->w2 = w2;
    m->freqK[0] = p0;
    m->freqK[1] = p1;
    m->freqK[2] = (1 - t) * p0 / t;
    m->freqK[3] = (1 - t) * p1 / t;
    m->wbar[0] = (p0 / t) * w0 + (p1 / t) * 1.0;
    m->wbar[1] = p0 * w0 + p1 * 1.0 + (1 - t) * w2;
    m->mr[0] = mean_rate(kappa, m->wbar[0], pi);
    m->mr[1] = mean_rate(kappa, m->wbar[1], pi);
    m->Qfactor[0] = 1 / m->mr[0];
    m->Qfactor[1] = 1 / m->mr[1];
    if (!need_eigen) return true;
    // the eigen systems of the distinct omegas are independent: all but the
    // first on helper threads.  A set whose omega equals another's (w2 = 1
    // under H0, or w0 = 1) takes that set's system.
    const double w[3] = {w0, 1.0, w2};
    int src[3] = {0, 1, 2};
    if (w0 == 1.0) src[0] = 1;
    if (w2 == 1.0) src[2] = 1;
    else if (w2 == w0) src[2] = src[0];
    bool ok[3] = {false, false, false};
    auto one = [&](int s) {
        double Q[kS * kS];
        build_q(kappa, w[s], pi, Q);
        ok[s] = eigen_rev(Q, pi, m->root[s], m->U[s], m->V[s]);
    };
    int own[3], nown = 0;
    for (int s = 0; s < 3; s++)
        if (src[s] == s) own[nown++] = s;
    if (threads && nown > 1) {
        const std::function<void(int)> task = [&](int k) { one(own[k]); };
        if (!EigenHelpers::get().try_run(nown, task)) {
            // the helpers serve another engine right now: threads of our own
            std::thread helpers[2];
            for (int k = 1; k < nown; k++) helpers[k - 1] = std::thread(one, own[k]);
            one(own[0]);
            for (int k = 1; k < nown; k++) helpers[k - 1].join();
        }
    } else {
        for (int k = 0; k < nown; k++) one(own[k]);
    }
    for (int s = 0; s < 3; s++)
        if (src[s] != s) {
            std::memcpy(m->root[s], m->root[src[s]], sizeof m->root[s]);
            std::memcpy(m->U[s], m->U[src[s]], sizeof m->U[s]);
            std::memcpy(m->V[s], m->V[src[s]], sizeof m->V[s]);
            ok[s] = ok[src[s]];
        }
    return ok[0] && ok[1] && ok[2];
}

// <Z, dQ/dkappa> and <Z, dQ/domega> at Q(kappa, omega); Z has row stride 64.
// Off-diagonal dQ_ij/dtheta contributes dQ_ij (Z_ij - Z_ii) because the
// diagonal is minus the row sum.  Codons with pi = 0 are held fixed by the
// eigen split (identity rows), so their rows and columns are skipped.
static void q_param_grads(double kappa, double omega, const double *pi, const double *Z,
                          double *gk, double *gw)
{
    const CodonTables &T = tables();
    double sk = 0, sw = 0;
    for (int i = 0; i < kS; i++) {
        if (!(pi[i] > 1e-100)) continue;
        for (int j = 0; j < kS; j++) {
            unsigned char kd = T.kind[i][j];
            if (!kd || !(pi[j] > 1e-100)) continue;
            const double dz = Z[i * 64 + j] - Z[i * 64 + i];
            const double w = (kd & 2) ? omega : 1.0, k = (kd & 1) ? kappa : 1.0;
            if (kd & 1) sk += dz * w * pi[j];
            if (kd & 2) sw += dz * k * pi[j];
        }
    }
    *gk = sk;
    *gw = sw;
}

// d mr / d kappa and d mr / d omega of the mean rate sum_i pi_i sum_j Q_ij
static void mean_rate_grads(double kappa, double omega, const double *pi, double *mk, double *mw)
{
    const CodonTables &T = tables();
    double a = 0, b = 0;
    for (int i = 0; i < kS; i++)
        for (int j = 0; j < kS; j++) {
            unsigned char kd = T.kind[i][j];
            if (!kd) continue;
            const double pp = pi[i] * pi[j];
            if (kd & 1) a += pp * ((kd & 2) ? omega : 1.0);
            if (kd & 2) b += pp * ((kd & 1) ? kappa : 1.0);
        }
    *mk = a;
    *mw = b;
}

void model_a_param_grad(const ModelA &m, const double *pi, const double *Z, const double *dQf,
                        const double *Sf, ModelGrad *g)
{
    const double w[3] = {m.w0, 1.0, m.w2};
    double gk = 0, gw[3] = {0, 0, 0};
    for (int s = 0; s < 3; s++) {
        double k, ww;
        q_param_grads(m.kappa, w[s], pi, Z + (size_t)s * 4096, &k, &ww);
        gk += k;
        gw[s] = ww;
    }
    const double p0 = m.p0, p1 = m.p1, t = p0 + p1;
    // Qfactor_l = 1 / mr(kappa, wbar_l)
    double mk[2], mw[2];
    for (int l = 0; l < 2; l++) mean_rate_grads(m.kappa, m.wbar[l], pi, &mk[l], &mw[l]);
    const double cQ[2] = {-dQf[0] * m.Qfactor[0] * m.Qfactor[0],
                          -dQf[1] * m.Qfactor[1] * m.Qfactor[1]};   // d lnL / d mr_l
    // wbar_bg = (p0 w0 + p1) / t ; wbar_fg = p0 w0 + p1 + (1 - t) w2
    const double db_dw0 = p0 / t, db_dp0 = p1 * (m.w0 - 1) / (t * t), db_dp1 = p0 * (1 - m.w0) / (t * t);
    const double df_dw0 = p0, df_dw2 = 1 - t, df_dp0 = m.w0 - m.w2, df_dp1 = 1 - m.w2;
    g->kappa = gk + cQ[0] * mk[0] + cQ[1] * mk[1];
    g->w0 = gw[0] + cQ[0] * mw[0] * db_dw0 + cQ[1] * mw[1] * df_dw0;
    g->w2 = gw[2] + cQ[1] * mw[1] * df_dw2;
    // class proportions (p0, p1, (1-t) p0 / t, (1-t) p1 / t)
    const double t2 = t * t;
    g->p0 = Sf[0] + Sf[2] * (p1 / t2 - 1) - Sf[3] * (p1 / t2) + cQ[0] * mw[0] * db_dp0 +
            cQ[1] * mw[1] * df_dp0;
    g->p1 = Sf[1] - Sf[2] * (p0 / t2) + Sf[3] * (p0 / t2 - 1) + cQ[0] * mw[0] * db_dp1 +
            cQ[1] * mw[1] * df_dp1;
}

}  // namespace bsa
