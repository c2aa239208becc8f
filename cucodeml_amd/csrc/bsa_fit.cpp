// Bounded quasi-Newton maximum-likelihood fit of branch-site model A on one
// engine: the optimiser that takes the place of the reference's ming2
// (tools.c:6361-6545) at its call in the ML driver (codeml.c:688), with the
// engine's adjoint gradient (or its forward differences, gradientB
// tools.c:6326-6350) instead of the reference's finite differences.
//
// Method (an active-set projected BFGS; Nocedal & Wright, Numerical
// Optimization, 2nd ed., sections 6.1 and 16.7):
//   * x is kept inside the box [lo, hi]; a variable sitting on a bound whose
//     gradient points out of the box is held there for the iteration, and
//     one with lo == hi is held there throughout;
//   * direction d = -H g on the free variables, H the inverse-Hessian
//     approximation (identity, scaled by s'y / y'y after the first step, as
//     in N&W eq. 6.20);
//   * projected backtracking line search: x(a) = clamp(x + a d), Armijo
//     condition f(x(a)) <= f + 1e-4 g'(x(a) - x), step cut by safeguarded
//     quadratic interpolation; a trial costs f alone, the gradient is asked
//     for only where the trial is accepted;
//   * BFGS inverse update restricted to the free variables, skipped when the
//     curvature s'y is not positive;
//   * stop (the reason is returned, BSA_STOP_*):
//       BSA_STOP_PGRAD  the projected gradient is below
//                       gtol = 1e-8 + 1e-10 max(1, |f|), about what an FP64
//                       gradient of a sum over sites of size |f| resolves;
//       BSA_STOP_HEND   two successive steps change f by less than e |f| and
//                       x by less than e |x| (the reference's H_end rule,
//                       tools.c:5724-5740), confirmed once more after
//                       restarting H from the scaled identity;
//       BSA_STOP_STALL  f has fallen by less than 1e-3 e |f| over the last
//                       50 iterations (on degenerate likelihood surfaces,
//                       e.g. p0 -> 1 where w2 and p1 no longer matter, x can
//                       keep drifting along a flat direction without an
//                       H_end stop; the 1e-3 keeps slow but real progress
//                       along a ridge going);
//       BSA_STOP_NODESCENT  no descent step along the scaled steepest
//                       descent (or a zero direction on the free variables);
//       BSA_STOP_LIMIT  the iteration limit.
//   * optional Newton finish (opts->newton > 0 iterations, after the
//     quasi-Newton phase, which max_iter < 0 skips): the Hessian of the
//     free variables by forward differences of the gradient (one gradient
//     evaluation per free variable, steps of 1e-5 max(1, |x_i|) towards the
//     inside of the box), Levenberg damping until it is positive definite,
//     Newton step with a projected Armijo line search.  With more than 160
//     free variables the step comes from conjugate gradients on
//     Hessian-vector products instead (forward differences of the gradient
//     along each direction), preconditioned by the exact Hessian block of
//     the last nblock variables (codeml's model parameters) and a scalar on
//     the rest: at most nblock + 32 gradient differences per step.  Where
//     the last step is below 1e-6 max(1, |x_i|) in every variable, the fit
//     stops with BSA_STOP_NEWTON: the point is the maximum to that
//     precision, also in flat directions (w2 on a ridge) that BFGS's
//     curvature estimate, and the H_end rule, resolve poorly.
//     With `verify` (exact FP64 gradients), a stall or no-descent exit is
//     checked the same way by up to four Newton iterations.
//     BSA_STOP_HEND, BSA_STOP_PGRAD and BSA_STOP_NEWTON are convergence.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <vector>

#include "bsa.h"
#include "bsa_fit.h"

int bsa_error(int code, const char *msg);   // sets bsa_last_error (engine)
int bsa_engine_precision(const bsa_engine *eng);   // BSA_FP64 / BSA_FP32 (engine)
bool bsa_engine_graph(const bsa_engine *eng);      // evaluations replayed from graphs
int bsa_engine_nbranch(const bsa_engine *eng);     // branch lengths at the head of x
// an evaluation with gradient in two calls (engine): -lnL at x, then the
// gradient at that x
int bsa_engine_lnl_part1(bsa_engine *eng, const double *x, int np, double *neg_lnl);
int bsa_engine_grad_part2(bsa_engine *eng, int np, double *grad);

namespace {
inline double clampd(double v, double a, double b) { return v < a ? a : (v > b ? b : v); }
}  // namespace

namespace bsa {

// Newton iterations on more free variables than this take their step from
// preconditioned conjugate gradients on Hessian-vector products (one
// gradient each) instead of the Hessian itself (one gradient per variable:
// over a minute per iteration at a thousand variables and 60 ms a gradient)
constexpr int kDenseNewtonMax = 160;
constexpr int kCgMax = 32;
constexpr double kCgTol = 1e-4;

// Hessian of the free variables fr at x by forward differences of the
// gradient, column by column, and the Newton step d from (H + mu I) d = -g,
// mu raised from 0 until the Cholesky factorisation succeeds.  Returns
// BSA_OK with *ok = false where no step is found (a failed evaluation, or
// no positive definite H + mu I).
template <class FG>
static int dense_step(FG &fg, int n, const double *x, const double *lo, const double *hi,
                      const std::vector<int> &fr, const std::vector<double> &gf,
                      std::vector<double> &dd, bool *ok)
{
    const int m = (int)fr.size();
    std::vector<double> xt(n), gc(n), Hm((size_t)m * m), L((size_t)m * m);
    *ok = false;
    for (int a = 0; a < m; a++) {
        const int i = fr[a];
        double h = 1e-5 * std::max(1.0, std::fabs(x[i]));
        if (x[i] + h > hi[i]) h = -h;
        if (x[i] + h < lo[i]) h = 0.5 * (hi[i] - lo[i]) * (x[i] - lo[i] < hi[i] - x[i] ? 1 : -1);
        std::memcpy(xt.data(), x, sizeof(double) * n);
        xt[i] += h;
        double ft;
        int rc = fg(xt.data(), &ft, gc.data());
        if (rc) return rc == BSA_ENUMERIC ? BSA_OK : rc;
        for (int b = 0; b < m; b++) Hm[(size_t)b * m + a] = (gc[fr[b]] - gf[b]) / h;
    }
    double dmax = 0;
    for (int a = 0; a < m; a++) {
        for (int b = 0; b < a; b++) {
            const double v = 0.5 * (Hm[(size_t)a * m + b] + Hm[(size_t)b * m + a]);
            Hm[(size_t)a * m + b] = Hm[(size_t)b * m + a] = v;
        }
        dmax = std::max(dmax, std::fabs(Hm[(size_t)a * m + a]));
    }
    bool pd = false;
    for (double mu = 0; !pd && mu < 1e6 * (dmax + 1); mu = mu ? 10 * mu : 1e-12 * (dmax + 1)) {
        pd = true;
        for (int a = 0; a < m && pd; a++)
            for (int b = 0; b <= a; b++) {
                double v = Hm[(size_t)a * m + b] + (a == b ? mu : 0);
                for (int c = 0; c < b; c++) v -= L[(size_t)a * m + c] * L[(size_t)b * m + c];
                if (a == b) {
                    if (!(v > 0)) {
                        pd = false;
                        break;
                    }
                    L[(size_t)a * m + a] = std::sqrt(v);
                } else {
                    L[(size_t)a * m + b] = v / L[(size_t)b * m + b];
                }
            }
    }
    if (!pd) return BSA_OK;
    for (int a = 0; a < m; a++) {
        double v = -gf[a];
        for (int c = 0; c < a; c++) v -= L[(size_t)a * m + c] * dd[c];
        dd[a] = v / L[(size_t)a * m + a];
    }
    for (int a = m - 1; a >= 0; a--) {
        double v = dd[a];
        for (int c = a + 1; c < m; c++) v -= L[(size_t)c * m + a] * dd[c];
        dd[a] = v / L[(size_t)a * m + a];
    }
    *ok = true;
    return BSA_OK;
}

// q = H v on the free variables fr (v, q: one entry per free variable) by a
// forward difference of the gradient along v, the step 1e-5 max(1, |x_i|)
// / |v|_inf (the largest |x_i| where v is nonzero) taken towards the inside
// of the box.  *valid = false where no step fits in the box or the
// evaluation fails (BSA_ENUMERIC).
template <class FG>
static int hess_dir(FG &fg, int n, const double *x, const double *lo, const double *hi,
                    const std::vector<int> &fr, const std::vector<double> &gf,
                    const std::vector<double> &v, std::vector<double> &q, bool *valid)
{
    const int m = (int)fr.size();
    double vm = 0, xs = 0, up = INFINITY, dn = INFINITY;
    for (int a = 0; a < m; a++) {
        const int i = fr[a];
        if (v[a] == 0) continue;
        vm = std::max(vm, std::fabs(v[a]));
        xs = std::max(xs, std::fabs(x[i]));
        if (v[a] > 0) {
            up = std::min(up, (hi[i] - x[i]) / v[a]);
            dn = std::min(dn, (x[i] - lo[i]) / v[a]);
        } else {
            up = std::min(up, (lo[i] - x[i]) / v[a]);
            dn = std::min(dn, (x[i] - hi[i]) / v[a]);
        }
    }
    *valid = false;
    if (vm == 0) return BSA_OK;
    const double h = 1e-5 * std::max(1.0, xs) / vm;
    const double st = h <= up ? h : h <= dn ? -h : up >= dn ? 0.5 * up : -0.5 * dn;
    if (!(st != 0)) return BSA_OK;
    std::vector<double> xt(x, x + n), gt(n);
    for (int a = 0; a < m; a++) {
        const int i = fr[a];
        if (v[a] != 0) xt[i] = clampd(x[i] + st * v[a], lo[i], hi[i]);
    }
    double ft;
    const int rc = fg(xt.data(), &ft, gt.data());
    if (rc) return rc == BSA_ENUMERIC ? BSA_OK : rc;
    for (int a = 0; a < m; a++) q[a] = (gt[fr[a]] - gf[a]) / st;
    *valid = true;
    return BSA_OK;
}

// Newton step d on the free variables fr by preconditioned conjugate
// gradients on H d = -g (truncated Newton; Nocedal & Wright, 2nd ed.,
// algorithms 5.3 and 7.1).  The free variables among the last nblock of x
// (block S: in codeml's x, the model parameters after the branch lengths)
// get their Hessian columns exactly, by one gradient difference each; the
// preconditioner is H_SS on S and a scalar on the rest (R), the Rayleigh
// quotient of H_RR along g_R.  On a likelihood surface whose flat
// directions (p2 w2 ~ const) lie in S, this resolves them in a few
// products, where plain conjugate gradients on a Hessian whose eigenvalues
// span nine decades barely move along them.  Each product H p takes one
// gradient difference along p's R part, H_.S p_S coming from the columns.
// Ends at (r'z / r0'z0)^(1/2) below kCgTol, after kCgMax products, or at a
// direction of non-positive curvature (with the step so far); *ok = false
// where no step is found.
template <class FG>
static int pcg_step(FG &fg, int n, const double *x, const double *lo, const double *hi,
                    const std::vector<int> &fr, int nblock, const std::vector<double> &gf,
                    std::vector<double> &dd, bool *ok)
{
    const int m = (int)fr.size();
    std::vector<int> S;
    for (int a = 0; a < m; a++)
        if (fr[a] >= n - nblock) S.push_back(a);
    const int ns = (int)S.size();
    std::vector<char> inS(m, 0);
    for (int a : S) inS[a] = 1;
    std::vector<double> HS((size_t)m * ns), L((size_t)ns * ns), v(m), q(m);
    *ok = false;
    bool valid;
    int rc;
    for (int c = 0; c < ns; c++) {   // columns of S, all free rows
        std::fill(v.begin(), v.end(), 0.0);
        v[S[c]] = 1;
        if ((rc = hess_dir(fg, n, x, lo, hi, fr, gf, v, q, &valid))) return rc;
        if (!valid) return BSA_OK;
        for (int a = 0; a < m; a++) HS[(size_t)a * ns + c] = q[a];
    }
    // Cholesky of H_SS + mu I, mu raised from 0 until it succeeds
    double dmax = 0;
    for (int c = 0; c < ns; c++) dmax = std::max(dmax, std::fabs(HS[(size_t)S[c] * ns + c]));
    bool pd = ns == 0;
    for (double mu = 0; !pd && mu < 1e6 * (dmax + 1); mu = mu ? 10 * mu : 1e-12 * (dmax + 1)) {
        pd = true;
        for (int a = 0; a < ns && pd; a++)
            for (int b = 0; b <= a; b++) {
                double w = 0.5 * (HS[(size_t)S[a] * ns + b] + HS[(size_t)S[b] * ns + a]) +
                           (a == b ? mu : 0);
                for (int c = 0; c < b; c++) w -= L[(size_t)a * ns + c] * L[(size_t)b * ns + c];
                if (a == b) {
                    if (!(w > 0)) {
                        pd = false;
                        break;
                    }
                    L[(size_t)a * ns + a] = std::sqrt(w);
                } else {
                    L[(size_t)a * ns + b] = w / L[(size_t)b * ns + b];
                }
            }
    }
    if (!pd) return BSA_OK;
    // H p: a gradient difference along p_R, plus the S columns times p_S
    auto hprod = [&](const std::vector<double> &p, std::vector<double> &hp, bool *hv) {
        std::fill(hp.begin(), hp.end(), 0.0);
        bool any = false;
        for (int a = 0; a < m; a++) {
            v[a] = inS[a] ? 0 : p[a];
            any = any || v[a] != 0;
        }
        *hv = true;
        if (any) {
            const int r2 = hess_dir(fg, n, x, lo, hi, fr, gf, v, hp, hv);
            if (r2 || !*hv) return r2;
        }
        for (int a = 0; a < m; a++)
            for (int c = 0; c < ns; c++) hp[a] += HS[(size_t)a * ns + c] * p[S[c]];
        return BSA_OK;
    };
    // scalar of R: Rayleigh quotient along -g_R (the product is also the
    // first direction's, scaled)
    std::vector<double> r(m), z(m), p(m), hq(m);
    for (int a = 0; a < m; a++) r[a] = -gf[a];
    double sig = 1, rr = 0;
    std::fill(hq.begin(), hq.end(), 0.0);
    for (int a = 0; a < m; a++)
        if (!inS[a]) rr += r[a] * r[a];
    if (rr > 0) {
        for (int a = 0; a < m; a++) v[a] = inS[a] ? 0 : r[a];
        if ((rc = hess_dir(fg, n, x, lo, hi, fr, gf, v, hq, &valid))) return rc;
        if (!valid) return BSA_OK;
        double vq = 0;
        for (int a = 0; a < m; a++)
            if (!inS[a]) vq += r[a] * hq[a];
        sig = vq / rr;
        if (!(sig > 0)) return BSA_OK;
    }
    auto precond = [&](const std::vector<double> &rv, std::vector<double> &zv) {
        for (int a = 0; a < m; a++) zv[a] = inS[a] ? 0 : rv[a] / sig;
        std::vector<double> t(ns);
        for (int a = 0; a < ns; a++) {
            double w = rv[S[a]];
            for (int c = 0; c < a; c++) w -= L[(size_t)a * ns + c] * t[c];
            t[a] = w / L[(size_t)a * ns + a];
        }
        for (int a = ns - 1; a >= 0; a--) {
            double w = t[a];
            for (int c = a + 1; c < ns; c++) w -= L[(size_t)c * ns + a] * t[c];
            t[a] = w / L[(size_t)a * ns + a];
        }
        for (int a = 0; a < ns; a++) zv[S[a]] = t[a];
    };
    precond(r, z);
    p = z;
    double rz = 0;
    for (int a = 0; a < m; a++) {
        rz += r[a] * z[a];
        dd[a] = 0;
    }
    const double rz0 = rz;
    for (int k = 0; k < kCgMax && rz > 0; k++) {
        if (k == 0) {
            // H p0 from the Rayleigh product: p0_R = -g_R / sig
            for (int a = 0; a < m; a++) {
                q[a] = hq[a] / sig;
                for (int c = 0; c < ns; c++) q[a] += HS[(size_t)a * ns + c] * p[S[c]];
            }
        } else {
            if ((rc = hprod(p, q, &valid))) return rc;
            if (!valid) break;
        }
        double pq = 0;
        for (int a = 0; a < m; a++) pq += p[a] * q[a];
        if (!(pq > 0)) break;
        const double al = rz / pq;
        for (int a = 0; a < m; a++) {
            dd[a] += al * p[a];
            r[a] -= al * q[a];
        }
        *ok = true;
        precond(r, z);
        double rzn = 0;
        for (int a = 0; a < m; a++) rzn += r[a] * z[a];
        if (!(rzn > kCgTol * kCgTol * rz0)) break;
        const double be = rzn / rz;
        for (int a = 0; a < m; a++) p[a] = z[a] + be * p[a];
        rz = rzn;
    }
    return BSA_OK;
}

// Newton iterations from x (f, g its value and gradient, updated in place);
// *converged = 1 when the last Newton step was below 1e-6 max(1, |x_i|)
template <class FG, class FV, class GR>
static int newton_finish(FG &fg, FV &fv, GR &gr, int n, double *x, const double *lo,
                         const double *hi, int iters, int nblock, double *f,
                         std::vector<double> &g, int *converged)
{
    std::vector<double> xt(n), gt(n);
    *converged = 0;
    for (int k = 0; k < iters; k++) {
        std::vector<int> fr;
        for (int i = 0; i < n; i++) {
            const double tol = 1e-12 * (1 + std::fabs(lo[i]) + std::fabs(hi[i]));
            const bool out = lo[i] == hi[i] || (x[i] <= lo[i] + tol && g[i] > 0) ||
                             (x[i] >= hi[i] - tol && g[i] < 0);
            if (!out) fr.push_back(i);
        }
        const int m = (int)fr.size();
        if (m == 0) {
            *converged = 1;
            return BSA_OK;
        }
        std::vector<double> gf(m), dd(m);
        for (int a = 0; a < m; a++) gf[a] = g[fr[a]];
        bool ok = false;
        const int rc = m > kDenseNewtonMax ? pcg_step(fg, n, x, lo, hi, fr, nblock, gf, dd, &ok)
                                           : dense_step(fg, n, x, lo, hi, fr, gf, dd, &ok);
        if (rc) return rc;
        if (!ok) return BSA_OK;
        double rel = 0;
        for (int a = 0; a < m; a++)
            rel = std::max(rel, std::fabs(dd[a]) / std::max(1.0, std::fabs(x[fr[a]])));
        if (rel < 1e-10) {   // at the maximum to rounding
            *converged = 1;
            return BSA_OK;
        }
        // projected Armijo line search along the Newton step
        double step = 1.0, ft = 0;
        bool acc = false;
        for (int ls = 0; ls < 30; ls++) {
            double slope = 0;
            std::memcpy(xt.data(), x, sizeof(double) * n);
            for (int a = 0; a < m; a++) {
                const int i = fr[a];
                xt[i] = clampd(x[i] + step * dd[a], lo[i], hi[i]);
                slope += g[i] * (xt[i] - x[i]);
            }
            int rc = fv(xt.data(), &ft);
            if (rc && rc != BSA_ENUMERIC) return rc;
            if (rc == BSA_OK && std::isfinite(ft) && ft <= *f + 1e-4 * slope) {
                if ((rc = gr(gt.data()))) return rc;   // the accepted point's gradient
                acc = true;
                break;
            }
            step *= 0.5;
        }
        if (!acc) {
            // no decrease: at the maximum where the step is below rounding
            *converged = rel < 1e-6;
            return BSA_OK;
        }
        std::memcpy(x, xt.data(), sizeof(double) * n);
        g = gt;
        *f = ft;
        *converged = rel * step < 1e-6;
        if (rel * step < 1e-8) return BSA_OK;
    }
    return BSA_OK;
}

// Newton iterations that check a stall or no-descent exit: a BFGS crawling
// along a curved, nearly flat ridge (p2 w2 ~ const) near the maximum can end
// by the stall rule a few 1e-6 of lnL short of it; from there the first two
// Newton iterations close the gap and the later ones show the step vanish
constexpr int kVerifyNewton = 4;

int minimize_box(const Objective &obj, int n, double *x, const double *lo, const double *hi,
                 const bsa_fit_opts *opts, FitResult *res, bool verify, int nblock)
{
    const int max_iter = !opts || opts->max_iter == 0 ? 2000 : std::max(0, opts->max_iter);
    const int newton = opts ? std::max(0, opts->newton) : 0;
    const double e = opts && opts->e > 0 ? opts->e : 1e-8;
    const double max_step = opts && opts->max_step > 0 ? opts->max_step : 5.0;
    int nfev = 0;
    auto fv = [&](const double *xx, double *ff) {
        nfev++;
        return obj.value(xx, ff);
    };
    auto gr = [&](double *gg) { return obj.grad(gg); };
    auto fg = [&](const double *xx, double *ff, double *gg) {
        if (obj.both) {
            nfev++;
            return obj.both(xx, ff, gg);
        }
        const int rc = fv(xx, ff);
        return rc ? rc : gr(gg);
    };
    std::vector<double> g(n), xt(n), gt(n), d(n), s(n), y(n), Hy(n), H((size_t)n * n, 0.0);
    std::vector<char> act(n);
    for (int i = 0; i < n; i++) x[i] = clampd(x[i], lo[i], hi[i]);
    double f;
    int rc = fg(x, &f, g.data());
    if (rc) return rc;
    auto reset_h = [&](double gamma) {
        std::fill(H.begin(), H.end(), 0.0);
        for (int i = 0; i < n; i++) H[(size_t)i * n + i] = gamma;
    };
    reset_h(1.0);
    bool scaled = false, fresh = true;
    double gamma = 1.0;
    int it = 0, good = 0, stop = BSA_STOP_LIMIT;
    bool confirmed = false;
    // f after the last kStall + 1 accepted steps (ring buffer)
    constexpr int kStall = 50;
    std::vector<double> fring(kStall + 1);
    long nacc = 0;
    fring[0] = f;
    for (; it < max_iter; it++) {
        // working set and projected-gradient test
        double pg = 0;
        for (int i = 0; i < n; i++) {
            const double tol = 1e-12 * (1 + std::fabs(lo[i]) + std::fabs(hi[i]));
            act[i] = lo[i] == hi[i] || (x[i] <= lo[i] + tol && g[i] > 0) ||
                     (x[i] >= hi[i] - tol && g[i] < 0);
            pg = std::max(pg, std::fabs(clampd(x[i] - g[i], lo[i], hi[i]) - x[i]));
        }
        const double gtol = 1e-8 + 1e-10 * std::max(1.0, std::fabs(f));
        if (pg < gtol) {
            stop = BSA_STOP_PGRAD;
            break;
        }
        double gd = 0, dmax = 0;
        for (int i = 0; i < n; i++) {
            d[i] = 0;
            if (act[i]) continue;
            for (int j = 0; j < n; j++)
                if (!act[j]) d[i] -= H[(size_t)i * n + j] * g[j];
            gd += g[i] * d[i];
        }
        if (!(gd < 0)) {   // not a descent direction: restart from scaled identity
            reset_h(gamma);
            fresh = true;
            gd = 0;
            for (int i = 0; i < n; i++) {
                d[i] = act[i] ? 0 : -gamma * g[i];
                gd += g[i] * d[i];
            }
        }
        for (int i = 0; i < n; i++) dmax = std::max(dmax, std::fabs(d[i]));
        if (!(dmax > 0)) {
            stop = BSA_STOP_NODESCENT;
            break;
        }
        double a = 1.0;
        if (!scaled) a = std::min(1.0, 0.1 / dmax);   // first step: move at most 0.1
        a = std::min(a, max_step / dmax);
        // projected backtracking line search
        double ft = 0;
        bool ok = false;
        for (int ls = 0; ls < 60; ls++) {
            double slope = 0, dx = 0;
            for (int i = 0; i < n; i++) {
                xt[i] = clampd(x[i] + a * d[i], lo[i], hi[i]);
                slope += g[i] * (xt[i] - x[i]);
                dx = std::max(dx, std::fabs(xt[i] - x[i]));
            }
            if (dx == 0) break;
            rc = fv(xt.data(), &ft);
            if (rc && rc != BSA_ENUMERIC) return rc;
            if (rc == BSA_OK && std::isfinite(ft) && ft <= f + 1e-4 * slope) {
                if ((rc = gr(gt.data()))) return rc;   // the accepted point's gradient
                ok = true;
                break;
            }
            double an = 0.5 * a;
            if (rc == BSA_OK && std::isfinite(ft)) {
                // minimiser of the quadratic through f, slope and ft along the path
                const double sl = slope / a, den = 2 * (ft - f - sl * a);
                if (den > 0) an = -sl * a * a / den;
            }
            a = clampd(an, 0.1 * a, 0.5 * a);
        }
        if (!ok) {
            if (fresh) {   // no progress even along the scaled steepest descent
                stop = BSA_STOP_NODESCENT;
                break;
            }
            reset_h(gamma);
            fresh = true;
            continue;
        }
        // curvature pair on the free variables
        double sy = 0, yy = 0, s2 = 0;
        for (int i = 0; i < n; i++) {
            s[i] = xt[i] - x[i];
            y[i] = act[i] ? 0 : gt[i] - g[i];
            if (act[i]) s[i] = 0;
            sy += s[i] * y[i];
            yy += y[i] * y[i];
            s2 += s[i] * s[i];
        }
        if (sy > 1e-12 * std::sqrt(yy * s2) && yy > 0) {
            if (!scaled) {
                gamma = sy / yy;
                reset_h(gamma);
                scaled = true;
            }
            double yHy = 0;
            for (int i = 0; i < n; i++) {
                double v = 0;
                for (int j = 0; j < n; j++) v += H[(size_t)i * n + j] * y[j];
                Hy[i] = v;
                yHy += y[i] * v;
            }
            const double c1 = (sy + yHy) / (sy * sy);
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++)
                    H[(size_t)i * n + j] += c1 * s[i] * s[j] - (Hy[i] * s[j] + s[i] * Hy[j]) / sy;
            fresh = false;
        }
        // two successive steps that change f by less than e |f| and x by
        // less than e |x| (2-norms, as H_end) end the fit, after one restart
        // of H from the scaled identity has confirmed it
        double sn = 0, xn = 0;
        for (int i = 0; i < n; i++) {
            sn += (xt[i] - x[i]) * (xt[i] - x[i]);
            xn += x[i] * x[i];
        }
        const bool small_step = std::fabs(f - ft) < e * std::max(1.0, std::fabs(f)) &&
                                std::sqrt(sn) < e * std::max(1.0, std::sqrt(xn));
        std::memcpy(x, xt.data(), sizeof(double) * n);
        std::swap(g, gt);
        f = ft;
        nacc++;
        fring[nacc % (kStall + 1)] = f;
        if (nacc >= kStall &&
            fring[(nacc - kStall) % (kStall + 1)] - f < 1e-3 * e * std::max(1.0, std::fabs(f))) {
            stop = BSA_STOP_STALL;
            it++;
            break;
        }
        good = small_step ? good + 1 : 0;
        if (good >= 2) {
            if (confirmed) {
                stop = BSA_STOP_HEND;
                it++;
                break;
            }
            confirmed = true;
            good = 0;
            reset_h(gamma);
            fresh = true;
        }
    }
    if (newton > 0) {
        int st = 0;
        if ((rc = newton_finish(fg, fv, gr, n, x, lo, hi, newton, nblock, &f, g, &st))) return rc;
        if (st) stop = BSA_STOP_NEWTON;
        it += newton;
    } else if (verify && (stop == BSA_STOP_STALL || stop == BSA_STOP_NODESCENT)) {
        int st = 0;
        if ((rc = newton_finish(fg, fv, gr, n, x, lo, hi, kVerifyNewton, nblock, &f, g, &st))) return rc;
        if (st) stop = BSA_STOP_NEWTON;
    }
    res->f = f;
    res->iterations = it;
    res->evaluations = nfev;
    res->stop = stop;
    return BSA_OK;
}

}  // namespace bsa

using bsa::FitResult;
using bsa::ObjFn;
using bsa::Objective;
using bsa::minimize_box;

extern "C" int bsa_fit(bsa_engine *eng, double *x, int np, const double *lo, const double *hi,
                       const bsa_fit_opts *opts, double *neg_lnl, int *info)
{
    if (!eng || !x || !lo || !hi || !neg_lnl) return bsa_error(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(eng)) return bsa_error(BSA_EINVAL, "np does not match the engine");
    for (int i = 0; i < np; i++)
        if (!(lo[i] <= hi[i])) return bsa_error(BSA_EINVAL, "bounds with lo > hi");
    const bool fd = opts && opts->gradient == BSA_GRAD_FD;
    // the gradient only where the optimiser accepts a point: a rejected
    // trial costs one likelihood (the forward pass of the adjoint's
    // evaluation, or the one likelihood of the differences), not a gradient
    Objective obj;
    std::vector<double> xl(np);
    if (fd) {
        obj.value = [&](const double *xx, double *f) {
            std::copy(xx, xx + np, xl.begin());
            return bsa_lnl(eng, xx, np, f);
        };
        obj.grad = [&](double *g) {
            double f;
            return bsa_lnl_grad_fd(eng, xl.data(), np, 0, 1e-6, &f, g);
        };
    } else if (bsa_engine_graph(eng)) {
        // graph replay evaluates lnL and gradient together
        obj = bsa::combined(
            [&](const double *xx, double *f, double *g) { return bsa_lnl_grad(eng, xx, np, f, g); },
            np);
    } else {
        obj.value = [&](const double *xx, double *f) {
            return bsa_engine_lnl_part1(eng, xx, np, f);
        };
        obj.grad = [&](double *g) { return bsa_engine_grad_part2(eng, np, g); };
    }
    FitResult r;
    int rc = minimize_box(obj, np, x, lo, hi, opts, &r,
                          !fd && bsa_engine_precision(eng) == BSA_FP64,
                          np - bsa_engine_nbranch(eng));
    if (rc) return rc;
    // leave the engine's state (per-site values, parameters) at the returned x
    double f;
    if ((rc = bsa_lnl(eng, x, np, &f))) return rc;
    *neg_lnl = f;
    if (info) {
        info[0] = r.iterations;
        info[1] = r.evaluations + 1;
        info[2] = r.stop;
    }
    return BSA_OK;
}

extern "C" int bsa_debug_minimize(int n, bsa_objective fn, void *ctx, double *x, const double *lo,
                                  const double *hi, const bsa_fit_opts *opts, int nblock,
                                  double *fmin, int *info)
{
    if (n <= 0 || !fn || !x || !lo || !hi || !fmin || nblock < 0 || nblock > n)
        return bsa_error(BSA_EINVAL, "bad argument");
    for (int i = 0; i < n; i++)
        if (!(lo[i] <= hi[i])) return bsa_error(BSA_EINVAL, "bounds with lo > hi");
    ObjFn fg = [&](const double *xx, double *f, double *g) {
        return fn(xx, f, g, ctx) ? BSA_ENUMERIC : BSA_OK;
    };
    FitResult r;
    int rc = minimize_box(fg, n, x, lo, hi, opts, &r, false, nblock);
    if (rc) return rc;
    *fmin = r.f;
    if (info) {
        info[0] = r.iterations;
        info[1] = r.evaluations;
        info[2] = r.stop;
    }
    return BSA_OK;
}
