// C ABI of the single-gene branch-site model A engine: problem set-up, one
// likelihood evaluation per call, adjoint and finite-difference gradients,
// per-site outputs, Bayes empirical Bayes.  The engine is a forest of one
// gene (bsa_forest.h).  Declarations and the reference entry points each
// call replaces are in include/bsa.h.
#include "bsa_forest.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace bsa;

struct bsa_engine {
    Gene g;
    Forest f;
    std::vector<double> part_x;   // x of the last part-1 evaluation
    bool part_ready = false;      // its forward state is on the device
};

int bsa_engine_precision(const bsa_engine *eng) { return eng->f.precision; }
bool bsa_engine_graph(const bsa_engine *eng) { return eng->f.use_graph; }
int bsa_engine_nbranch(const bsa_engine *eng) { return eng->g.nbranch; }

extern "C" int bsa_create(const bsa_problem *p, bsa_engine **out)
{
    if (!p || !out) return fail(BSA_EINVAL, "null argument");
    *out = nullptr;
    HIP_SHARED;
    if (p->precision != BSA_FP64 && p->precision != BSA_FP32)
        return fail(BSA_EINVAL, "unknown precision %d", p->precision);
    bsa_engine *e = new bsa_engine();
    int rc = gene_init(e->g, p, p->precision == BSA_FP32 ? 4 : 8);
    if (rc) { delete e; return rc; }
    rc = forest_build(e->f, {&e->g}, p->device, p->precision, p->kernel, p->tile);
    if (rc) {
        forest_free(e->f);   // sets no error text
        delete e;
        return rc;
    }
    *out = e;
    return BSA_OK;
}

extern "C" int bsa_destroy(bsa_engine *e)
{
    if (!e) return BSA_OK;
    HIP_SHARED;
    forest_free(e->f);
    delete e;
    return BSA_OK;
}

extern "C" int bsa_set_model(bsa_engine *e, const bsa_model_opts *o)
{
    if (!e || !o) return fail(BSA_EINVAL, "null argument");
    if (o->fix_omega && !(o->omega_fix > 0)) return fail(BSA_EINVAL, "omega_fix must be > 0");
    e->g.opts = *o;
    e->part_ready = false;   // x means something else now: no sweep of a kept forward pass
    return BSA_OK;
}

extern "C" int bsa_num_params(const bsa_engine *e)
{
    if (!e) return -1;
    return gene_num_params(e->g);
}

// ---------------------------------------------------------------- evaluation

// One evaluation at natural parameters; branch lengths from x.  qf_fixed,
// when given, replaces the two Qfactors the model would give (BEB keeps them
// at the MLE, GetPMatBranch treesub.c:7289-7301).
static int evaluate_natural(bsa_engine *e, const double *x, double kappa, double p0, double p1,
                            double w0, double w2, const double *qf_fixed, double *lnl,
                            bool adjoint)
{
    e->part_ready = false;   // the device buffers no longer hold a part-1 state
    int rc = gene_stage(e->g, e->f, x, kappa, p0, p1, w0, w2, qf_fixed, adjoint);
    if (rc) return rc;
    if ((rc = forest_run(e->f, adjoint))) return rc;
    *lnl = gene_collect(e->g, e->f);
    return BSA_OK;
}

static int evaluate(bsa_engine *e, const double *x, double *lnl, bool adjoint = false)
{
    auto t0 = std::chrono::steady_clock::now();
    double kappa, p0, p1, w0, w2;
    natural_params(e->g, x, &kappa, &p0, &p1, &w0, &w2);
    const int rc = evaluate_natural(e, x, kappa, p0, p1, w0, w2, nullptr, lnl, adjoint);
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc == BSA_OK) {
        e->f.timing[6] = ms;
        e->f.timing_sum[6] += ms;
        e->f.n_eval += 1;
    }
    return rc;
}

extern "C" int bsa_lnl(bsa_engine *e, const double *x, int np, double *neg_lnl)
{
    if (!e || !x || !neg_lnl) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    const int rg = forest_prepare_graph(e->f, false);
    if (rg) return rg;
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->f.dev));
    double lnl;
    int rc = evaluate(e, x, &lnl);
    if (rc) return rc;
    *neg_lnl = -lnl;
    return BSA_OK;
}

// An evaluation with gradient in two calls, for an optimiser that needs the
// gradient only at the points it accepts: part 1 returns -lnL at x and keeps
// the forward state the adjoint sweep needs; part 2 returns d(-lnL)/dx at
// that x.  Together they give bsa_lnl_grad's values bitwise; a rejected
// trial point costs the forward pass only.
int bsa_engine_lnl_part1(bsa_engine *e, const double *x, int np, double *neg_lnl)
{
    if (!e || !x || !neg_lnl) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->f.dev));
    e->part_ready = false;
    int rc = forest_setup_adjoint(e->f);
    if (rc) return rc;
    auto t0 = std::chrono::steady_clock::now();
    double kappa, p0, p1, w0, w2;
    natural_params(e->g, x, &kappa, &p0, &p1, &w0, &w2);
    if ((rc = gene_stage(e->g, e->f, x, kappa, p0, p1, w0, w2, nullptr, true))) return rc;
    if ((rc = forest_run_part(e->f, 1))) return rc;
    const double lnl = gene_collect(e->g, e->f);
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    e->f.timing[6] = ms;
    e->f.timing_sum[6] += ms;
    e->f.n_eval += 1;
    e->part_x.assign(x, x + np);
    e->part_ready = true;
    *neg_lnl = -lnl;
    return BSA_OK;
}

int bsa_engine_grad_part2(bsa_engine *e, int np, double *grad)
{
    if (!e || !grad) return fail(BSA_EINVAL, "null argument");
    if (!e->part_ready || (int)e->part_x.size() != np)
        return fail(BSA_EINVAL, "no part-1 evaluation to differentiate");
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->f.dev));
    const int rc = forest_run_part(e->f, 2);
    if (rc) return rc;
    gene_chain_rule(e->g, e->f, e->part_x.data(), grad);
    return BSA_OK;
}

// -lnL and d(-lnL)/dx by one forward pass and one adjoint (preorder) sweep
extern "C" int bsa_lnl_grad(bsa_engine *e, const double *x, int np, double *neg_lnl, double *grad)
{
    if (!e || !x || !neg_lnl || !grad) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    const int rg = forest_prepare_graph(e->f, true);
    if (rg) return rg;
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->f.dev));
    int rc = forest_setup_adjoint(e->f);
    if (rc) return rc;
    double lnl;
    if ((rc = evaluate(e, x, &lnl, true))) return rc;
    gene_chain_rule(e->g, e->f, x, grad);
    *neg_lnl = -lnl;
    return BSA_OK;
}

// finite-difference gradient of -lnL in codeml's variables, as gradientB
// (tools.c:6326-6350): step eh = small*(|x_i|+1); forward difference, or
// central with eh^0.67 when `central`.
extern "C" int bsa_lnl_grad_fd(bsa_engine *e, const double *x, int np, int central, double small,
                               double *neg_lnl, double *grad)
{
    if (!e || !x || !neg_lnl || !grad) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    if (!(small > 0)) small = 1e-6;
    const int rg = forest_prepare_graph(e->f, false);
    if (rg) return rg;
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->f.dev));
    std::vector<double> x1(x, x + np);
    double f0, f1, f2;
    int rc = evaluate(e, x, &f0);
    if (rc) return rc;
    for (int i = 0; i < np; i++) {
        double eh = small * (std::fabs(x[i]) + 1);
        if (central) {
            eh = std::pow(eh, 0.67);
            x1[i] = x[i] + eh;
            if ((rc = evaluate(e, x1.data(), &f1))) return rc;
            x1[i] = x[i] - eh;
            if ((rc = evaluate(e, x1.data(), &f2))) return rc;
            grad[i] = (-f1 + f2) / (2 * eh);
        } else {
            x1[i] = x[i] + eh;
            if ((rc = evaluate(e, x1.data(), &f1))) return rc;
            grad[i] = (-f1 + f0) / eh;
        }
        x1[i] = x[i];
    }
    // leave the engine's per-site state at x
    if ((rc = evaluate(e, x, &f0))) return rc;
    *neg_lnl = -f0;
    return BSA_OK;
}

// Bayes empirical Bayes for model A (lfunNSsites_ACD, codeml.c:6617-6826;
// get_grid_para_like_ACD 6419-6607; get_pclassM_iw_ACD 6609-6615 with
// GetIndexTernary, tools.c:4441-4461).  Branch lengths, kappa and the two
// Qfactors are fixed at x; the 121 w sets (w0 on 10 points in (0, 1), w2 on
// 10 points in (1, 11), w = 1) come from 100 forward passes with
// (w0, w2) on the 10 x 10 grid (class 2a gives (w0, w2), class 0 (w0, w0),
// class 2b (1, w2), class 1 (1, 1)); the 10^4-point grid over
// (p0-p1 triangle, w0, w2) is integrated on the GPU.
namespace {
// device buffer released when it goes out of scope (release errors cannot
// be reported from a destructor and are ignored)
struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

extern "C" int bsa_beb(bsa_engine *e, const double *x, int np, double *post, double *ml3,
                       double *postw20, double *postp0p1, double *lnf_sets)
{
    if (!e || !x || !post) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    const int rg = forest_prepare_graph(e->f, false);
    if (rg) return rg;
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->f.dev));
    constexpr int n1d = 10, dim = 4, ngrid = 10000, nset = n1d + 1 + n1d * n1d + n1d;
    const int npatt = e->g.npatt, npad = e->f.npad;
    double kappa, p0, p1, w0, w2;
    natural_params(e->g, x, &kappa, &p0, &p1, &w0, &w2);
    ModelA mle;
    if (!setup_model_a(kappa, p0, p1, w0, w2, e->g.pi.data(), &mle, false))
        return fail(BSA_EINVAL, "parameters out of range for BEB");
    const double qf[2] = {mle.Qfactor[0], mle.Qfactor[1]};
    double w0g[n1d], w2g[n1d];
    for (int i = 0; i < n1d; i++) {
        w0g[i] = (i + 0.5) / n1d;
        w2g[i] = 1.0 + (i + 0.5) * (11.0 - 1.0) / n1d;
    }
    DevBuf F, mx, pc, iw, lnfx, gl, W, dpost;
    HIPCHK(hipMalloc(&F.p, sizeof(double) * nset * (size_t)npad));
    hipStream_t s = e->f.stream;
    double *dF = (double *)F.p;
    auto row = [&](int set) { return dF + (size_t)set * npad; };
    for (int i = 0; i < n1d; i++)
        for (int j = 0; j < n1d; j++) {
            double l;
            int rc = evaluate_natural(e, x, kappa, p0, p1, w0g[i], w2g[j], qf, &l, false);
            if (rc) return rc;
            const size_t rb = sizeof(double) * npad;
            const double *fh = e->f.d_fh;
            HIPCHK(hipMemcpyAsync(row(n1d + 1 + i * n1d + j), fh + 2 * (size_t)npad, rb,
                                  hipMemcpyDeviceToDevice, s));
            if (j == 0)
                HIPCHK(hipMemcpyAsync(row(i), fh, rb, hipMemcpyDeviceToDevice, s));
            if (i == 0)
                HIPCHK(hipMemcpyAsync(row(n1d + 1 + n1d * n1d + j), fh + 3 * (size_t)npad, rb,
                                      hipMemcpyDeviceToDevice, s));
            if (i == 0 && j == 0)
                HIPCHK(hipMemcpyAsync(row(n1d), fh + (size_t)npad, rb, hipMemcpyDeviceToDevice, s));
        }
    if (lnf_sets) {
        HIPCHK(hipStreamSynchronize(s));
        for (int k = 0; k < nset; k++)
            HIPCHK(hipMemcpy(lnf_sets + (size_t)k * npatt, row(k), sizeof(double) * npatt,
                             hipMemcpyDeviceToHost));
    }
    // normalise per pattern; S1 = sum_h fpatt_h max_k log f
    HIPCHK(hipMalloc(&mx.p, sizeof(double) * npad));
    launch_beb_norm(s, dF, nset, npatt, npad, (double *)mx.p);
    HIPCHK(hipGetLastError());
    std::vector<double> hmx(npatt);
    HIPCHK(hipMemcpyAsync(hmx.data(), mx.p, sizeof(double) * npatt, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double S1 = 0, S1c = 0;
    for (int h = 0; h < npatt; h++) {
        const double v = e->g.fpatt[h] * hmx[h], t = S1 + v;
        S1c += std::fabs(S1) >= std::fabs(v) ? (S1 - t) + v : (v - t) + S1;
        S1 = t;
    }
    S1 += S1c;
    // grid: ip = (p0-p1 triangle row, column, w0 index, w2 index)
    std::vector<double> hpc(4 * ngrid);
    std::vector<int> hiw(4 * ngrid);
    for (int g = 0; g < ngrid; g++) {
        const int ip0 = g / 1000, ip1 = (g / 100) % 10, ip2 = (g / 10) % 10, ip3 = g % 10;
        const int tri = ip0 * n1d + ip1;
        const int ix = (int)std::sqrt((double)tri), iy = tri - ix * ix;
        const double q0 = (1 + (iy / 2) * 3 + (iy % 2)) / (3.0 * n1d);
        const double q1 = (1 + (n1d - 1 - ix) * 3 + (iy % 2)) / (3.0 * n1d);
        const double q2 = 1 - q0 - q1;
        hpc[g * 4 + 0] = q0;
        hpc[g * 4 + 1] = q1;
        hpc[g * 4 + 2] = q2 * q0 / (1 - q2);
        hpc[g * 4 + 3] = q2 * q1 / (1 - q2);
        hiw[g * 4 + 0] = ip2;
        hiw[g * 4 + 1] = n1d;
        hiw[g * 4 + 2] = n1d + 1 + ip2 * n1d + ip3;
        hiw[g * 4 + 3] = n1d + 1 + n1d * n1d + ip3;
    }
    HIPCHK(hipMalloc(&pc.p, sizeof(double) * 4 * ngrid));
    HIPCHK(hipMalloc(&iw.p, sizeof(int) * 4 * ngrid));
    HIPCHK(hipMalloc(&lnfx.p, sizeof(double) * ngrid));
    HIPCHK(hipMemcpyAsync(pc.p, hpc.data(), sizeof(double) * 4 * ngrid, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(iw.p, hiw.data(), sizeof(int) * 4 * ngrid, hipMemcpyHostToDevice, s));
    launch_beb_lnfx(s, dF, npad, npatt, e->f.d_fpatt, (const double *)pc.p, (const int *)iw.p, ngrid,
                    (double *)lnfx.p);
    HIPCHK(hipGetLastError());
    std::vector<double> hl(ngrid);
    HIPCHK(hipMemcpyAsync(hl.data(), lnfx.p, sizeof(double) * ngrid, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // marginal likelihood and posterior on the grid, accumulated in the
    // reference's order with its rescaling (codeml.c:6724-6750)
    double fX = 1, S2 = -1e300, pp[dim][n1d] = {}, p01[n1d * n1d] = {};
    for (int g = 0; g < ngrid; g++) {
        const int ip[4] = {g / 1000, (g / 100) % 10, (g / 10) % 10, g % 10};
        double t = hl[g] - S2;
        if (t > 10) {
            t = t < 200 ? std::exp(-t) : 0;
            fX = fX * t + 1;
            for (int j = 0; j < dim; j++)
                for (int k = 0; k < n1d; k++) pp[j][k] *= t;
            for (int k = 0; k < n1d * n1d; k++) p01[k] *= t;
            for (int j = 0; j < dim; j++) pp[j][ip[j]] += 1;
            p01[ip[0] * n1d + ip[1]] += 1;
            S2 = hl[g];
        } else if (t > -200) {
            t = std::exp(t);
            fX += t;
            for (int j = 0; j < dim; j++) pp[j][ip[j]] += t;
            p01[ip[0] * n1d + ip[1]] += t;
        }
    }
    for (int j = 0; j < dim; j++)
        for (int k = 0; k < n1d; k++) pp[j][k] /= fX;
    for (int k = 0; k < n1d * n1d; k++) p01[k] /= fX;
    const double lfX = std::log(fX) + S2;
    // grid points whose weight exp(lnfX - log f(X)) is representable
    std::vector<int> hg;
    std::vector<double> hw;
    for (int g = 0; g < ngrid; g++) {
        const double L = hl[g] - lfX;
        if (L > -745) {
            hg.push_back(g);
            hw.push_back(std::exp(L));
        }
    }
    const int ng = (int)hg.size();
    HIPCHK(hipMalloc(&gl.p, sizeof(int) * std::max(ng, 1)));
    HIPCHK(hipMalloc(&W.p, sizeof(double) * std::max(ng, 1)));
    HIPCHK(hipMalloc(&dpost.p, sizeof(double) * 4 * (size_t)npatt));
    HIPCHK(hipMemcpyAsync(gl.p, hg.data(), sizeof(int) * ng, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(W.p, hw.data(), sizeof(double) * ng, hipMemcpyHostToDevice, s));
    launch_beb_post(s, dF, npad, npatt, (const double *)pc.p, (const int *)iw.p,
                    (const int *)gl.p, (const double *)W.p, ng, (double *)dpost.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(post, dpost.p, sizeof(double) * 4 * (size_t)npatt, hipMemcpyDeviceToHost,
                          s));
    HIPCHK(hipStreamSynchronize(s));
    if (ml3) {
        ml3[0] = lfX + S1 - dim * std::log((double)n1d);   // log f(X) as the reference prints it
        ml3[1] = S1;
        ml3[2] = lfX;
    }
    if (postw20)
        for (int k = 0; k < n1d; k++) {
            postw20[k] = pp[2][k];
            postw20[n1d + k] = pp[3][k];
        }
    if (postp0p1)
        for (int k = 0; k < n1d * n1d; k++) postp0p1[k] = p01[k];
    // leave the engine's last evaluation at x
    double l;
    return evaluate(e, x, &l);
}

extern "C" int bsa_class_lnl(bsa_engine *e, double *fhK, double *site)
{
    if (!e) return fail(BSA_EINVAL, "null engine");
    if (!e->g.evaluated) return fail(BSA_EINVAL, "no evaluation yet");
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->f.dev));
    const int npatt = e->g.npatt, npad = e->f.npad;
    if (fhK)
        for (int c = 0; c < 4; c++)
            HIPCHK(hipMemcpy(fhK + (size_t)c * npatt, e->f.d_fh + (size_t)c * npad,
                             sizeof(double) * npatt, hipMemcpyDeviceToHost));
    if (site)
        HIPCHK(hipMemcpy(site, e->f.d_site, sizeof(double) * npatt, hipMemcpyDeviceToHost));
    return BSA_OK;
}

extern "C" int bsa_last_params(const bsa_engine *e, double *o)
{
    if (!e || !o) return fail(BSA_EINVAL, "null argument");
    if (!e->g.evaluated) return fail(BSA_EINVAL, "no evaluation yet");
    const ModelA &m = e->g.model;
    double v[11] = {m.kappa, m.p0, m.p1, m.w0, m.w2, m.Qfactor[0], m.Qfactor[1],
                    m.freqK[0], m.freqK[1], m.freqK[2], m.freqK[3]};
    std::memcpy(o, v, sizeof v);
    return BSA_OK;
}

extern "C" int bsa_timing(bsa_engine *e, double *o, int reset)
{
    if (!e || !o) return fail(BSA_EINVAL, "null argument");
    Forest &f = e->f;
    std::memcpy(o, f.timing, sizeof f.timing);
    std::memcpy(o + 7, f.timing_sum, sizeof f.timing_sum);
    o[14] = f.n_eval;
    o[15] = f.n_launch;
    o[16] = f.n_adj_launch;
    if (reset) {
        for (double &v : f.timing_sum) v = 0;
        f.n_eval = 0;
        f.n_launch = 0;
        f.n_adj_launch = 0;
    }
    return BSA_OK;
}

extern "C" int bsa_model_a_setup(double kappa, double p0, double p1, double w0, double w2,
                                 const double *pi, double *freqK, double *Qfactor, double *Root,
                                 double *U, double *V)
{
    if (!pi || !freqK || !Qfactor) return fail(BSA_EINVAL, "null argument");
    static thread_local ModelA m;
    bool eig = Root && U && V;
    if (!setup_model_a(kappa, p0, p1, w0, w2, pi, &m, eig))
        return fail(BSA_ENUMERIC, "model set-up failed");
    std::memcpy(freqK, m.freqK, sizeof m.freqK);
    std::memcpy(Qfactor, m.Qfactor, sizeof m.Qfactor);
    if (eig) {
        std::memcpy(Root, m.root, sizeof m.root);
        std::memcpy(U, m.U, sizeof m.U);
        std::memcpy(V, m.V, sizeof m.V);
    }
    return BSA_OK;
}

extern "C" int bsa_work(bsa_engine *e, double *o)
{
    if (!e || !o) return fail(BSA_EINVAL, "null argument");
    HIP_SHARED;
    int rc = forest_setup_adjoint(e->f);
    if (rc) return rc;
    std::memcpy(o, e->f.work, sizeof e->f.work);
    return BSA_OK;
}

extern "C" int bsa_set_graph(bsa_engine *e, int enable)
{
    if (!e) return fail(BSA_EINVAL, "null engine");
    e->f.use_graph = enable != 0;
    return BSA_OK;
}

extern "C" int bsa_set_shapes(bsa_engine *e, int tile, int tpb, int tpc)
{
    if (!e) return fail(BSA_EINVAL, "null engine");
    e->part_ready = false;
    return forest_set_shapes(e->f, tile, tpb, tpc);
}

extern "C" int bsa_set_prune(bsa_engine *e, int mode)
{
    if (!e) return fail(BSA_EINVAL, "null engine");
    e->part_ready = false;
    return forest_set_prune(e->f, mode);
}

extern "C" int bsa_tune(bsa_engine *e, const double *x, int np, int reps, double *out7)
{
    if (!e || !x) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    std::vector<double> g(np);
    double f;
    return forest_tune(e->f, [&]() { return bsa_lnl_grad(e, x, np, &f, g.data()); }, reps, out7);
}

extern "C" int bsa_default_bounds(const bsa_engine *e, double *lo, double *hi)
{
    if (!e || !lo || !hi) return fail(BSA_EINVAL, "null argument");
    default_bounds(e->g, lo, hi);
    return BSA_OK;
}

int bsa_error(int code, const char *msg) { return fail(code, "%s", msg); }

extern "C" const char *bsa_last_error(void) { return last_error_text(); }
extern "C" const char *bsa_version(void) { return "bsa 0.2 gfx950"; }

// In-bounds check of the matrix-core lane layout: runs one 16x16x4 MFMA in
// FP64 and FP32 on A (16x4) and B (4x16) and returns each lane's 4 outputs.
extern "C" int bsa_debug_mfma_probe(int device, const double *A, const double *B, double *D64,
                                    double *D32)
{
    HIP_SHARED;
    HIPCHK(hipSetDevice(device));
    DevBuf bA, bB, bD, cA, cB, cD;
    HIPCHK(hipMalloc(&bA.p, 64 * 8));
    HIPCHK(hipMalloc(&bB.p, 64 * 8));
    HIPCHK(hipMalloc(&bD.p, 256 * 8));
    HIPCHK(hipMalloc(&cA.p, 64 * 4));
    HIPCHK(hipMalloc(&cB.p, 64 * 4));
    HIPCHK(hipMalloc(&cD.p, 256 * 4));
    float hA[64], hB[64], hD[256];
    for (int i = 0; i < 64; i++) { hA[i] = (float)A[i]; hB[i] = (float)B[i]; }
    HIPCHK(hipMemcpy(bA.p, A, 64 * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(bB.p, B, 64 * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(cA.p, hA, 64 * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(cB.p, hB, 64 * 4, hipMemcpyHostToDevice));
    launch_mfma_probe(nullptr, (const double *)bA.p, (const double *)bB.p, (double *)bD.p,
                      (const float *)cA.p, (const float *)cB.p, (float *)cD.p);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(D64, bD.p, 256 * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hD, cD.p, 256 * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < 256; i++) D32[i] = hD[i];
    return BSA_OK;
}
