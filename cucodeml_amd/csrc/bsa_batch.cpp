// C ABI of the batch engine: many genes, each with its own tree, alignment
// and model options, evaluated together.  Their jobs share one forest
// (bsa_forest.h), so one evaluation of all of them costs one launch per tree
// level of the deepest tree rather than one per level of every gene; the
// host model set-up (three eigen systems per gene) and the chain rule run on
// a pool of host threads.  bsa_batch_fit runs the genes' ML fits in
// lockstep: one optimiser thread per gene, and an evaluation of the whole
// forest whenever every gene still being fitted has asked for one.
#include "bsa_fit.h"
#include "bsa_forest.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

using namespace bsa;

struct bsa_batch {
    std::vector<std::unique_ptr<Gene>> genes;
    Forest f;
};

namespace {

int host_threads()
{
    const unsigned hc = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hc == 0 ? 1u : hc));
}

// fn(i) for i in [0, n) on up to 16 host threads; the first error (with its
// text) is returned
template <class Fn>
int parallel_for(int n, Fn fn)
{
    const int nt = std::max(1, std::min(n, host_threads()));
    std::vector<int> rcs(nt, BSA_OK);
    std::vector<std::string> errs(nt);
    auto work = [&](int t) {
        for (int i = t; i < n; i += nt) {
            const int rc = fn(i);
            if (rc && !rcs[t]) {
                rcs[t] = rc;
                errs[t] = last_error_text();
            }
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work, t);
    work(0);
    for (auto &x : th) x.join();
    for (int t = 0; t < nt; t++)
        if (rcs[t]) return fail(rcs[t], "%s", errs[t].c_str());
    return BSA_OK;
}

// Lockstep evaluation for the fit threads: each thread stages its gene and
// arrives; the thread that completes the set of genes still being fitted
// runs the forest's evaluation and releases the others.  A thread whose fit
// has ended leaves, which may complete the set itself.  Genes that have
// left are evaluated again at their last staged point (their results are
// not read).
struct Lockstep {
    std::mutex mu;
    std::condition_variable cv;
    Forest *f = nullptr;
    int active = 0, arrived = 0, rc = BSA_OK;
    std::string err;
    unsigned long generation = 0;

    void run_locked()
    {
        rc = forest_run(*f, true);
        if (rc) err = last_error_text();
        arrived = 0;
        generation++;
        cv.notify_all();
    }
    int arrive()
    {
        std::unique_lock<std::mutex> lk(mu);
        const unsigned long gen = generation;
        if (++arrived == active)
            run_locked();
        else
            cv.wait(lk, [&] { return generation != gen; });
        return rc ? fail(rc, "%s", err.c_str()) : BSA_OK;
    }
    void leave()
    {
        std::unique_lock<std::mutex> lk(mu);
        active--;
        if (active > 0 && arrived == active) run_locked();
    }
};

int check_gene(const bsa_batch *b, int gene)
{
    if (!b) return fail(BSA_EINVAL, "null batch");
    if (gene < 0 || gene >= (int)b->genes.size())
        return fail(BSA_EINVAL, "gene %d out of range (batch of %zu)", gene, b->genes.size());
    return BSA_OK;
}

}  // namespace

extern "C" int bsa_batch_create(const bsa_problem *const *probs, int ngene, bsa_batch **out)
{
    if (!probs || !out || ngene <= 0) return fail(BSA_EINVAL, "bad argument");
    *out = nullptr;
    HIP_SHARED;
    const bsa_problem *p0 = probs[0];
    if (!p0) return fail(BSA_EINVAL, "null problem 0");
    if (p0->precision != BSA_FP64 && p0->precision != BSA_FP32)
        return fail(BSA_EINVAL, "unknown precision %d", p0->precision);
    for (int i = 1; i < ngene; i++) {
        const bsa_problem *p = probs[i];
        if (!p) return fail(BSA_EINVAL, "null problem %d", i);
        if (p->device != p0->device || p->precision != p0->precision || p->kernel != p0->kernel ||
            p->tile != p0->tile)
            return fail(BSA_EINVAL, "gene %d: the genes of a batch share device, precision, kernel "
                                    "and tile", i);
    }
    auto b = std::make_unique<bsa_batch>();
    b->genes.resize(ngene);
    std::vector<Gene *> gp(ngene);
    const int wb = p0->precision == BSA_FP32 ? 4 : 8;
    for (int i = 0; i < ngene; i++) {
        b->genes[i] = std::make_unique<Gene>();
        gp[i] = b->genes[i].get();
        const int rc = gene_init(*gp[i], probs[i], wb);
        if (rc) return fail(rc, "gene %d: %s", i, last_error_text());
    }
    const int rc = forest_build(b->f, gp, p0->device, p0->precision, p0->kernel, p0->tile);
    if (rc) {
        forest_free(b->f);   // sets no error text
        return rc;
    }
    *out = b.release();
    return BSA_OK;
}

extern "C" int bsa_batch_destroy(bsa_batch *b)
{
    if (!b) return BSA_OK;
    HIP_SHARED;
    forest_free(b->f);
    delete b;
    return BSA_OK;
}

extern "C" int bsa_batch_size(const bsa_batch *b) { return b ? (int)b->genes.size() : -1; }

extern "C" int bsa_batch_set_model(bsa_batch *b, int gene, const bsa_model_opts *o)
{
    if (!b || !o) return fail(BSA_EINVAL, "null argument");
    if (o->fix_omega && !(o->omega_fix > 0)) return fail(BSA_EINVAL, "omega_fix must be > 0");
    if (gene < 0) {
        for (auto &g : b->genes) g->opts = *o;
        return BSA_OK;
    }
    if (int rc = check_gene(b, gene)) return rc;
    b->genes[gene]->opts = *o;
    return BSA_OK;
}

extern "C" int bsa_batch_num_params(const bsa_batch *b, int gene)
{
    if (check_gene(b, gene)) return -1;
    return gene_num_params(*b->genes[gene]);
}

extern "C" int bsa_batch_default_bounds(const bsa_batch *b, int gene, double *lo, double *hi)
{
    if (int rc = check_gene(b, gene)) return rc;
    if (!lo || !hi) return fail(BSA_EINVAL, "null argument");
    default_bounds(*b->genes[gene], lo, hi);
    return BSA_OK;
}

static int batch_eval(bsa_batch *b, const double *const *x, double *neg_lnl, double *const *grad)
{
    if (!b || !x || !neg_lnl) return fail(BSA_EINVAL, "null argument");
    const int G = (int)b->genes.size();
    for (int i = 0; i < G; i++)
        if (!x[i] || (grad && !grad[i])) return fail(BSA_EINVAL, "null array for gene %d", i);
    const bool adjoint = grad != nullptr;
    Forest &f = b->f;
    const int rg = forest_prepare_graph(f, adjoint);
    if (rg) return rg;
    HIP_SHARED;
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(f.dev));
    int rc = adjoint ? forest_setup_adjoint(f) : BSA_OK;
    if (rc) return rc;
    rc = parallel_for(G, [&](int i) { return gene_stage_x(*b->genes[i], f, x[i], adjoint); });
    if (rc) return rc;
    if ((rc = forest_run(f, adjoint))) return rc;
    for (int i = 0; i < G; i++) neg_lnl[i] = -gene_collect(*b->genes[i], f);
    if (adjoint)
        parallel_for(G, [&](int i) {
            gene_chain_rule(*b->genes[i], f, x[i], grad[i]);
            return BSA_OK;
        });
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    f.timing[6] = ms;
    f.timing_sum[6] += ms;
    f.n_eval += 1;
    return BSA_OK;
}

extern "C" int bsa_batch_lnl(bsa_batch *b, const double *const *x, double *neg_lnl)
{
    return batch_eval(b, x, neg_lnl, nullptr);
}

extern "C" int bsa_batch_lnl_grad(bsa_batch *b, const double *const *x, double *neg_lnl,
                                  double *const *grad)
{
    if (!grad) return fail(BSA_EINVAL, "null argument");
    return batch_eval(b, x, neg_lnl, grad);
}

extern "C" int bsa_batch_fit(bsa_batch *b, double *const *x, const double *const *lo,
                             const double *const *hi, const bsa_fit_opts *opts, double *neg_lnl,
                             int *info)
{
    if (!b || !x || !neg_lnl) return fail(BSA_EINVAL, "null argument");
    if (opts && opts->gradient != BSA_GRAD_ADJOINT)
        return fail(BSA_EINVAL, "batch fits use the adjoint gradient");
    const int G = (int)b->genes.size();
    std::vector<std::vector<double>> blo(G), bhi(G);
    for (int i = 0; i < G; i++) {
        if (!x[i]) return fail(BSA_EINVAL, "null x for gene %d", i);
        const Gene &g = *b->genes[i];
        const int np = gene_num_params(g);
        blo[i].resize(np);
        bhi[i].resize(np);
        default_bounds(g, blo[i].data(), bhi[i].data());
        if (lo && lo[i]) std::memcpy(blo[i].data(), lo[i], sizeof(double) * np);
        if (hi && hi[i]) std::memcpy(bhi[i].data(), hi[i], sizeof(double) * np);
        for (int k = 0; k < np; k++)
            if (!(blo[i][k] <= bhi[i][k])) return fail(BSA_EINVAL, "gene %d: bounds with lo > hi", i);
    }
    Forest &f = b->f;
    int rc = forest_prepare_graph(f, true);
    if (!rc) rc = forest_prepare_graph(f, false);   // the closing evaluation
    if (rc) return rc;
    HIP_SHARED;
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(f.dev));
    if ((rc = forest_setup_adjoint(f))) return rc;
    Lockstep ls;
    ls.f = &f;
    ls.active = G;
    std::vector<FitResult> res(G);
    std::vector<int> rcs(G, BSA_OK);
    std::vector<std::string> errs(G);
    auto fit_one = [&](int i) {
        Gene &g = *b->genes[i];
        const int np = gene_num_params(g);
        ObjFn fg = [&](const double *xx, double *fv, double *gv) {
            int r = gene_stage_x(g, f, xx, true);
            if (r) return r;
            if ((r = ls.arrive())) return r;
            *fv = -gene_collect(g, f);
            gene_chain_rule(g, f, xx, gv);
            return BSA_OK;
        };
        const int r = minimize_box(fg, np, x[i], blo[i].data(), bhi[i].data(), opts, &res[i]);
        if (r) {
            rcs[i] = r;
            errs[i] = last_error_text();
        }
        ls.leave();
    };
    {
        std::vector<std::thread> th;
        th.reserve(G);
        for (int i = 0; i < G; i++) th.emplace_back(fit_one, i);
        for (auto &t : th) t.join();
    }
    for (int i = 0; i < G; i++)
        if (rcs[i]) return fail(rcs[i], "gene %d: %s", i, errs[i].c_str());
    // leave every gene's last evaluation at its estimates
    rc = parallel_for(G, [&](int i) { return gene_stage_x(*b->genes[i], f, x[i], false); });
    if (rc) return rc;
    if ((rc = forest_run(f, false))) return rc;
    for (int i = 0; i < G; i++) {
        neg_lnl[i] = -gene_collect(*b->genes[i], f);
        if (info) {
            info[3 * i] = res[i].iterations;
            info[3 * i + 1] = res[i].evaluations + 1;
            info[3 * i + 2] = res[i].converged;
        }
    }
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    f.timing[6] = ms;
    f.timing_sum[6] += ms;
    return BSA_OK;
}

extern "C" int bsa_batch_class_lnl(bsa_batch *b, int gene, double *fhK, double *site)
{
    if (int rc = check_gene(b, gene)) return rc;
    const Gene &g = *b->genes[gene];
    if (!g.evaluated) return fail(BSA_EINVAL, "no evaluation yet");
    HIP_SHARED;
    const Forest &f = b->f;
    HIPCHK(hipSetDevice(f.dev));
    const size_t np = f.npad;
    if (fhK)
        for (int c = 0; c < 4; c++)
            HIPCHK(hipMemcpy(fhK + (size_t)c * g.npatt, f.d_fh + (4 * (size_t)g.gi + c) * np,
                             sizeof(double) * g.npatt, hipMemcpyDeviceToHost));
    if (site)
        HIPCHK(hipMemcpy(site, f.d_site + (size_t)g.gi * np, sizeof(double) * g.npatt,
                         hipMemcpyDeviceToHost));
    return BSA_OK;
}

extern "C" int bsa_batch_last_params(const bsa_batch *b, int gene, double *o)
{
    if (int rc = check_gene(b, gene)) return rc;
    if (!o) return fail(BSA_EINVAL, "null argument");
    const Gene &g = *b->genes[gene];
    if (!g.evaluated) return fail(BSA_EINVAL, "no evaluation yet");
    const ModelA &m = g.model;
    const double v[11] = {m.kappa, m.p0, m.p1, m.w0, m.w2, m.Qfactor[0], m.Qfactor[1],
                          m.freqK[0], m.freqK[1], m.freqK[2], m.freqK[3]};
    std::memcpy(o, v, sizeof v);
    return BSA_OK;
}

extern "C" int bsa_batch_timing(bsa_batch *b, double *o, int reset)
{
    if (!b || !o) return fail(BSA_EINVAL, "null argument");
    Forest &f = b->f;
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

extern "C" int bsa_batch_work(bsa_batch *b, double *o)
{
    if (!b || !o) return fail(BSA_EINVAL, "null argument");
    HIP_SHARED;
    const int rc = forest_setup_adjoint(b->f);
    if (rc) return rc;
    std::memcpy(o, b->f.work, sizeof b->f.work);
    return BSA_OK;
}

extern "C" int bsa_batch_set_graph(bsa_batch *b, int enable)
{
    if (!b) return fail(BSA_EINVAL, "null batch");
    b->f.use_graph = enable != 0;
    return BSA_OK;
}

// Plan a forest of the given genes on the host only (no device work) and
// check every index of its forward and adjoint schedules; out[8] = {genes,
// padded patterns, partial slots, pruning launches, pruning jobs, adjoint
// jobs, eigen-basis jobs, branch-matrix chunks}.
extern "C" int bsa_debug_forest_plan(const bsa_problem *const *probs, int ngene, double *out8)
{
    if (!probs || ngene <= 0 || !out8) return fail(BSA_EINVAL, "bad argument");
    std::vector<std::unique_ptr<Gene>> genes(ngene);
    std::vector<Gene *> gp(ngene);
    const int prec = probs[0] ? probs[0]->precision : BSA_FP64;
    for (int i = 0; i < ngene; i++) {
        if (!probs[i]) return fail(BSA_EINVAL, "null problem %d", i);
        genes[i] = std::make_unique<Gene>();
        gp[i] = genes[i].get();
        const int rc = gene_init(*gp[i], probs[i], prec == BSA_FP32 ? 4 : 8);
        if (rc) return fail(rc, "gene %d: %s", i, last_error_text());
    }
    Forest f;
    int rc = forest_plan(f, gp, prec, probs[0]->kernel, probs[0]->tile);
    if (rc) return rc;
    forest_plan_adjoint(f);
    if ((rc = forest_check(f))) return rc;
    const double v[8] = {(double)f.G, (double)f.npad, (double)f.nslot,
                         (double)(f.level_off.size() - 1), (double)f.prune.size(),
                         (double)f.adj.size(), (double)f.ejobs.size(), (double)f.nchunk};
    std::memcpy(out8, v, sizeof v);
    return BSA_OK;
}
