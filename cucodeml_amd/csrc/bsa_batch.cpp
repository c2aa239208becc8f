// C ABI of the batch engine: many genes, each with its own tree, alignment
// and model options, evaluated together.  Their jobs share one forest
// (bsa_forest.h), so one evaluation of all of them costs one launch per tree
// level of the deepest tree rather than one per level of every gene; the
// host model set-up (three eigen systems per gene) and the chain rule run on
// a pool of host threads.  bsa_batch_fit runs one optimiser thread per gene
// being fitted (at most 256 at once) and serves their requests in rounds
// (continuous batching): each round evaluates the genes whose points are
// staged, on the device, while the other genes' threads prepare theirs.
#include "bsa_fit.h"
#include "bsa_forest.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <functional>
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

constexpr int kMaxFitThreads = 256;   // optimiser threads of one batch fit

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

// Continuous batching of the fits' evaluation requests.  A fit thread stages
// its gene's point (host model set-up) and queues the gene; the server
// thread runs one round over every gene queued, whenever the device is
// free, while further requests queue for the next round.  A round launches
// the jobs of its genes only, so a gene costs device time only for the
// points its own fit asks for, and a gene's result does not depend on which
// other genes share its round.
struct RoundServer {
    // what a request asks for: the forward pass at the gene's staged point
    // (kept for a later sweep), the adjoint sweep of its last forward pass,
    // or both
    enum { kForward = 1, kSweep = 2, kBoth = 3 };
    std::mutex mu;
    std::condition_variable cv_server;
    std::unique_ptr<std::condition_variable[]> cv_gene;
    Forest *f = nullptr;
    std::vector<std::pair<int, int>> queue;   // (gene, kind) waiting for a round
    std::vector<char> ready;         // per gene: the results of its request are in
    std::vector<int> grc;            // per gene: status of its last round
    std::vector<std::string> gerr;
    int fitting = 0;                 // fits still running
    long rounds = 0;

    explicit RoundServer(Forest *fo) : cv_gene(new std::condition_variable[fo->G]), f(fo)
    {
        ready.assign(f->G, 0);
        grc.assign(f->G, BSA_OK);
        gerr.assign(f->G, std::string());
        fitting = f->G;
    }
    int request(int g, int kind)
    {
        std::unique_lock<std::mutex> lk(mu);
        ready[g] = 0;
        queue.emplace_back(g, kind);
        cv_server.notify_one();
        cv_gene[g].wait(lk, [&] { return ready[g] != 0; });
        return grc[g] ? fail(grc[g], "%s", gerr[g].c_str()) : BSA_OK;
    }
    void finished()
    {
        std::unique_lock<std::mutex> lk(mu);
        fitting--;
        cv_server.notify_one();
    }
    void serve()
    {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            cv_server.wait(lk, [&] { return !queue.empty() || fitting == 0; });
            if (queue.empty()) return;   // every fit has finished
            std::vector<std::pair<int, int>> q;
            q.swap(queue);
            lk.unlock();
            std::vector<int> fwd, adj;
            for (const auto &r : q) {
                if (r.second & kForward) fwd.push_back(r.first);
                if (r.second & kSweep) adj.push_back(r.first);
            }
            std::sort(fwd.begin(), fwd.end());
            std::sort(adj.begin(), adj.end());
            const int rc = forest_run_round(*f, fwd, adj, true);
            const std::string err = rc ? last_error_text() : std::string();
            lk.lock();
            rounds++;
            for (const auto &r : q) {
                const int g = r.first;
                grc[g] = rc;
                gerr[g] = err;
                ready[g] = 1;
                cv_gene[g].notify_one();
            }
        }
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

extern "C" int bsa_batch_eval_genes(bsa_batch *b, const int *genes, int n, const double *const *x,
                                    double *neg_lnl, double *const *grad)
{
    if (!b || !genes || n <= 0 || !x || !neg_lnl) return fail(BSA_EINVAL, "bad argument");
    const int G = (int)b->genes.size();
    std::vector<int> gl(genes, genes + n);
    for (int i = 0; i < n; i++) {
        if (gl[i] < 0 || gl[i] >= G || (i > 0 && gl[i] <= gl[i - 1]))
            return fail(BSA_EINVAL, "genes must be ascending indices below %d", G);
        if (!x[i] || (grad && !grad[i])) return fail(BSA_EINVAL, "null array for entry %d", i);
    }
    const bool adjoint = grad != nullptr;
    Forest &f = b->f;
    HIP_SHARED;
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(f.dev));
    int rc = adjoint ? forest_setup_adjoint(f) : BSA_OK;
    if (rc) return rc;
    rc = parallel_for(n, [&](int i) { return gene_stage_x(*b->genes[gl[i]], f, x[i], adjoint); });
    if (rc) return rc;
    if ((rc = forest_run_round(f, gl, adjoint))) return rc;
    for (int i = 0; i < n; i++) neg_lnl[i] = -gene_collect(*b->genes[gl[i]], f);
    if (adjoint)
        parallel_for(n, [&](int i) {
            gene_chain_rule(*b->genes[gl[i]], f, x[i], grad[i]);
            return BSA_OK;
        });
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    f.timing[6] = ms;
    f.timing_sum[6] += ms;
    f.n_eval += 1;
    return BSA_OK;
}

// One bounded BFGS fit of a gene under its current model options, from x
// (overwritten with the estimates); its evaluations go through the round
// server.
using GeneFit = std::function<int(double *x, const double *lo, const double *hi, FitResult *res)>;

// Runs chain(i, fit) for every gene i on at most kMaxFitThreads optimiser
// threads (a worker takes the next gene when a chain ends) and serves the
// fits' evaluation requests in rounds on this thread.  A chain may change
// its own gene's model options between fits: only the gene's own thread
// stages it, and no request of the gene is pending then.
static int run_fit_chains(bsa_batch *b, const bsa_fit_opts *opts,
                          const std::function<int(int, const GeneFit &)> &chain)
{
    const int G = (int)b->genes.size();
    Forest &f = b->f;
    int rc = forest_setup_adjoint(f);
    if (rc) return rc;
    RoundServer srv(&f);
    std::vector<int> rcs(G, BSA_OK);
    std::vector<std::string> errs(G);
    auto chain_one = [&](int i) {
        Gene &g = *b->genes[i];
        const GeneFit fit = [&](double *x0, const double *lo, const double *hi, FitResult *res) {
            // a line-search trial asks for the forward pass alone; the sweep
            // follows in a later round where the optimiser accepts the point
            const int np = gene_num_params(g);
            std::vector<double> xk(np);   // point of the gene's last forward pass
            auto run = [&](const double *xx, double *fv, int kind) {
                int r = gene_stage_x(g, f, xx, true);
                if (r) return r;
                if ((r = srv.request(i, kind))) return r;
                std::copy(xx, xx + np, xk.begin());
                *fv = -gene_collect(g, f);
                return BSA_OK;
            };
            Objective obj;
            obj.value = [&](const double *xx, double *fv) {
                return run(xx, fv, RoundServer::kForward);
            };
            obj.grad = [&](double *gv) {
                const int r = srv.request(i, RoundServer::kSweep);
                if (r) return r;
                gene_chain_rule(g, f, xk.data(), gv);
                return BSA_OK;
            };
            obj.both = [&](const double *xx, double *fv, double *gv) {
                const int r = run(xx, fv, RoundServer::kBoth);
                if (r) return r;
                gene_chain_rule(g, f, xx, gv);
                return BSA_OK;
            };
            return minimize_box(obj, np, x0, lo, hi, opts, res, f.precision == BSA_FP64,
                                np - g.nbranch);
        };
        const int r = chain(i, fit);
        if (r) {
            rcs[i] = r;
            errs[i] = last_error_text();
        }
        srv.finished();
    };
    {
        std::atomic<int> next{0};
        auto worker = [&]() {
            for (int i = next++; i < G; i = next++) chain_one(i);
        };
        const int nw = std::min(G, kMaxFitThreads);
        std::vector<std::thread> th;
        th.reserve(nw);
        for (int t = 0; t < nw; t++) th.emplace_back(worker);
        srv.serve();   // HIP work on this thread only
        for (auto &t : th) t.join();
    }
    f.n_eval += (double)srv.rounds;
    for (int i = 0; i < G; i++)
        if (rcs[i]) return fail(rcs[i], "gene %d: %s", i, errs[i].c_str());
    return BSA_OK;
}

// every gene evaluated at x[gene] (its last evaluation left there), -lnL out
static int eval_all_at(bsa_batch *b, double *const *x, double *neg_lnl)
{
    Forest &f = b->f;
    const int G = (int)b->genes.size();
    int rc = parallel_for(G, [&](int i) { return gene_stage_x(*b->genes[i], f, x[i], false); });
    if (rc) return rc;
    std::vector<int> all(G);
    for (int i = 0; i < G; i++) all[i] = i;
    if ((rc = forest_run_round(f, all, false))) return rc;
    for (int i = 0; i < G; i++) neg_lnl[i] = -gene_collect(*b->genes[i], f);
    return BSA_OK;
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
    HIP_SHARED;
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(f.dev));
    std::vector<FitResult> res(G);
    int rc = run_fit_chains(b, opts, [&](int i, const GeneFit &fit) {
        return fit(x[i], blo[i].data(), bhi[i].data(), &res[i]);
    });
    if (rc) return rc;
    // leave every gene's last evaluation at its estimates
    if ((rc = eval_all_at(b, x, neg_lnl))) return rc;
    if (info)
        for (int i = 0; i < G; i++) {
            info[3 * i] = res[i].iterations;
            info[3 * i + 1] = res[i].evaluations + 1;
            info[3 * i + 2] = res[i].stop;
        }
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    f.timing[6] = ms;
    f.timing_sum[6] += ms;
    return BSA_OK;
}

extern "C" int bsa_batch_fit_lrt(bsa_batch *b, double *const *x1, double *const *x0,
                                 const bsa_fit_opts *opts, double *neg_lnl, int *info)
{
    if (!b || !x1 || !x0 || !neg_lnl) return fail(BSA_EINVAL, "null argument");
    if (opts && opts->gradient != BSA_GRAD_ADJOINT)
        return fail(BSA_EINVAL, "batch fits use the adjoint gradient");
    const int G = (int)b->genes.size();
    const bsa_model_opts h1{0, 1.0, 1}, h0{1, 1.0, 1};
    for (int i = 0; i < G; i++) {
        if (!x1[i] || !x0[i]) return fail(BSA_EINVAL, "null x for gene %d", i);
        b->genes[i]->opts = h1;
    }
    Forest &f = b->f;
    HIP_SHARED;
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(f.dev));
    std::vector<FitResult> r1(G), r0(G), r1b(G);
    std::vector<double> f1(G), f0(G);
    std::vector<char> refit(G, 0);
    int rc = run_fit_chains(b, opts, [&](int i, const GeneFit &fit) {
        Gene &g = *b->genes[i];
        const int np1 = gene_num_params(g);
        std::vector<double> lo(np1), hi(np1);
        default_bounds(g, lo.data(), hi.data());
        int r = fit(x1[i], lo.data(), hi.data(), &r1[i]);
        if (r) return r;
        f1[i] = r1[i].f;
        // H0 from the H1 estimates without w2 (the last variable)
        g.opts = h0;
        std::memcpy(x0[i], x1[i], sizeof(double) * (np1 - 1));
        default_bounds(g, lo.data(), hi.data());
        r = fit(x0[i], lo.data(), hi.data(), &r0[i]);
        g.opts = h1;
        if (r) return r;
        f0[i] = r0[i].f;
        if (-f0[i] > -f1[i] + 1e-6) {
            // H1 contains H0 (w2 = 1): a lower H1 maximum is a local one;
            // fit H1 again from the H0 estimates and keep the higher maximum
            std::vector<double> xb(np1);
            std::memcpy(xb.data(), x0[i], sizeof(double) * (np1 - 1));
            xb[np1 - 1] = 1.0;
            default_bounds(g, lo.data(), hi.data());
            if ((r = fit(xb.data(), lo.data(), hi.data(), &r1b[i]))) return r;
            refit[i] = 1;
            if (r1b[i].f < f1[i]) {
                std::memcpy(x1[i], xb.data(), sizeof(double) * np1);
                f1[i] = r1b[i].f;
            }
        }
        return BSA_OK;
    });
    if (rc) return rc;
    // leave every gene's last evaluation at its H1 estimates
    std::vector<double> nl(G);
    if ((rc = eval_all_at(b, x1, nl.data()))) return rc;
    for (int i = 0; i < G; i++) {
        neg_lnl[2 * i] = nl[i];
        neg_lnl[2 * i + 1] = f0[i];
        if (info) {
            int *o = info + 7 * i;
            o[0] = r1[i].iterations + (refit[i] ? r1b[i].iterations : 0);
            o[1] = r1[i].evaluations + (refit[i] ? r1b[i].evaluations : 0) + 1;
            // stop reason of the H1 fit whose maximum was kept
            o[2] = refit[i] && r1b[i].f < r1[i].f ? r1b[i].stop : r1[i].stop;
            o[3] = r0[i].iterations;
            o[4] = r0[i].evaluations;
            o[5] = r0[i].stop;
            o[6] = refit[i];
        }
    }
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    f.timing[6] = ms;
    f.timing_sum[6] += ms;
    return BSA_OK;
}

extern "C" int bsa_batch_tune(bsa_batch *b, const double *const *x, int reps, double *out7)
{
    if (!b || !x) return fail(BSA_EINVAL, "null argument");
    const int G = (int)b->genes.size();
    std::vector<std::vector<double>> gs(G);
    std::vector<double *> gp(G);
    std::vector<double> negl(G);
    for (int i = 0; i < G; i++) {
        gs[i].resize(gene_num_params(*b->genes[i]));
        gp[i] = gs[i].data();
    }
    return forest_tune(b->f, [&]() { return batch_eval(b, x, negl.data(), gp.data()); }, reps,
                       out7);
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

extern "C" int bsa_batch_set_prune(bsa_batch *b, int mode)
{
    if (!b) return fail(BSA_EINVAL, "null batch");
    return forest_set_prune(b->f, mode);
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
