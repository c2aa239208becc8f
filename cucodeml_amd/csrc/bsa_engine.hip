// C ABI of the branch-site model A engine: problem set-up (tree schedule,
// device buffers), one likelihood evaluation per call, finite-difference
// gradients, per-site outputs.  Declarations and the reference entry points
// each call replaces are in include/bsa.h.
#include "bsa.h"
#include "bsa_kernels.h"
#include "bsa_model.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <vector>

using namespace bsa;

static thread_local std::string g_err;

// A graph capture must not see HIP calls of other threads that would
// invalidate it (allocations, synchronous copies): every entry point that
// calls HIP holds this lock shared, a capture holds it exclusively.
static std::shared_mutex g_hip_mu;
#define HIP_SHARED std::shared_lock<std::shared_mutex> hip_shared_lock_(g_hip_mu)

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return fail(BSA_EDEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                          \
    } while (0)

constexpr int kUVRSet = 64 * 64 * 2 + 64;
// per-evaluation model block on the device: the three U/V/root sets, then
// the four class proportions
constexpr int kModelBlock = 3 * kUVRSet + 8;

struct bsa_engine {
    int dev = 0;
    int ns = 0, nnode = 0, root = 0, nbranch = 0, npatt = 0, npad = 0, ncodes = 0, namb = 0;
    int precision = BSA_FP64, kernel = BSA_KERNEL_MFMA, tile = 64;
    std::vector<int> father, label, ibranch, height;
    std::vector<std::vector<int>> sons;
    std::vector<char> onpath;
    std::vector<double> pi, fpatt;
    bsa_model_opts opts{0, 1.0, 1};

    // schedule
    std::vector<PmatJob> pjobs;
    std::vector<PruneJob> prune;          // all jobs, grouped by level
    std::vector<int> level_off;           // level l: prune[level_off[l] .. level_off[l+1])
    int nslot = 0;
    int root_slot[4];

    // device
    hipStream_t stream = nullptr;
    hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    // device work of one evaluation captured as a graph (lnL only, with gradient)
    bool use_graph = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    double *d_uvr = nullptr;
    PmatJob *d_pjobs = nullptr;
    PruneJob *d_prune = nullptr;
    void *d_PT = nullptr, *d_AT = nullptr, *d_L = nullptr;
    int *d_E = nullptr;
    unsigned char *d_z = nullptr, *d_mask_amb = nullptr;
    double *d_pi = nullptr, *d_fpatt = nullptr, *d_fh = nullptr, *d_site = nullptr;
    double *d_partial = nullptr, *d_out = nullptr;
    int nroot_blk = 0;

    // host staging (pinned)
    double *h_uvr = nullptr;
    double *h_out = nullptr;
    PmatJob *h_pjobs = nullptr;

    // model cache and last evaluation
    ModelA model;
    bool have_eigen = false;
    double cache_kappa = NAN, cache_w[3] = {NAN, NAN, NAN};
    double last_lnl = NAN;
    // pmat, prune, root, child adjoints, branch matrices, adjoint finish, wall
    double timing[7] = {0, 0, 0, 0, 0, 0, 0};
    double timing_sum[7] = {0, 0, 0, 0, 0, 0, 0};
    double n_eval = 0, n_launch = 0, n_adj_launch = 0;
    bool evaluated = false;
    double work[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

    // adjoint (allocated on the first gradient call)
    bool adj_ready = false;
    std::vector<AdjJob> adj;              // grouped by depth of the child node
    std::vector<int> adj_off;             // depth d: adj[adj_off[d-1] .. adj_off[d])
    std::vector<int> adj_nint;            // depth d: the first adj_nint[d-1] jobs have an interior child
    std::vector<EigJob> ejobs;
    std::vector<int> ejob_node;           // node of each EigJob
    std::vector<int> set_off, set_list;
    int nchunk = 0, tpc = 0, nmrow = 0;   // branch-matrix chunks, tiles per chunk
    int root_aslot[4];
    AdjJob *d_adj = nullptr;
    EigJob *d_ejobs = nullptr;
    EigJob *h_ejobs = nullptr;
    int *d_set_off = nullptr, *d_set_list = nullptr;
    void *d_A = nullptr, *d_C = nullptr;
    int *d_EA = nullptr;
    double *d_Mpart = nullptr, *d_W = nullptr, *d_G = nullptr, *d_dtau = nullptr, *d_Z = nullptr;
    double *d_mask64 = nullptr;
    double *h_adjout = nullptr;           // dtau[nej] then Z[3*4096]
    std::vector<unsigned char> code_mask; // ncodes x 61
    std::vector<std::vector<int>> slot;   // inside slot per node and class
};

static size_t tsize(const bsa_engine *e) { return e->precision == BSA_FP32 ? 4 : 8; }

static void free_all(bsa_engine *e)
{
    if (e->stream) hipSetDevice(e->dev);
    void *dev_ptrs[] = {e->d_uvr, e->d_pjobs, e->d_prune, e->d_PT, e->d_AT, e->d_L, e->d_E,
                        e->d_z, e->d_mask_amb, e->d_pi, e->d_fpatt, e->d_fh, e->d_site,
                        e->d_partial, e->d_out, e->d_adj, e->d_ejobs, e->d_set_off,
                        e->d_set_list, e->d_A, e->d_C, e->d_EA, e->d_Mpart, e->d_W, e->d_G,
                        e->d_dtau, e->d_Z, e->d_mask64};
    for (void *p : dev_ptrs)
        if (p) hipFree(p);
    if (e->h_uvr) hipHostFree(e->h_uvr);
    if (e->h_out) hipHostFree(e->h_out);
    if (e->h_pjobs) hipHostFree(e->h_pjobs);
    if (e->h_ejobs) hipHostFree(e->h_ejobs);
    if (e->h_adjout) hipHostFree(e->h_adjout);
    for (auto &v : e->ev)
        if (v) hipEventDestroy(v);
    for (auto &g : e->graph)
        if (g) hipGraphExecDestroy(g);
    if (e->stream) hipStreamDestroy(e->stream);
}

// ---------------------------------------------------------------- set-up

static int build_schedule(bsa_engine *e)
{
    const int ns = e->ns, nn = e->nnode;
    // heights (tips 0), levels of interior nodes
    e->height.assign(nn, 0);
    std::vector<int> order;   // postorder
    {
        std::vector<std::pair<int, int>> st{{e->root, 0}};
        while (!st.empty()) {
            auto &[v, k] = st.back();
            if (k < (int)e->sons[v].size()) {
                int c = e->sons[v][k++];
                st.push_back({c, 0});
            } else {
                order.push_back(v);
                st.pop_back();
            }
        }
    }
    if ((int)order.size() != nn) return fail(BSA_EINVAL, "tree is not connected from the root");
    for (int v : order) {
        int h = 0;
        for (int c : e->sons[v]) h = std::max(h, e->height[c] + 1);
        e->height[v] = (v < ns) ? 0 : h;
    }
    // nodes whose site classes 2/3 differ from 0/1: proper ancestors of a
    // foreground branch
    e->onpath.assign(nn, 0);
    int nfg = 0;
    for (int v = 0; v < nn; v++) {
        if (v == e->root || e->label[v] != 1) continue;
        nfg++;
        for (int a = e->father[v]; a >= 0; a = e->father[a]) e->onpath[a] = 1;
    }
    if (nfg == 0) return fail(BSA_EINVAL, "branch-site model A needs a foreground (#1) branch");
    // partial slots
    std::vector<std::array<int, 4>> slot(nn, std::array<int, 4>{-1, -1, -1, -1});
    int ns_ = 0;
    for (int v = ns; v < nn; v++) {
        slot[v][0] = ns_++;
        slot[v][1] = ns_++;
        if (e->onpath[v]) {
            slot[v][2] = ns_++;
            slot[v][3] = ns_++;
        } else {
            slot[v][2] = slot[v][0];
            slot[v][3] = slot[v][1];
        }
    }
    e->nslot = ns_;
    for (int c = 0; c < 4; c++) e->root_slot[c] = slot[e->root][c];
    e->slot.assign(nn, std::vector<int>(4, -1));
    for (int v = 0; v < nn; v++)
        for (int c = 0; c < 4; c++) e->slot[v][c] = slot[v][c];
    // pruning jobs grouped by height
    int hmax = e->height[e->root];
    e->prune.clear();
    e->level_off.assign(1, 0);
    for (int lev = 1; lev <= hmax; lev++) {
        for (int v = ns; v < nn; v++) {
            if (e->height[v] != lev) continue;
            int ncls = e->onpath[v] ? 4 : 2;
            for (int c = 0; c < ncls; c++) {
                PruneJob j{};
                j.out_slot = slot[v][c];
                j.nch = (int)e->sons[v].size();
                for (int k = 0; k < j.nch; k++) {
                    int s = e->sons[v][k];
                    j.ch_tip[k] = s < ns ? s : -1;
                    j.ch_slot[k] = s < ns ? -1 : slot[s][c];
                    j.ch_pmat[k] = s * 3 + uvr_set(c, e->label[s]);
                    j.ch_cslot[k] = s < ns ? -1 : (s - ns) * 4 + c;
                }
                e->prune.push_back(j);
            }
        }
        e->level_off.push_back((int)e->prune.size());
    }
    // work per evaluation of the pruning pass, counted on 61 states
    {
        const double w = e->precision == BSA_FP32 ? 4.0 : 8.0, np_ = e->npatt;
        double fl = 0, by = 0;
        for (const PruneJob &j : e->prune) {
            int nint = 0, ntip = 0;
            for (int k = 0; k < j.nch; k++) (j.ch_tip[k] >= 0 ? ntip : nint)++;
            fl += 2.0 * kS * kS * np_ * nint;
            by += np_ * (kS * w * (nint + 1) + ntip + 4.0 * (nint + 1));
        }
        int nint_nodes = nn - ns, nint_edges = 0;
        for (int v = ns; v < nn; v++)
            if (v != e->root) nint_edges++;
        e->work[0] = fl;
        e->work[1] = by;
        e->work[2] = 4.0 * np_ * nint_edges * 2.0 * kS * kS;
        e->work[3] = 4.0 * np_ * kS * w * (nint_nodes + nint_edges);
        e->work[4] = (double)(e->level_off.size() - 1);
        e->work[5] = (double)e->prune.size();
    }
    // P(t) jobs: every branch for eigen sets w0 and 1; foreground also w2
    e->pjobs.clear();
    for (int v = 0; v < nn; v++) {
        if (v == e->root) continue;
        int nset = e->label[v] == 1 ? 3 : 2;
        for (int s = 0; s < nset; s++) {
            PmatJob p{};
            p.node = v;
            p.set = s;
            p.pm = v * 3 + s;
            p.tip = v < ns ? v : -1;
            p.t = 0;
            e->pjobs.push_back(p);
        }
    }
    return BSA_OK;
}

static int upload_static(bsa_engine *e, const bsa_problem *p)
{
    const size_t ts = tsize(e);
    const int np = e->npad;
    HIPCHK(hipSetDevice(e->dev));
    HIPCHK(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    for (auto &v : e->ev) HIPCHK(hipEventCreate(&v));
    HIPCHK(hipMalloc(&e->d_uvr, sizeof(double) * kModelBlock));
    HIPCHK(hipMalloc(&e->d_pjobs, sizeof(PmatJob) * e->pjobs.size()));
    HIPCHK(hipMalloc(&e->d_prune, sizeof(PruneJob) * std::max<size_t>(1, e->prune.size())));
    HIPCHK(hipMalloc(&e->d_PT, sizeof(double) * 4096 * (size_t)e->nnode * 3));
    HIPCHK(hipMemset(e->d_PT, 0, sizeof(double) * 4096 * (size_t)e->nnode * 3));
    // tip table: every code's row for every tip branch and eigen set
    HIPCHK(hipMalloc(&e->d_AT, sizeof(double) * 64 * (size_t)e->ncodes * e->ns * 3));
    if (e->namb > 0) {
        HIPCHK(hipMalloc(&e->d_mask_amb, 64 * (size_t)e->namb));
    }
    HIPCHK(hipMalloc(&e->d_L, ts * 64 * (size_t)np * e->nslot));
    HIPCHK(hipMalloc(&e->d_E, sizeof(int) * (size_t)np * e->nslot));
    HIPCHK(hipMalloc(&e->d_z, (size_t)np * e->ns));
    HIPCHK(hipMalloc(&e->d_pi, sizeof(double) * 64));
    HIPCHK(hipMalloc(&e->d_fpatt, sizeof(double) * np));
    HIPCHK(hipMalloc(&e->d_fh, sizeof(double) * 4 * (size_t)np));
    HIPCHK(hipMalloc(&e->d_site, sizeof(double) * np));
    e->nroot_blk = e->npad / 16;   // k_root: 16 patterns per block, padding included
    HIPCHK(hipMalloc(&e->d_partial, sizeof(double) * 2 * kRootSums * e->nroot_blk));
    HIPCHK(hipMalloc(&e->d_out, sizeof(double) * 8));
    HIPCHK(hipHostMalloc((void **)&e->h_uvr, sizeof(double) * kModelBlock, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&e->h_out, sizeof(double) * 8, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&e->h_pjobs, sizeof(PmatJob) * e->pjobs.size(),
                         hipHostMallocDefault));

    // static inputs
    std::vector<unsigned char> z((size_t)np * e->ns, 0);
    for (int s = 0; s < e->ns; s++)
        std::memcpy(&z[(size_t)s * np], p->z + (size_t)s * e->npatt, e->npatt);
    HIPCHK(hipMemcpy(e->d_z, z.data(), z.size(), hipMemcpyHostToDevice));
    if (e->namb > 0) {
        std::vector<unsigned char> m(64 * (size_t)e->namb, 0);
        for (int a = 0; a < e->namb; a++)
            for (int k = 0; k < kS; k++) m[a * 64 + k] = p->code_mask[(size_t)(kS + a) * kS + k] ? 1 : 0;
        HIPCHK(hipMemcpy(e->d_mask_amb, m.data(), m.size(), hipMemcpyHostToDevice));
    }
    std::vector<double> pi64(64, 0.0), fp(np, 0.0);
    for (int i = 0; i < kS; i++) pi64[i] = e->pi[i];
    for (int h = 0; h < e->npatt; h++) fp[h] = e->fpatt[h];
    HIPCHK(hipMemcpy(e->d_pi, pi64.data(), sizeof(double) * 64, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_fpatt, fp.data(), sizeof(double) * np, hipMemcpyHostToDevice));
    if (!e->prune.empty())
        HIPCHK(hipMemcpy(e->d_prune, e->prune.data(), sizeof(PruneJob) * e->prune.size(),
                         hipMemcpyHostToDevice));
    return BSA_OK;
}

extern "C" int bsa_create(const bsa_problem *p, bsa_engine **out)
{
    if (!p || !out) return fail(BSA_EINVAL, "null argument");
    *out = nullptr;
    HIP_SHARED;
    if (p->ns < 3) return fail(BSA_EINVAL, "need at least 3 sequences (ns=%d)", p->ns);
    if (p->nnode <= p->ns || p->nnode > 2 * p->ns - 1)
        return fail(BSA_EINVAL, "nnode=%d inconsistent with ns=%d", p->nnode, p->ns);
    if (p->root < p->ns || p->root >= p->nnode) return fail(BSA_EINVAL, "root must be interior");
    if (p->npatt <= 0) return fail(BSA_EINVAL, "npatt must be positive");
    if (p->ncodes < kS || p->ncodes > 256) return fail(BSA_EINVAL, "ncodes=%d out of range", p->ncodes);
    if (!p->father || !p->label || !p->ibranch || !p->z || !p->fpatt || !p->pi)
        return fail(BSA_EINVAL, "null array in problem");
    if (p->ncodes > kS && !p->code_mask) return fail(BSA_EINVAL, "ambiguity codes need code_mask");
    if (p->precision != BSA_FP64 && p->precision != BSA_FP32)
        return fail(BSA_EINVAL, "unknown precision %d", p->precision);
    if (p->kernel != BSA_KERNEL_MFMA && p->kernel != BSA_KERNEL_VALU)
        return fail(BSA_EINVAL, "unknown kernel %d", p->kernel);

    bsa_engine *e = new bsa_engine();
    e->dev = p->device;
    e->ns = p->ns;
    e->nnode = p->nnode;
    e->root = p->root;
    e->nbranch = p->nnode - 1;
    e->npatt = p->npatt;
    e->npad = (p->npatt + kPadP - 1) / kPadP * kPadP;
    e->tile = p->tile == 0 ? 64 : p->tile;
    if (e->tile != 64 && e->tile != 128) {
        int t = e->tile;
        delete e;
        return fail(BSA_EINVAL, "tile %d: 64 or 128 patterns per block", t);
    }
    e->ncodes = p->ncodes;
    e->namb = p->ncodes - kS;
    e->precision = p->precision;
    e->kernel = p->kernel;
    e->father.assign(p->father, p->father + p->nnode);
    e->label.assign(p->label, p->label + p->nnode);
    e->ibranch.assign(p->ibranch, p->ibranch + p->nnode);
    e->sons.assign(p->nnode, {});
    std::vector<char> seen_b(e->nbranch, 0);
    for (int v = 0; v < p->nnode; v++) {
        int f = e->father[v];
        if (v == p->root) {
            if (f != -1) { delete e; return fail(BSA_EINVAL, "father[root] must be -1"); }
            continue;
        }
        if (f < p->ns || f >= p->nnode) {
            delete e;
            return fail(BSA_EINVAL, "node %d has invalid father %d", v, f);
        }
        if (e->label[v] != 0 && e->label[v] != 1) {
            delete e;
            return fail(BSA_EINVAL, "branch label of node %d is %d; model A allows 0 and 1", v,
                        e->label[v]);
        }
        int b = e->ibranch[v];
        if (b < 0 || b >= e->nbranch || seen_b[b]) {
            delete e;
            return fail(BSA_EINVAL, "ibranch of node %d (%d) is not a permutation", v, b);
        }
        seen_b[b] = 1;
        e->sons[f].push_back(v);
    }
    for (int v = p->ns; v < p->nnode; v++) {
        if (e->sons[v].empty()) { delete e; return fail(BSA_EINVAL, "interior node %d has no sons", v); }
        if ((int)e->sons[v].size() > kMaxCh) {
            delete e;
            return fail(BSA_EINVAL, "node %d has %zu sons (max %d)", v, e->sons[v].size(), kMaxCh);
        }
    }
    for (size_t i = 0; i < (size_t)p->ns * p->npatt; i++)
        if (p->z[i] >= p->ncodes) { delete e; return fail(BSA_EINVAL, "code %d >= ncodes", p->z[i]); }
    e->pi.assign(p->pi, p->pi + kS);
    e->fpatt.assign(p->fpatt, p->fpatt + p->npatt);
    e->code_mask.assign((size_t)p->ncodes * kS, 0);
    for (int c = 0; c < kS; c++) e->code_mask[(size_t)c * kS + c] = 1;
    for (size_t i = (size_t)kS * kS; i < (size_t)p->ncodes * kS; i++)
        e->code_mask[i] = p->code_mask[i] ? 1 : 0;
    int rc = build_schedule(e);
    if (rc) { delete e; return rc; }
    rc = upload_static(e, p);
    if (rc) { std::string m = g_err; free_all(e); delete e; g_err = m; return rc; }
    *out = e;
    return BSA_OK;
}

extern "C" int bsa_destroy(bsa_engine *e)
{
    if (!e) return BSA_OK;
    HIP_SHARED;
    free_all(e);
    delete e;
    return BSA_OK;
}

extern "C" int bsa_set_model(bsa_engine *e, const bsa_model_opts *o)
{
    if (!e || !o) return fail(BSA_EINVAL, "null argument");
    if (o->fix_omega && !(o->omega_fix > 0)) return fail(BSA_EINVAL, "omega_fix must be > 0");
    e->opts = *o;
    return BSA_OK;
}

extern "C" int bsa_num_params(const bsa_engine *e)
{
    if (!e) return -1;
    return e->nbranch + (e->opts.fix_omega ? 4 : 5);
}

// ---------------------------------------------------------------- evaluation

// x (codeml variables) -> natural parameters; f_and_x, tools.c:1280-1305
static void natural_params(const bsa_engine *e, const double *x, double *kappa, double *p0,
                           double *p1, double *w0, double *w2)
{
    const int nb = e->nbranch;
    *kappa = x[nb];
    if (e->opts.p_transformed) {
        double a = std::exp(x[nb + 1]), b = std::exp(x[nb + 2]), tot = 1 + a + b;
        *p0 = a / tot;
        *p1 = b / tot;
    } else {
        *p0 = x[nb + 1];
        *p1 = x[nb + 2];
    }
    *w0 = x[nb + 3];
    *w2 = e->opts.fix_omega ? e->opts.omega_fix : x[nb + 4];
}

template <typename T>
static int enqueue_device(bsa_engine *e, hipStream_t s, bool adjoint, bool events)
{
#define REC(i) \
    if (events) HIPCHK(hipEventRecord(e->ev[i], s))
    REC(0);
    launch_pmat(s, e->d_uvr, e->d_pjobs, (int)e->pjobs.size(), (double *)e->d_PT, e->d_mask_amb,
                e->ncodes, (double *)e->d_AT);
    HIPCHK(hipGetLastError());
    REC(1);
    for (size_t l = 0; l + 1 < e->level_off.size(); l++) {
        int a = e->level_off[l], b = e->level_off[l + 1];
        launch_prune<T>(s, e->kernel, e->tile, e->d_prune + a, b - a, e->npad, e->d_z,
                        (const double *)e->d_PT, (const double *)e->d_AT, e->ncodes, (T *)e->d_L,
                        e->d_E, adjoint ? (T *)e->d_C : nullptr);
    }
    HIPCHK(hipGetLastError());
    REC(2);
    RootArgs ra;
    ra.npatt = e->npatt;
    ra.npad = e->npad;
    for (int c = 0; c < 4; c++) {
        ra.root_slot[c] = e->root_slot[c];
        ra.seed_aslot[c] = adjoint ? e->root_aslot[c] : -1;
    }
    launch_root<T>(s, ra, (const T *)e->d_L, e->d_E, e->d_pi, e->d_uvr + 3 * kUVRSet, e->d_fpatt,
                   e->d_fh, e->d_site, e->d_partial, e->nroot_blk, e->d_out, (T *)e->d_A,
                   e->d_EA);
    HIPCHK(hipGetLastError());
    REC(3);
    if (adjoint) {
        for (size_t d = 0; d + 1 < e->adj_off.size(); d++)
            launch_adjoint_level<T>(s, e->d_adj + e->adj_off[d], e->adj_nint[d], e->npad, e->d_z,
                                    (const double *)e->d_PT, (const double *)e->d_AT, e->ncodes,
                                    e->d_E, (const T *)e->d_C, (T *)e->d_A, e->d_EA);
        REC(6);
        launch_adjoint_m<T>(s, e->d_adj, (int)e->adj.size(), e->npad, e->nchunk, e->tpc, e->d_z,
                            (const double *)e->d_PT, (const double *)e->d_AT, e->ncodes, e->d_mask64,
                            (const T *)e->d_L, e->d_E, (const T *)e->d_C, (const T *)e->d_A,
                            e->d_EA, e->d_Mpart);
        REC(4);
        launch_adjoint_finish(s, e->d_Mpart, e->nchunk, e->d_ejobs, (int)e->ejobs.size(),
                              e->d_set_off, e->d_set_list, e->d_uvr, e->d_G, e->d_dtau, e->d_W,
                              e->d_Z);
        HIPCHK(hipGetLastError());
        REC(5);
    }
#undef REC
    return BSA_OK;
}

template <typename T>
static int run_device(bsa_engine *e, bool adjoint)
{
    hipStream_t s = e->stream;
    HIPCHK(hipMemcpyAsync(e->d_uvr, e->h_uvr, sizeof(double) * kModelBlock, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(e->d_pjobs, e->h_pjobs, sizeof(PmatJob) * e->pjobs.size(),
                          hipMemcpyHostToDevice, s));
    if (adjoint)
        HIPCHK(hipMemcpyAsync(e->d_ejobs, e->h_ejobs, sizeof(EigJob) * e->ejobs.size(),
                              hipMemcpyHostToDevice, s));
    if (e->use_graph) {
        // the launch sequence depends only on the tree and the data, so it is
        // captured once and replayed; per-evaluation inputs live in device
        // buffers refreshed by the copies above
        hipGraphExec_t g = e->graph[adjoint ? 1 : 0];
        if (!g) return fail(BSA_EDEVICE, "evaluation graph not captured");
        HIPCHK(hipEventRecord(e->ev[0], s));
        HIPCHK(hipGraphLaunch(g, s));
        HIPCHK(hipEventRecord(e->ev[5], s));
    } else {
        const int rc = enqueue_device<T>(e, s, adjoint, true);
        if (rc) return rc;
    }
    if (adjoint) {
        HIPCHK(hipMemcpyAsync(e->h_adjout, e->d_dtau, sizeof(double) * e->ejobs.size(),
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(e->h_adjout + e->ejobs.size(), e->d_Z, sizeof(double) * 3 * 4096,
                              hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipMemcpyAsync(e->h_out, e->d_out, sizeof(double) * kRootSums, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    float ms;
    if (e->use_graph) {
        // one device time for the whole evaluation (out[0] of bsa_timing)
        HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[5]));
        for (int i = 0; i < 6; i++) e->timing[i] = 0;
        e->timing[0] = ms;
        for (int i = 0; i < 6; i++) e->timing_sum[i] += e->timing[i];
        return BSA_OK;
    }
    HIPCHK(hipEventElapsedTime(&ms, e->ev[0], e->ev[1]));
    e->timing[0] = ms;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[1], e->ev[2]));
    e->timing[1] = ms;
    HIPCHK(hipEventElapsedTime(&ms, e->ev[2], e->ev[3]));
    e->timing[2] = ms;
    e->timing[3] = e->timing[4] = e->timing[5] = 0;
    if (adjoint) {
        HIPCHK(hipEventElapsedTime(&ms, e->ev[3], e->ev[6]));
        e->timing[3] = ms;
        HIPCHK(hipEventElapsedTime(&ms, e->ev[6], e->ev[4]));
        e->timing[4] = ms;
        HIPCHK(hipEventElapsedTime(&ms, e->ev[4], e->ev[5]));
        e->timing[5] = ms;
        int nl = 1;   // the branch-matrix launch, plus one per depth with interior children
        for (int v : e->adj_nint) nl += v > 0;
        e->n_adj_launch += nl;
    }
    for (int i = 0; i < 6; i++) e->timing_sum[i] += e->timing[i];
    e->n_launch += (double)(e->level_off.size() - 1);
    return BSA_OK;
}

// Capture the device work of one evaluation as a graph (the launch sequence
// depends only on the tree and the data; per-evaluation inputs are copied to
// device buffers before each replay).
template <typename T>
static int capture_graph(bsa_engine *e, bool adjoint)
{
    hipStream_t s = e->stream;
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_device<T>(e, s, adjoint, false);
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (rc) {
        if (graph) hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(BSA_EDEVICE, "graph capture: %s", hipGetErrorString(ec));
    hipGraphExec_t g = nullptr;
    const hipError_t ei = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(BSA_EDEVICE, "graph instantiation: %s", hipGetErrorString(ei));
    e->graph[adjoint ? 1 : 0] = g;
    return BSA_OK;
}

static int setup_adjoint(bsa_engine *e);

// In graph mode, capture the evaluation's graph before its first use, with
// no other HIP work of this process running.
static int prepare_graph(bsa_engine *e, bool adjoint)
{
    if (!e->use_graph || e->graph[adjoint ? 1 : 0]) return BSA_OK;
    std::unique_lock<std::shared_mutex> lock(g_hip_mu);
    if (e->graph[adjoint ? 1 : 0]) return BSA_OK;
    HIPCHK(hipSetDevice(e->dev));
    if (adjoint) {
        const int rc = setup_adjoint(e);
        if (rc) return rc;
    }
    return e->precision == BSA_FP32 ? capture_graph<float>(e, adjoint)
                                    : capture_graph<double>(e, adjoint);
}

static int evaluate_inner(bsa_engine *e, const double *x, double *lnl, bool adjoint);

static int evaluate(bsa_engine *e, const double *x, double *lnl, bool adjoint = false)
{
    auto t0 = std::chrono::steady_clock::now();
    int rc = evaluate_inner(e, x, lnl, adjoint);
    double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rc == BSA_OK) {
        e->timing[6] = ms;
        e->timing_sum[6] += ms;
        e->n_eval += 1;
    }
    return rc;
}

static int evaluate_natural(bsa_engine *e, const double *x, double kappa, double p0, double p1,
                            double w0, double w2, const double *qf_fixed, double *lnl,
                            bool adjoint);

static int evaluate_inner(bsa_engine *e, const double *x, double *lnl, bool adjoint)
{
    double kappa, p0, p1, w0, w2;
    natural_params(e, x, &kappa, &p0, &p1, &w0, &w2);
    return evaluate_natural(e, x, kappa, p0, p1, w0, w2, nullptr, lnl, adjoint);
}

// One evaluation at natural parameters; branch lengths from x.  qf_fixed,
// when given, replaces the two Qfactors the model would give (BEB keeps them
// at the MLE, GetPMatBranch treesub.c:7289-7301).
static int evaluate_natural(bsa_engine *e, const double *x, double kappa, double p0, double p1,
                            double w0, double w2, const double *qf_fixed, double *lnl,
                            bool adjoint)
{
    for (int b = 0; b < e->nbranch; b++)
        if (!std::isfinite(x[b])) return fail(BSA_EINVAL, "branch length %d not finite", b);
    if (!(kappa > 0) || !(w0 >= 0) || !(w2 >= 0) || !(p0 >= 0) || !(p1 >= 0) || !(p0 + p1 > 0) ||
        p0 + p1 > 1 + 1e-12)
        return fail(BSA_EINVAL, "parameters out of range: kappa=%g p0=%g p1=%g w0=%g w2=%g", kappa,
                    p0, p1, w0, w2);
    const bool same = e->have_eigen && kappa == e->cache_kappa && w0 == e->cache_w[0] &&
                      w2 == e->cache_w[2];
    if (same) {
        // eigen systems unchanged: only proportions and Qfactors are recomputed
        if (!setup_model_a(kappa, p0, p1, w0, w2, e->pi.data(), &e->model, false))
            return fail(BSA_ENUMERIC, "model set-up failed");
    } else {
        if (!setup_model_a(kappa, p0, p1, w0, w2, e->pi.data(), &e->model, true)) {
            e->have_eigen = false;
            return fail(BSA_ENUMERIC, "eigen decomposition failed");
        }
        e->have_eigen = true;
        e->cache_kappa = kappa;
        e->cache_w[0] = w0;
        e->cache_w[1] = 1.0;
        e->cache_w[2] = w2;
        // padded U, V, root per eigen set
        for (int s = 0; s < 3; s++) {
            double *U = e->h_uvr + s * kUVRSet, *V = U + 4096, *R = U + 8192;
            std::memset(U, 0, sizeof(double) * kUVRSet);
            for (int i = 0; i < kS; i++)
                for (int k = 0; k < kS; k++) {
                    U[i * 64 + k] = e->model.U[s][i * kS + k];
                    V[i * 64 + k] = e->model.V[s][i * kS + k];
                }
            for (int k = 0; k < kS; k++) R[k] = e->model.root[s][k];
        }
    }
    if (qf_fixed) {
        e->model.Qfactor[0] = qf_fixed[0];
        e->model.Qfactor[1] = qf_fixed[1];
    }
    for (int c = 0; c < 4; c++) e->h_uvr[3 * kUVRSet + c] = e->model.freqK[c];
    for (size_t i = 0; i < e->pjobs.size(); i++) {
        PmatJob p = e->pjobs[i];
        p.t = x[e->ibranch[p.node]] * e->model.Qfactor[e->label[p.node]];
        e->h_pjobs[i] = p;
    }
    if (adjoint)
        for (size_t i = 0; i < e->ejobs.size(); i++) {
            EigJob j = e->ejobs[i];
            const int v = e->ejob_node[i];
            j.tau = x[e->ibranch[v]] * e->model.Qfactor[e->label[v]];
            e->h_ejobs[i] = j;
        }
    int rc = e->precision == BSA_FP32 ? run_device<float>(e, adjoint) : run_device<double>(e, adjoint);
    if (rc) return rc;
    *lnl = e->h_out[0];
    e->last_lnl = *lnl;
    e->evaluated = true;
    return BSA_OK;
}

extern "C" int bsa_lnl(bsa_engine *e, const double *x, int np, double *neg_lnl)
{
    if (!e || !x || !neg_lnl) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    const int rg = prepare_graph(e, false);
    if (rg) return rg;
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->dev));
    double lnl;
    int rc = evaluate(e, x, &lnl);
    if (rc) return rc;
    *neg_lnl = -lnl;
    return BSA_OK;
}

// ---------------------------------------------------------------- adjoint

static int setup_adjoint(bsa_engine *e)
{
    if (e->adj_ready) return BSA_OK;
    const int ns = e->ns, nn = e->nnode;
    std::vector<int> depth(nn, 0), order{e->root};
    for (size_t i = 0; i < order.size(); i++)
        for (int c : e->sons[order[i]]) {
            depth[c] = depth[order[i]] + 1;
            order.push_back(c);
        }
    int maxd = 0;
    for (int v = 0; v < nn; v++) maxd = std::max(maxd, depth[v]);
    auto aslot = [&](int v, int c) { return (v - ns) * 4 + c; };
    for (int c = 0; c < 4; c++) e->root_aslot[c] = aslot(e->root, c);
    // Nodes off the foreground path with a background branch hold one
    // adjoint per merged class pair (0+2, 1+3): their partials and P(t) are
    // the same for both classes, so the two adjoints can be summed.
    auto merged = [&](int v) {
        return v != e->root && !e->onpath[v] && e->label[v] == 0;
    };
    e->adj.clear();
    e->adj_off.assign(1, 0);
    e->adj_nint.clear();
    e->ejobs.clear();
    e->ejob_node.clear();
    std::vector<int> lists[3];
    // contributions of a's children were kept by the forward job (a, c') with
    // c' = c when a has its own class-c job (c < 2 or a on the path), else c - 2
    auto fill_sibs = [&](AdjJob &j, int term, int v, int a, int c) {
        const int cf = (c < 2 || e->onpath[a]) ? c : c - 2;
        int n = 0;
        for (int s2 : e->sons[a]) {
            if (s2 == v) continue;
            j.sib_tip[n] = s2 < ns ? s2 : -1;
            j.sib_slot[term][n] = s2 < ns ? -1 : e->slot[s2][c];
            j.sib_cslot[term][n] = s2 < ns ? -1 : (s2 - ns) * 4 + cf;
            j.sib_pmat[term][n] = s2 * 3 + uvr_set(c, e->label[s2]);
            n++;
        }
        j.nsib = n;
    };
    auto add_eig = [&](int m_index, int set, int v) {
        EigJob ej{};
        ej.m_index = m_index;
        ej.set = set;
        ej.tau = 0;
        lists[set].push_back((int)e->ejobs.size());
        e->ejobs.push_back(ej);
        e->ejob_node.push_back(v);
    };
    for (int d = 1; d <= maxd; d++) {
        // interior children first, so one launch per depth covers exactly them
        std::vector<int> at_d;
        for (int v = ns; v < nn; v++)
            if (depth[v] == d) at_d.push_back(v);
        for (int v = 0; v < ns; v++)
            if (depth[v] == d) at_d.push_back(v);
        size_t nint_jobs = 0;
        for (int v : at_d) {
            if (v >= ns) nint_jobs += merged(v) ? 2 : 4;
            const int a = e->father[v];
            if (merged(v)) {
                for (int m = 0; m < 2; m++) {
                    AdjJob j{};
                    j.v_tip = v < ns ? v : -1;
                    j.v_slot = v < ns ? -1 : e->slot[v][m];
                    j.v_pmat = v * 3 + uvr_set(m, 0);
                    j.v_aslot = v < ns ? -1 : aslot(v, m);
                    j.m_index = (int)e->adj.size();
                    if (merged(a)) {
                        j.nterm = 1;
                        j.a_aslot[0] = aslot(a, m);
                        fill_sibs(j, 0, v, a, m);
                    } else {
                        j.nterm = 2;
                        for (int t = 0; t < 2; t++) {
                            j.a_aslot[t] = aslot(a, m + 2 * t);
                            fill_sibs(j, t, v, a, m + 2 * t);
                        }
                    }
                    e->adj.push_back(j);
                    add_eig(j.m_index, uvr_set(m, 0), v);
                }
            } else {
                for (int c = 0; c < 4; c++) {
                    AdjJob j{};
                    j.v_tip = v < ns ? v : -1;
                    j.v_slot = v < ns ? -1 : e->slot[v][c];
                    j.v_pmat = v * 3 + uvr_set(c, e->label[v]);
                    j.v_aslot = v < ns ? -1 : aslot(v, c);
                    j.m_index = (int)e->adj.size();
                    j.nterm = 1;
                    j.a_aslot[0] = aslot(a, c);   // the parent of an unmerged node is unmerged
                    fill_sibs(j, 0, v, a, c);
                    e->adj.push_back(j);
                    add_eig(j.m_index, uvr_set(c, e->label[v]), v);
                }
            }
        }
        e->adj_off.push_back((int)e->adj.size());
        e->adj_nint.push_back((int)nint_jobs);
    }
    e->set_off = {0};
    e->set_list.clear();
    for (int s = 0; s < 3; s++) {
        e->set_list.insert(e->set_list.end(), lists[s].begin(), lists[s].end());
        e->set_off.push_back((int)e->set_list.size());
    }
    e->nmrow = (int)e->adj.size();   // one branch matrix per job
    const int ntile = e->npad / 64;
    // tiles per branch-matrix block: up to kAdjTilesPerChunk, fewer when the
    // gene is small, so that the launch still has about 1024 blocks
    e->tpc = std::max(1, std::min(kAdjTilesPerChunk, ntile * (int)e->adj.size() / 1024));
    e->nchunk = (ntile + e->tpc - 1) / e->tpc;
    const size_t ts = tsize(e), nej = e->ejobs.size();
    const size_t nasl = (size_t)(nn - ns) * 4;
    HIPCHK(hipSetDevice(e->dev));
    HIPCHK(hipMalloc(&e->d_adj, sizeof(AdjJob) * e->adj.size()));
    HIPCHK(hipMemcpy(e->d_adj, e->adj.data(), sizeof(AdjJob) * e->adj.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&e->d_ejobs, sizeof(EigJob) * nej));
    HIPCHK(hipHostMalloc((void **)&e->h_ejobs, sizeof(EigJob) * nej, hipHostMallocDefault));
    HIPCHK(hipMalloc(&e->d_set_off, sizeof(int) * 4));
    HIPCHK(hipMemcpy(e->d_set_off, e->set_off.data(), sizeof(int) * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&e->d_set_list, sizeof(int) * std::max<size_t>(1, e->set_list.size())));
    HIPCHK(hipMemcpy(e->d_set_list, e->set_list.data(), sizeof(int) * e->set_list.size(),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&e->d_A, ts * 64 * (size_t)e->npad * nasl));
    HIPCHK(hipMalloc(&e->d_C, ts * 64 * (size_t)e->npad * nasl));
    HIPCHK(hipMalloc(&e->d_EA, sizeof(int) * (size_t)e->npad * nasl));
    HIPCHK(hipMalloc(&e->d_Mpart, sizeof(double) * 4096 * (size_t)e->nmrow * e->nchunk));
    HIPCHK(hipMemset(e->d_Mpart, 0, sizeof(double) * 4096 * (size_t)e->nmrow * e->nchunk));
    HIPCHK(hipMalloc(&e->d_W, sizeof(double) * 4096 * 3 * kWParts));
    HIPCHK(hipMalloc(&e->d_G, sizeof(double) * 4096 * nej));
    HIPCHK(hipMalloc(&e->d_dtau, sizeof(double) * nej));
    HIPCHK(hipMalloc(&e->d_Z, sizeof(double) * 3 * 4096));
    std::vector<double> m64((size_t)e->ncodes * 64, 0.0);
    for (int c = 0; c < e->ncodes; c++)
        for (int k = 0; k < kS; k++) m64[(size_t)c * 64 + k] = e->code_mask[(size_t)c * kS + k];
    HIPCHK(hipMalloc(&e->d_mask64, sizeof(double) * m64.size()));
    HIPCHK(hipMemcpy(e->d_mask64, m64.data(), sizeof(double) * m64.size(), hipMemcpyHostToDevice));
    HIPCHK(hipHostMalloc((void **)&e->h_adjout, sizeof(double) * (nej + 3 * 4096), hipHostMallocDefault));
    // work of one gradient on 61 states, per kernel.  Child adjoints (interior
    // children only): one 61 x 61 contraction per pattern; bytes: the
    // parent's adjoint per term, the siblings' kept contributions (interior)
    // or codes (tips), exponents, the child's adjoint and exponent written.
    // Branch matrices (every job): one 61 x 61 outer-product accumulation per
    // pattern; bytes: the same X operands, plus the child's partials and
    // exponent (interior) or code (tip).  The contributions C kept by the
    // forward pass are the siblings' partials times P, so no contraction is
    // repeated for them.
    {
        const double w = (double)ts, np_ = e->npatt;
        double fa = 0, ba = 0, fm = 0, bm = 0;
        for (const AdjJob &j : e->adj) {
            int nint = 0;
            for (int q = 0; q < j.nsib; q++) nint += j.sib_tip[q] < 0;
            const bool vint = j.v_tip < 0;
            const int nt = j.nterm;
            const double xbytes = kS * w * nt * (1 + nint) + 4.0 * nt * (1 + nint) +
                                  (double)nt * (j.nsib - nint);
            if (vint) {
                fa += 2.0 * kS * kS * np_;
                ba += np_ * (xbytes + kS * w + 4.0);
            }
            fm += 2.0 * kS * kS * np_;
            bm += np_ * (xbytes + (vint ? kS * w + 4.0 : 1.0));
        }
        e->work[6] = fa;
        e->work[7] = ba;
        e->work[8] = fm;
        e->work[9] = bm;
    }
    e->adj_ready = true;
    return BSA_OK;
}

// -lnL and d(-lnL)/dx by one forward pass and one adjoint (preorder) sweep
extern "C" int bsa_lnl_grad(bsa_engine *e, const double *x, int np, double *neg_lnl, double *grad)
{
    if (!e || !x || !neg_lnl || !grad) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    const int rg = prepare_graph(e, true);
    if (rg) return rg;
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->dev));
    int rc = setup_adjoint(e);
    if (rc) return rc;
    double lnl;
    if ((rc = evaluate(e, x, &lnl, true))) return rc;
    const size_t nej = e->ejobs.size();
    const double *dtau = e->h_adjout, *Z = e->h_adjout + nej;
    std::vector<double> dt_node(e->nnode, 0.0);
    for (size_t i = 0; i < nej; i++) dt_node[e->ejob_node[i]] += dtau[i];
    double dQf[2] = {0, 0};
    const int nb = e->nbranch;
    for (int v = 0; v < e->nnode; v++) {
        if (v == e->root) continue;
        const int b = e->ibranch[v], l = e->label[v];
        dQf[l] += x[b] * dt_node[v];
        grad[b] = -e->model.Qfactor[l] * dt_node[v];
    }
    const double Sf[4] = {e->h_out[1], e->h_out[2], e->h_out[3], e->h_out[4]};
    ModelGrad mg;
    model_a_param_grad(e->model, e->pi.data(), Z, dQf, Sf, &mg);
    grad[nb] = -mg.kappa;
    if (e->opts.p_transformed) {
        const double p0 = e->model.p0, p1 = e->model.p1;
        grad[nb + 1] = -(mg.p0 * p0 * (1 - p0) - mg.p1 * p0 * p1);
        grad[nb + 2] = -(-mg.p0 * p0 * p1 + mg.p1 * p1 * (1 - p1));
    } else {
        grad[nb + 1] = -mg.p0;
        grad[nb + 2] = -mg.p1;
    }
    grad[nb + 3] = -mg.w0;
    if (!e->opts.fix_omega) grad[nb + 4] = -mg.w2;
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
    const int rg = prepare_graph(e, false);
    if (rg) return rg;
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->dev));
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
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() { if (p) hipFree(p); }
};
}  // namespace

extern "C" int bsa_beb(bsa_engine *e, const double *x, int np, double *post, double *ml3,
                       double *postw20, double *postp0p1, double *lnf_sets)
{
    if (!e || !x || !post) return fail(BSA_EINVAL, "null argument");
    if (np != bsa_num_params(e)) return fail(BSA_EINVAL, "np=%d, engine expects %d", np, bsa_num_params(e));
    const int rg = prepare_graph(e, false);
    if (rg) return rg;
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->dev));
    constexpr int n1d = 10, dim = 4, ngrid = 10000, nset = n1d + 1 + n1d * n1d + n1d;
    const int npatt = e->npatt, npad = e->npad;
    double kappa, p0, p1, w0, w2;
    natural_params(e, x, &kappa, &p0, &p1, &w0, &w2);
    ModelA mle;
    if (!setup_model_a(kappa, p0, p1, w0, w2, e->pi.data(), &mle, false))
        return fail(BSA_EINVAL, "parameters out of range for BEB");
    const double qf[2] = {mle.Qfactor[0], mle.Qfactor[1]};
    double w0g[n1d], w2g[n1d];
    for (int i = 0; i < n1d; i++) {
        w0g[i] = (i + 0.5) / n1d;
        w2g[i] = 1.0 + (i + 0.5) * (11.0 - 1.0) / n1d;
    }
    DevBuf F, mx, pc, iw, lnfx, gl, W, dpost;
    HIPCHK(hipMalloc(&F.p, sizeof(double) * nset * (size_t)npad));
    hipStream_t s = e->stream;
    double *dF = (double *)F.p;
    auto row = [&](int set) { return dF + (size_t)set * npad; };
    for (int i = 0; i < n1d; i++)
        for (int j = 0; j < n1d; j++) {
            double l;
            int rc = evaluate_natural(e, x, kappa, p0, p1, w0g[i], w2g[j], qf, &l, false);
            if (rc) return rc;
            const size_t rb = sizeof(double) * npad;
            const double *fh = e->d_fh;
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
        const double v = e->fpatt[h] * hmx[h], t = S1 + v;
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
    launch_beb_lnfx(s, dF, npad, npatt, e->d_fpatt, (const double *)pc.p, (const int *)iw.p, ngrid,
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
    if (!e->evaluated) return fail(BSA_EINVAL, "no evaluation yet");
    HIP_SHARED;
    HIPCHK(hipSetDevice(e->dev));
    if (fhK)
        for (int c = 0; c < 4; c++)
            HIPCHK(hipMemcpy(fhK + (size_t)c * e->npatt, e->d_fh + (size_t)c * e->npad,
                             sizeof(double) * e->npatt, hipMemcpyDeviceToHost));
    if (site)
        HIPCHK(hipMemcpy(site, e->d_site, sizeof(double) * e->npatt, hipMemcpyDeviceToHost));
    return BSA_OK;
}

extern "C" int bsa_last_params(const bsa_engine *e, double *o)
{
    if (!e || !o) return fail(BSA_EINVAL, "null argument");
    if (!e->evaluated) return fail(BSA_EINVAL, "no evaluation yet");
    const ModelA &m = e->model;
    double v[11] = {m.kappa, m.p0, m.p1, m.w0, m.w2, m.Qfactor[0], m.Qfactor[1],
                    m.freqK[0], m.freqK[1], m.freqK[2], m.freqK[3]};
    std::memcpy(o, v, sizeof v);
    return BSA_OK;
}

extern "C" int bsa_timing(bsa_engine *e, double *o, int reset)
{
    if (!e || !o) return fail(BSA_EINVAL, "null argument");
    std::memcpy(o, e->timing, sizeof e->timing);
    std::memcpy(o + 7, e->timing_sum, sizeof e->timing_sum);
    o[14] = e->n_eval;
    o[15] = e->n_launch;
    o[16] = e->n_adj_launch;
    if (reset) {
        for (double &v : e->timing_sum) v = 0;
        e->n_eval = 0;
        e->n_launch = 0;
        e->n_adj_launch = 0;
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
    int rc = setup_adjoint(e);
    if (rc) return rc;
    std::memcpy(o, e->work, sizeof e->work);
    return BSA_OK;
}

extern "C" int bsa_set_graph(bsa_engine *e, int enable)
{
    if (!e) return fail(BSA_EINVAL, "null engine");
    e->use_graph = enable != 0;
    return BSA_OK;
}

extern "C" int bsa_default_bounds(const bsa_engine *e, double *lo, double *hi)
{
    if (!e || !lo || !hi) return fail(BSA_EINVAL, "null argument");
    const int nb = e->nbranch;
    for (int b = 0; b < nb; b++) {
        lo[b] = 4e-6;
        hi[b] = 50;
    }
    lo[nb] = 1e-4;   // kappa
    hi[nb] = 999;
    for (int k = 1; k <= 2; k++) {
        lo[nb + k] = e->opts.p_transformed ? -99 : 1e-6;
        hi[nb + k] = e->opts.p_transformed ? 99 : 1 - 1e-6;
    }
    lo[nb + 3] = 1e-6;   // w0 < 1
    hi[nb + 3] = 1;
    if (!e->opts.fix_omega) {   // w2 > 1
        lo[nb + 4] = 1;
        hi[nb + 4] = 999;
    }
    return BSA_OK;
}

int bsa_error(int code, const char *msg) { return fail(code, "%s", msg); }

extern "C" const char *bsa_last_error(void) { return g_err.c_str(); }
extern "C" const char *bsa_version(void) { return "bsa 0.1 gfx950"; }

// In-bounds check of the matrix-core lane layout: runs one 16x16x4 MFMA in
// FP64 and FP32 on A (16x4) and B (4x16) and returns each lane's 4 outputs.
extern "C" int bsa_debug_mfma_probe(int device, const double *A, const double *B, double *D64,
                                    double *D32)
{
    HIP_SHARED;
    HIPCHK(hipSetDevice(device));
    double *dA, *dB, *dD;
    float *fA, *fB, *fD;
    HIPCHK(hipMalloc(&dA, 64 * 8));
    HIPCHK(hipMalloc(&dB, 64 * 8));
    HIPCHK(hipMalloc(&dD, 256 * 8));
    HIPCHK(hipMalloc(&fA, 64 * 4));
    HIPCHK(hipMalloc(&fB, 64 * 4));
    HIPCHK(hipMalloc(&fD, 256 * 4));
    float hA[64], hB[64], hD[256];
    for (int i = 0; i < 64; i++) { hA[i] = (float)A[i]; hB[i] = (float)B[i]; }
    HIPCHK(hipMemcpy(dA, A, 64 * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dB, B, 64 * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(fA, hA, 64 * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(fB, hB, 64 * 4, hipMemcpyHostToDevice));
    launch_mfma_probe(nullptr, dA, dB, dD, fA, fB, fD);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    HIPCHK(hipMemcpy(D64, dD, 256 * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(hD, fD, 256 * 4, hipMemcpyDeviceToHost));
    for (int i = 0; i < 256; i++) D32[i] = hD[i];
    hipFree(dA); hipFree(dB); hipFree(dD); hipFree(fA); hipFree(fB); hipFree(fD);
    return BSA_OK;
}
