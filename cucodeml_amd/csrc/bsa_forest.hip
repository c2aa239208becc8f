// Genes and the forest that evaluates them: per-gene schedules (pruning by
// tree level, adjoint sweep by depth), their renumbering into shared device
// buffers, the launch sequence of one evaluation, the host model set-up
// before it and the chain rule after it.
#include "bsa_forest.h"

#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

namespace bsa {

static thread_local std::string g_err;
std::shared_mutex g_hip_mu;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char *last_error_text() { return g_err.c_str(); }

// ---------------------------------------------------------------- one gene

static int build_schedule(Gene &g, int wbytes)
{
    const int ns = g.ns, nn = g.nnode;
    // heights (tips 0), levels of interior nodes
    g.height.assign(nn, 0);
    std::vector<int> order;   // postorder
    {
        std::vector<std::pair<int, int>> st{{g.root, 0}};
        while (!st.empty()) {
            auto &[v, k] = st.back();
            if (k < (int)g.sons[v].size()) {
                int c = g.sons[v][k++];
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
        for (int c : g.sons[v]) h = std::max(h, g.height[c] + 1);
        g.height[v] = (v < ns) ? 0 : h;
    }
    // nodes whose site classes 2/3 differ from 0/1: proper ancestors of a
    // foreground branch
    g.onpath.assign(nn, 0);
    int nfg = 0;
    for (int v = 0; v < nn; v++) {
        if (v == g.root || g.label[v] != 1) continue;
        nfg++;
        for (int a = g.father[v]; a >= 0; a = g.father[a]) g.onpath[a] = 1;
    }
    if (nfg == 0) return fail(BSA_EINVAL, "branch-site model A needs a foreground (#1) branch");
    // partial slots
    g.slot.assign(nn, std::array<int, 4>{-1, -1, -1, -1});
    int ns_ = 0;
    for (int v = ns; v < nn; v++) {
        g.slot[v][0] = ns_++;
        g.slot[v][1] = ns_++;
        if (g.onpath[v]) {
            g.slot[v][2] = ns_++;
            g.slot[v][3] = ns_++;
        } else {
            g.slot[v][2] = g.slot[v][0];
            g.slot[v][3] = g.slot[v][1];
        }
    }
    g.nslot = ns_;
    for (int c = 0; c < 4; c++) g.root_slot[c] = g.slot[g.root][c];
    // pruning jobs grouped by height
    const int hmax = g.height[g.root];
    g.prune.clear();
    g.level_off.assign(1, 0);
    for (int lev = 1; lev <= hmax; lev++) {
        for (int v = ns; v < nn; v++) {
            if (g.height[v] != lev) continue;
            const int ncls = g.onpath[v] ? 4 : 2;
            for (int c = 0; c < ncls; c++) {
                PruneJob j{};
                j.out_slot = g.slot[v][c];
                j.nch = (int)g.sons[v].size();
                for (int k = 0; k < j.nch; k++) {
                    int s = g.sons[v][k];
                    j.ch_tip[k] = s < ns ? s : -1;
                    j.ch_slot[k] = s < ns ? -1 : g.slot[s][c];
                    j.ch_pmat[k] = s * 3 + uvr_set(c, g.label[s]);
                    j.ch_cslot[k] = s < ns ? -1 : (s - ns) * 4 + c;
                }
                g.prune.push_back(j);
            }
        }
        g.level_off.push_back((int)g.prune.size());
    }
    // work per evaluation of the pruning pass, counted on 61 states
    {
        const double w = wbytes, np_ = g.npatt;
        double fl = 0, by = 0;
        for (const PruneJob &j : g.prune) {
            int nint = 0, ntip = 0;
            for (int k = 0; k < j.nch; k++) (j.ch_tip[k] >= 0 ? ntip : nint)++;
            fl += 2.0 * kS * kS * np_ * nint;
            by += np_ * (kS * w * (nint + 1) + ntip + 4.0 * (nint + 1));
        }
        int nint_nodes = nn - ns, nint_edges = 0;
        for (int v = ns; v < nn; v++)
            if (v != g.root) nint_edges++;
        g.work[0] = fl;
        g.work[1] = by;
        g.work[2] = 4.0 * np_ * nint_edges * 2.0 * kS * kS;
        g.work[3] = 4.0 * np_ * kS * w * (nint_nodes + nint_edges);
        g.work[4] = (double)(g.level_off.size() - 1);
        g.work[5] = (double)g.prune.size();
    }
    // P(t) jobs: every branch for eigen sets w0 and 1; foreground also w2
    g.pjobs.clear();
    for (int v = 0; v < nn; v++) {
        if (v == g.root) continue;
        const int nset = g.label[v] == 1 ? 3 : 2;
        for (int s = 0; s < nset; s++) {
            PmatJob p{};
            p.node = v;
            p.set = s;
            p.pm = v * 3 + s;
            p.tip = v < ns ? v : -1;
            p.amb = 0;
            p.t = 0;
            g.pjobs.push_back(p);
        }
    }
    return BSA_OK;
}

int gene_init(Gene &g, const bsa_problem *p, int wbytes)
{
    if (p->ns < 3) return fail(BSA_EINVAL, "need at least 3 sequences (ns=%d)", p->ns);
    if (p->nnode <= p->ns || p->nnode > 2 * p->ns - 1)
        return fail(BSA_EINVAL, "nnode=%d inconsistent with ns=%d", p->nnode, p->ns);
    if (p->root < p->ns || p->root >= p->nnode) return fail(BSA_EINVAL, "root must be interior");
    if (p->npatt <= 0) return fail(BSA_EINVAL, "npatt must be positive");
    if (p->ncodes < kS || p->ncodes > 256) return fail(BSA_EINVAL, "ncodes=%d out of range", p->ncodes);
    if (!p->father || !p->label || !p->ibranch || !p->z || !p->fpatt || !p->pi)
        return fail(BSA_EINVAL, "null array in problem");
    if (p->ncodes > kS && !p->code_mask) return fail(BSA_EINVAL, "ambiguity codes need code_mask");
    g.ns = p->ns;
    g.nnode = p->nnode;
    g.root = p->root;
    g.nbranch = p->nnode - 1;
    g.npatt = p->npatt;
    g.ncodes = p->ncodes;
    g.namb = p->ncodes - kS;
    g.father.assign(p->father, p->father + p->nnode);
    g.label.assign(p->label, p->label + p->nnode);
    g.ibranch.assign(p->ibranch, p->ibranch + p->nnode);
    g.sons.assign(p->nnode, {});
    std::vector<char> seen_b(g.nbranch, 0);
    for (int v = 0; v < p->nnode; v++) {
        const int f = g.father[v];
        if (v == p->root) {
            if (f != -1) return fail(BSA_EINVAL, "father[root] must be -1");
            continue;
        }
        if (f < p->ns || f >= p->nnode) return fail(BSA_EINVAL, "node %d has invalid father %d", v, f);
        if (g.label[v] != 0 && g.label[v] != 1)
            return fail(BSA_EINVAL, "branch label of node %d is %d; model A allows 0 and 1", v,
                        g.label[v]);
        const int b = g.ibranch[v];
        if (b < 0 || b >= g.nbranch || seen_b[b])
            return fail(BSA_EINVAL, "ibranch of node %d (%d) is not a permutation", v, b);
        seen_b[b] = 1;
        g.sons[f].push_back(v);
    }
    for (int v = p->ns; v < p->nnode; v++) {
        if (g.sons[v].empty()) return fail(BSA_EINVAL, "interior node %d has no sons", v);
        if ((int)g.sons[v].size() > kMaxCh)
            return fail(BSA_EINVAL, "node %d has %zu sons (max %d)", v, g.sons[v].size(), kMaxCh);
    }
    for (size_t i = 0; i < (size_t)p->ns * p->npatt; i++)
        if (p->z[i] >= p->ncodes) return fail(BSA_EINVAL, "code %d >= ncodes", p->z[i]);
    g.z.assign(p->z, p->z + (size_t)p->ns * p->npatt);
    g.pi.assign(p->pi, p->pi + kS);
    g.fpatt.assign(p->fpatt, p->fpatt + p->npatt);
    g.code_mask.assign((size_t)p->ncodes * kS, 0);
    for (int c = 0; c < kS; c++) g.code_mask[(size_t)c * kS + c] = 1;
    for (size_t i = (size_t)kS * kS; i < (size_t)p->ncodes * kS; i++)
        g.code_mask[i] = p->code_mask[i] ? 1 : 0;
    return build_schedule(g, wbytes);
}

int gene_num_params(const Gene &g) { return g.nbranch + (g.opts.fix_omega ? 4 : 5); }

void natural_params(const Gene &g, const double *x, double *kappa, double *p0, double *p1,
                    double *w0, double *w2)
{
    const int nb = g.nbranch;
    *kappa = x[nb];
    if (g.opts.p_transformed) {
        const double a = std::exp(x[nb + 1]), b = std::exp(x[nb + 2]), tot = 1 + a + b;
        *p0 = a / tot;
        *p1 = b / tot;
    } else {
        *p0 = x[nb + 1];
        *p1 = x[nb + 2];
    }
    *w0 = x[nb + 3];
    *w2 = g.opts.fix_omega ? g.opts.omega_fix : x[nb + 4];
}

void default_bounds(const Gene &g, double *lo, double *hi)
{
    const int nb = g.nbranch;
    for (int b = 0; b < nb; b++) {
        lo[b] = 4e-6;
        hi[b] = 50;
    }
    lo[nb] = 1e-4;   // kappa
    hi[nb] = 999;
    for (int k = 1; k <= 2; k++) {
        lo[nb + k] = g.opts.p_transformed ? -99 : 1e-6;
        hi[nb + k] = g.opts.p_transformed ? 99 : 1 - 1e-6;
    }
    lo[nb + 3] = 1e-6;   // w0 < 1
    hi[nb + 3] = 1;
    if (!g.opts.fix_omega) {   // w2 > 1
        lo[nb + 4] = 1;
        hi[nb + 4] = 999;
    }
}

void gene_plan_adjoint(Gene &g, int wbytes)
{
    if (g.adj_planned) return;
    const int ns = g.ns, nn = g.nnode;
    std::vector<int> depth(nn, 0), order{g.root};
    for (size_t i = 0; i < order.size(); i++)
        for (int c : g.sons[order[i]]) {
            depth[c] = depth[order[i]] + 1;
            order.push_back(c);
        }
    int maxd = 0;
    for (int v = 0; v < nn; v++) maxd = std::max(maxd, depth[v]);
    auto aslot = [&](int v, int c) { return (v - ns) * 4 + c; };
    for (int c = 0; c < 4; c++) g.root_aslot[c] = aslot(g.root, c);
    // Nodes off the foreground path with a background branch hold one
    // adjoint per merged class pair (0+2, 1+3): their partials and P(t) are
    // the same for both classes, so the two adjoints can be summed.
    auto merged = [&](int v) { return v != g.root && !g.onpath[v] && g.label[v] == 0; };
    g.adj.clear();
    g.adj_off.assign(1, 0);
    g.adj_nint.clear();
    g.ejobs.clear();
    g.ejob_node.clear();
    // contributions of a's children were kept by the forward job (a, c') with
    // c' = c when a has its own class-c job (c < 2 or a on the path), else c - 2
    auto fill_sibs = [&](AdjJob &j, int term, int v, int a, int c) {
        const int cf = (c < 2 || g.onpath[a]) ? c : c - 2;
        int n = 0;
        for (int s2 : g.sons[a]) {
            if (s2 == v) continue;
            j.sib_tip[n] = s2 < ns ? s2 : -1;
            j.sib_slot[term][n] = s2 < ns ? -1 : g.slot[s2][c];
            j.sib_cslot[term][n] = s2 < ns ? -1 : (s2 - ns) * 4 + cf;
            j.sib_pmat[term][n] = s2 * 3 + uvr_set(c, g.label[s2]);
            n++;
        }
        j.nsib = n;
    };
    auto add_eig = [&](int m_index, int set, int v) {
        EigJob ej{};
        ej.m_index = m_index;
        ej.set = set;
        ej.tau = 0;
        g.ejobs.push_back(ej);
        g.ejob_node.push_back(v);
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
            const int a = g.father[v];
            if (merged(v)) {
                for (int m = 0; m < 2; m++) {
                    AdjJob j{};
                    j.v_tip = v < ns ? v : -1;
                    j.v_mask = 0;
                    j.v_slot = v < ns ? -1 : g.slot[v][m];
                    j.v_pmat = v * 3 + uvr_set(m, 0);
                    j.v_aslot = v < ns ? -1 : aslot(v, m);
                    j.m_index = (int)g.adj.size();
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
                    g.adj.push_back(j);
                    add_eig(j.m_index, uvr_set(m, 0), v);
                }
            } else {
                for (int c = 0; c < 4; c++) {
                    AdjJob j{};
                    j.v_tip = v < ns ? v : -1;
                    j.v_mask = 0;
                    j.v_slot = v < ns ? -1 : g.slot[v][c];
                    j.v_pmat = v * 3 + uvr_set(c, g.label[v]);
                    j.v_aslot = v < ns ? -1 : aslot(v, c);
                    j.m_index = (int)g.adj.size();
                    j.nterm = 1;
                    j.a_aslot[0] = aslot(a, c);   // the parent of an unmerged node is unmerged
                    fill_sibs(j, 0, v, a, c);
                    g.adj.push_back(j);
                    add_eig(j.m_index, uvr_set(c, g.label[v]), v);
                }
            }
        }
        g.adj_off.push_back((int)g.adj.size());
        g.adj_nint.push_back((int)nint_jobs);
    }
    // work of one gradient on 61 states, per kernel.  Child adjoints (interior
    // children only): one 61 x 61 contraction per pattern; bytes: the
    // parent's adjoint per term, the siblings' kept contributions (interior)
    // or codes (tips), exponents, the child's adjoint and exponent written.
    // Branch matrices (every job): for an interior child one 61 x 61
    // outer-product accumulation per pattern; for a tip, whose partials are
    // code indicators, one row of X added into the column of the tip's codon
    // per pattern.  Bytes: the same X operands, plus the child's partials and
    // exponent (interior) or code and place in the tip's pattern order (tip).
    // The contributions C kept by the
    // forward pass are the siblings' partials times P, so no contraction is
    // repeated for them.
    const double w = wbytes, np_ = g.npatt;
    double fa = 0, ba = 0, fm = 0, bm = 0;
    for (const AdjJob &j : g.adj) {
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
        fm += vint ? 2.0 * kS * kS * np_ : (double)kS * np_;
        bm += np_ * (xbytes + (vint ? kS * w + 4.0 : 1.0 + 4.0));
    }
    g.work[6] = fa;
    g.work[7] = ba;
    g.work[8] = fm;
    g.work[9] = bm;
    g.adj_planned = true;
}

// ---------------------------------------------------------------- forest

static size_t tsize(const Forest &f) { return f.precision == BSA_FP32 ? 4 : 8; }

// forest numbering of gene g's node v: the tips of all genes first (tip
// table rows), then the interior nodes
static int gnode(const Forest &f, const Gene &g, int v)
{
    return v < g.ns ? g.tip_base + v : f.ns_total + g.int_base + (v - g.ns);
}
static int gpm(const Forest &f, const Gene &g, int pm) { return gnode(f, g, pm / 3) * 3 + pm % 3; }
static int gslot(const Gene &g, int s) { return s < 0 ? -1 : g.slot_base + s; }
static int gaslot(const Gene &g, int s) { return s < 0 ? -1 : g.int_base * 4 + s; }
static int gtip(const Gene &g, int t) { return t < 0 ? -1 : g.tip_base + t; }

void forest_free(Forest &f)
{
    // best-effort release: errors cannot be reported here and are ignored
    if (f.stream) (void)hipSetDevice(f.dev);
    void *dev_ptrs[] = {f.d_uvr, f.d_pjobs, f.d_prune, f.d_wjobs, f.d_woff, f.d_witems, f.d_rargs, f.d_PT, f.d_AT, f.d_L, f.d_E,
                        f.d_z, f.d_mask_amb, f.d_pi, f.d_fpatt, f.d_fh, f.d_site, f.d_partial,
                        f.d_out, f.d_adj, f.d_ejobs, f.d_set_off, f.d_set_list, f.d_A, f.d_C,
                        f.d_EA, f.d_Mpart, f.d_W, f.d_G, f.d_dtau, f.d_Z, f.d_mask64, f.d_tperm,
                        f.d_sel};
    for (void *p : dev_ptrs)
        if (p) (void)hipFree(p);
    void *host_ptrs[] = {f.h_uvr, f.h_out, f.h_pjobs, f.h_ejobs, f.h_adjout, f.h_sel};
    for (void *p : host_ptrs)
        if (p) (void)hipHostFree(p);
    for (auto &v : f.ev)
        if (v) (void)hipEventDestroy(v);
    for (auto &gr : f.graph)
        if (gr) (void)hipGraphExecDestroy(gr);
    if (f.stream) (void)hipStreamDestroy(f.stream);
    f = Forest();
}

int forest_plan(Forest &f, const std::vector<Gene *> &genes, int precision, int kernel, int tile)
{
    if (genes.empty()) return fail(BSA_EINVAL, "no genes");
    if (precision != BSA_FP64 && precision != BSA_FP32)
        return fail(BSA_EINVAL, "unknown precision %d", precision);
    if (kernel != BSA_KERNEL_MFMA && kernel != BSA_KERNEL_VALU && kernel != BSA_KERNEL_MFMA32 &&
        kernel != BSA_KERNEL_FP32MC)
        return fail(BSA_EINVAL, "unknown kernel %d", kernel);
    if ((kernel == BSA_KERNEL_MFMA32 || kernel == BSA_KERNEL_FP32MC) && precision != BSA_FP32)
        return fail(BSA_EINVAL, "the FP32 matrix-core kernels need FP32 partials (BSA_FP32)");
    tile = tile == 0 ? 64 : tile;
    if (tile != 64 && tile != 128) return fail(BSA_EINVAL, "tile %d: 64 or 128 patterns per block", tile);
    f.precision = precision;
    f.kernel = kernel;
    f.tile = tile;
    f.genes = genes;
    f.G = (int)genes.size();
    int tip = 0, nint = 0, slot = 0, pj = 0, amb = 0, mask = 0, maxpatt = 0, maxcodes = kS;
    size_t nlev = 0;
    for (int i = 0; i < f.G; i++) {
        Gene &g = *genes[i];
        g.gi = i;
        g.tip_base = tip;
        g.int_base = nint;
        g.slot_base = slot;
        g.pj_base = pj;
        g.amb_base = amb;
        g.mask_base = mask;
        tip += g.ns;
        nint += g.nnode - g.ns;
        slot += g.nslot;
        pj += (int)g.pjobs.size();
        amb += g.namb;
        mask += g.ncodes;
        maxpatt = std::max(maxpatt, g.npatt);
        maxcodes = std::max(maxcodes, g.ncodes);
        nlev = std::max(nlev, g.level_off.size() - 1);
    }
    f.ns_total = tip;
    f.nint_total = nint;
    f.nslot = slot;
    f.ncodes = maxcodes;
    // the tip-table kernel builds rows for every code of the forest: a gene
    // with fewer ambiguity codes reads (and never uses) rows past its own
    f.namb_rows = amb > 0 ? amb + (maxcodes - kS) : 0;
    f.nmask_rows = mask;
    f.npad = (maxpatt + kPadP - 1) / kPadP * kPadP;
    // renumbered schedules
    f.pjobs.clear();
    for (Gene *gp : genes) {
        const Gene &g = *gp;
        for (const PmatJob &p : g.pjobs) {
            PmatJob q = p;
            q.node = gnode(f, g, p.node);
            q.set = 3 * g.gi + p.set;
            q.pm = gpm(f, g, p.pm);
            q.tip = gtip(g, p.tip);
            q.amb = g.amb_base;
            f.pjobs.push_back(q);
        }
    }
    f.prune.clear();
    f.level_off.assign(1, 0);
    f.level_maxint.clear();
    for (Gene *gp : genes) gp->prune_rng.assign(nlev, {0, 0});
    for (size_t l = 0; l < nlev; l++) {
        for (Gene *gp : genes) {
            Gene &g = *gp;
            if (l + 1 >= g.level_off.size()) continue;
            g.prune_rng[l][0] = (int)f.prune.size();
            g.prune_rng[l][1] = (int)f.prune.size() + g.level_off[l + 1] - g.level_off[l];
            for (int i = g.level_off[l]; i < g.level_off[l + 1]; i++) {
                PruneJob j = g.prune[i];
                j.out_slot = gslot(g, j.out_slot);
                for (int k = 0; k < j.nch; k++) {
                    j.ch_tip[k] = gtip(g, j.ch_tip[k]);
                    j.ch_slot[k] = gslot(g, j.ch_slot[k]);
                    j.ch_pmat[k] = gpm(f, g, j.ch_pmat[k]);
                    j.ch_cslot[k] = gaslot(g, j.ch_cslot[k]);
                }
                f.prune.push_back(j);
            }
        }
        int maxint = 0;
        for (int i = f.level_off.back(); i < (int)f.prune.size(); i++) {
            int n = 0;
            for (int k = 0; k < f.prune[i].nch; k++) n += f.prune[i].ch_tip[k] < 0;
            maxint = std::max(maxint, n);
        }
        f.level_maxint.push_back(maxint);
        f.level_off.push_back((int)f.prune.size());
    }
    // persistent pruning: each gene's jobs in level order, genes by falling
    // job count (the largest first, so the static round-robin of items over
    // blocks ends evenly), every tile of a gene
    f.wjobs.clear();
    f.woff.assign(1, 0);
    for (Gene *gp : genes) {
        for (const auto &rg : gp->prune_rng)
            for (int i = rg[0]; i < rg[1]; i++) f.wjobs.push_back(i);
        f.woff.push_back((int)f.wjobs.size());
    }
    f.walk_order.resize(f.G);
    for (int i = 0; i < f.G; i++) f.walk_order[i] = i;
    std::stable_sort(f.walk_order.begin(), f.walk_order.end(), [&](int a, int b) {
        return f.woff[a + 1] - f.woff[a] > f.woff[b + 1] - f.woff[b];
    });
    const int wt = f.npad / kPadP;
    f.witems.clear();
    for (int gi : f.walk_order)
        for (int t = 0; t < wt; t++) f.witems.push_back(gi * wt + t);
    for (double &v : f.work) v = 0;
    for (Gene *gp : genes)
        for (int k = 0; k < 4; k++) f.work[k] += gp->work[k];
    f.work[4] = (double)nlev;
    f.work[5] = (double)f.prune.size();
    return BSA_OK;
}

// Host images of a planned forest's static device inputs: tip codes,
// ambiguity masks, pi, pattern counts and the root jobs.
struct StaticInputs {
    std::vector<unsigned char> z, m;
    std::vector<double> pi64, fp;
    std::vector<RootArgs> ra;
};

static void static_inputs(const Forest &f, StaticInputs &si)
{
    const size_t np = f.npad, G = f.G;
    si.z.assign(np * f.ns_total, 0);
    si.m.assign(64 * (size_t)f.namb_rows, 0);
    si.pi64.assign(64 * G, 0.0);
    si.fp.assign(np * G, 0.0);
    si.ra.assign(2 * G, RootArgs());
    for (Gene *gp : f.genes) {
        const Gene &g = *gp;
        for (int s = 0; s < g.ns; s++)
            std::memcpy(&si.z[(size_t)(g.tip_base + s) * np], &g.z[(size_t)s * g.npatt], g.npatt);
        for (int a = 0; a < g.namb; a++)
            for (int k = 0; k < kS; k++)
                si.m[(size_t)(g.amb_base + a) * 64 + k] = g.code_mask[(size_t)(kS + a) * kS + k];
        for (int i = 0; i < kS; i++) si.pi64[(size_t)g.gi * 64 + i] = g.pi[i];
        for (int h = 0; h < g.npatt; h++) si.fp[(size_t)g.gi * np + h] = g.fpatt[h];
        for (int w = 0; w < 2; w++) {
            RootArgs &r = si.ra[(size_t)w * G + g.gi];
            r.npatt = g.npatt;
            r.npad = f.npad;
            for (int c = 0; c < 4; c++) {
                r.root_slot[c] = gslot(g, g.root_slot[c]);
                // adjoint slots of the root: interior node root, class c
                r.seed_aslot[c] = w ? gaslot(g, (g.root - g.ns) * 4 + c) : -1;
            }
        }
    }
}

int forest_upload(Forest &f, int device)
{
    f.dev = device;
    const size_t ts = tsize(f);
    const size_t np = f.npad, G = f.G;
    HIPCHK(hipSetDevice(f.dev));
    HIPCHK(hipStreamCreateWithFlags(&f.stream, hipStreamNonBlocking));
    for (auto &v : f.ev) HIPCHK(hipEventCreate(&v));
    HIPCHK(hipMalloc(&f.d_uvr, sizeof(double) * f.model_block()));
    HIPCHK(hipMalloc(&f.d_pjobs, sizeof(PmatJob) * f.pjobs.size()));
    HIPCHK(hipMalloc(&f.d_prune, sizeof(PruneJob) * std::max<size_t>(1, f.prune.size())));
    HIPCHK(hipMalloc(&f.d_rargs, sizeof(RootArgs) * 2 * G));
    const size_t nnode_total = (size_t)f.ns_total + f.nint_total;
    HIPCHK(hipMalloc(&f.d_PT, sizeof(double) * 4096 * nnode_total * 3));
    HIPCHK(hipMemset(f.d_PT, 0, sizeof(double) * 4096 * nnode_total * 3));
    // tip table: every code's row for every tip branch and eigen set
    HIPCHK(hipMalloc(&f.d_AT, sizeof(double) * 64 * (size_t)f.ncodes * f.ns_total * 3));
    if (f.namb_rows > 0) HIPCHK(hipMalloc(&f.d_mask_amb, 64 * (size_t)f.namb_rows));
    HIPCHK(hipMalloc(&f.d_L, ts * 64 * np * f.nslot));
    HIPCHK(hipMalloc(&f.d_E, sizeof(int) * np * f.nslot));
    HIPCHK(hipMalloc(&f.d_z, np * f.ns_total));
    HIPCHK(hipMalloc(&f.d_pi, sizeof(double) * 64 * G));
    HIPCHK(hipMalloc(&f.d_fpatt, sizeof(double) * np * G));
    HIPCHK(hipMalloc(&f.d_fh, sizeof(double) * 4 * np * G));
    HIPCHK(hipMalloc(&f.d_site, sizeof(double) * np * G));
    f.nroot_blk = f.npad / 16;   // k_root: 16 patterns per block, padding included
    HIPCHK(hipMalloc(&f.d_partial, sizeof(double) * 2 * kRootSums * f.nroot_blk * G));
    HIPCHK(hipMalloc(&f.d_out, sizeof(double) * kRootSums * G));
    HIPCHK(hipHostMalloc((void **)&f.h_uvr, sizeof(double) * f.model_block(), hipHostMallocDefault));
    std::memset(f.h_uvr, 0, sizeof(double) * f.model_block());
    HIPCHK(hipHostMalloc((void **)&f.h_out, sizeof(double) * kRootSums * G, hipHostMallocDefault));
    HIPCHK(hipHostMalloc((void **)&f.h_pjobs, sizeof(PmatJob) * f.pjobs.size(),
                         hipHostMallocDefault));
    std::memcpy(f.h_pjobs, f.pjobs.data(), sizeof(PmatJob) * f.pjobs.size());

    // static inputs
    StaticInputs si;
    static_inputs(f, si);
    HIPCHK(hipMemcpy(f.d_z, si.z.data(), si.z.size(), hipMemcpyHostToDevice));
    if (f.namb_rows > 0)
        HIPCHK(hipMemcpy(f.d_mask_amb, si.m.data(), si.m.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f.d_pi, si.pi64.data(), sizeof(double) * si.pi64.size(),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f.d_fpatt, si.fp.data(), sizeof(double) * si.fp.size(),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f.d_rargs, si.ra.data(), sizeof(RootArgs) * si.ra.size(),
                     hipMemcpyHostToDevice));
    if (!f.prune.empty())
        HIPCHK(hipMemcpy(f.d_prune, f.prune.data(), sizeof(PruneJob) * f.prune.size(),
                         hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&f.d_wjobs, sizeof(int) * std::max<size_t>(1, f.wjobs.size())));
    HIPCHK(hipMalloc(&f.d_woff, sizeof(int) * f.woff.size()));
    HIPCHK(hipMalloc(&f.d_witems, sizeof(int) * std::max<size_t>(1, f.witems.size())));
    if (!f.wjobs.empty())
        HIPCHK(hipMemcpy(f.d_wjobs, f.wjobs.data(), sizeof(int) * f.wjobs.size(),
                         hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(f.d_woff, f.woff.data(), sizeof(int) * f.woff.size(), hipMemcpyHostToDevice));
    if (!f.witems.empty())
        HIPCHK(hipMemcpy(f.d_witems, f.witems.data(), sizeof(int) * f.witems.size(),
                         hipMemcpyHostToDevice));
    int ncu = 0;
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, f.dev));
    f.walk_nblk = 2 * std::max(1, ncu);
    return BSA_OK;
}

void forest_plan_adjoint(Forest &f)
{
    if (!f.adj.empty()) return;
    const int wb = (int)tsize(f);
    size_t maxd = 0;
    for (Gene *gp : f.genes) {
        gene_plan_adjoint(*gp, wb);
        maxd = std::max(maxd, gp->adj_off.size() - 1);
    }
    // per depth: the interior-child jobs of every gene, then the tip-child
    // jobs; the branch-matrix row of a job is its position
    std::vector<std::vector<int>> gidx(f.G);   // gene-local job -> forest job
    for (Gene *gp : f.genes) {
        gp->adj_int_rng.assign(maxd, {0, 0});
        gp->adj_tip_rng.assign(maxd, {0, 0});
    }
    f.adj.clear();
    f.adj_off.assign(1, 0);
    f.adj_nint.clear();
    for (size_t d = 0; d < maxd; d++) {
        int nint = 0;
        for (int pass = 0; pass < 2; pass++)
            for (Gene *gp : f.genes) {
                Gene &g = *gp;
                if (d + 1 >= g.adj_off.size()) continue;
                const int a = g.adj_off[d], b = g.adj_off[d + 1], ni = g.adj_nint[d];
                const int lo = pass == 0 ? a : a + ni, hi = pass == 0 ? a + ni : b;
                if (pass == 0) nint += ni;
                std::array<int, 2> &rng = pass == 0 ? g.adj_int_rng[d] : g.adj_tip_rng[d];
                rng[0] = (int)f.adj.size();
                rng[1] = (int)f.adj.size() + (hi - lo);
                std::vector<int> &gx = gidx[g.gi];
                gx.resize(g.adj.size(), -1);
                for (int i = lo; i < hi; i++) {
                    AdjJob j = g.adj[i];
                    j.v_tip = gtip(g, j.v_tip);
                    j.v_mask = g.mask_base;
                    j.v_slot = gslot(g, j.v_slot);
                    j.v_pmat = gpm(f, g, j.v_pmat);
                    j.v_aslot = gaslot(g, j.v_aslot);
                    j.m_index = (int)f.adj.size();
                    for (int t = 0; t < j.nterm; t++) j.a_aslot[t] = gaslot(g, j.a_aslot[t]);
                    for (int q = 0; q < j.nsib; q++) {
                        j.sib_tip[q] = gtip(g, j.sib_tip[q]);
                        for (int t = 0; t < j.nterm; t++) {
                            j.sib_slot[t][q] = gslot(g, j.sib_slot[t][q]);
                            j.sib_cslot[t][q] = gaslot(g, j.sib_cslot[t][q]);
                            j.sib_pmat[t][q] = gpm(f, g, j.sib_pmat[t][q]);
                        }
                    }
                    gx[i] = j.m_index;
                    f.adj.push_back(j);
                }
            }
        f.adj_off.push_back((int)f.adj.size());
        f.adj_nint.push_back(nint);
    }
    // eigen-basis jobs, gene by gene; set lists per forest eigen set 3 g + s
    f.ejobs.clear();
    std::vector<std::vector<int>> lists(3 * (size_t)f.G);
    for (Gene *gp : f.genes) {
        Gene &g = *gp;
        g.ej_base = (int)f.ejobs.size();
        for (const EigJob &e : g.ejobs) {
            EigJob q = e;
            q.m_index = gidx[g.gi][e.m_index];
            q.set = 3 * g.gi + e.set;
            lists[q.set].push_back((int)f.ejobs.size());
            f.ejobs.push_back(q);
        }
    }
    f.set_off = {0};
    f.set_list.clear();
    for (auto &l : lists) {
        f.set_list.insert(f.set_list.end(), l.begin(), l.end());
        f.set_off.push_back((int)f.set_list.size());
    }
    f.nmrow = (int)f.adj.size();   // one branch matrix per job
    const int ntile = f.npad / 64;
    // tiles per branch-matrix block: up to kAdjTilesPerChunkAuto, fewer when
    // the genes are small, so that the launch still has about 1024 blocks
    f.tpc = f.tpc_set > 0 ? f.tpc_set
                          : std::max(1, std::min(kAdjTilesPerChunkAuto,
                                                 ntile * (int)f.adj.size() / 1024));
    f.nchunk = (ntile + f.tpc - 1) / f.tpc;
    for (int k = 6; k < 10; k++) {
        f.work[k] = 0;
        for (Gene *gp : f.genes) f.work[k] += gp->work[k];
    }
}

int forest_setup_adjoint(Forest &f)
{
    if (f.adj_ready) return BSA_OK;
    forest_plan_adjoint(f);
    int rc = forest_check(f);
    if (rc) return rc;
    const size_t ts = tsize(f), nej = f.ejobs.size(), G = f.G;
    const size_t nasl = (size_t)f.nint_total * 4;
    HIPCHK(hipSetDevice(f.dev));
    HIPCHK(hipMalloc(&f.d_adj, sizeof(AdjJob) * f.adj.size()));
    HIPCHK(hipMemcpy(f.d_adj, f.adj.data(), sizeof(AdjJob) * f.adj.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&f.d_ejobs, sizeof(EigJob) * nej));
    HIPCHK(hipHostMalloc((void **)&f.h_ejobs, sizeof(EigJob) * nej, hipHostMallocDefault));
    std::memcpy(f.h_ejobs, f.ejobs.data(), sizeof(EigJob) * nej);
    HIPCHK(hipMalloc(&f.d_set_off, sizeof(int) * f.set_off.size()));
    HIPCHK(hipMemcpy(f.d_set_off, f.set_off.data(), sizeof(int) * f.set_off.size(),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&f.d_set_list, sizeof(int) * std::max<size_t>(1, f.set_list.size())));
    HIPCHK(hipMemcpy(f.d_set_list, f.set_list.data(), sizeof(int) * f.set_list.size(),
                     hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&f.d_A, ts * 64 * (size_t)f.npad * nasl));
    HIPCHK(hipMalloc(&f.d_C, ts * 64 * (size_t)f.npad * nasl));
    HIPCHK(hipMalloc(&f.d_EA, sizeof(int) * (size_t)f.npad * nasl));
    HIPCHK(hipMalloc(&f.d_Mpart, sizeof(double) * 4096 * (size_t)f.nmrow * f.nchunk));
    HIPCHK(hipMemset(f.d_Mpart, 0, sizeof(double) * 4096 * (size_t)f.nmrow * f.nchunk));
    f.mpart_chunks = f.nchunk;
    HIPCHK(hipMalloc(&f.d_W, sizeof(double) * 4096 * 3 * G * kWParts));
    HIPCHK(hipMalloc(&f.d_G, sizeof(double) * 4096 * nej));
    HIPCHK(hipMalloc(&f.d_dtau, sizeof(double) * nej));
    HIPCHK(hipMalloc(&f.d_Z, sizeof(double) * 3 * G * 4096));
    // code-indicator rows (61 sense codons, then the ambiguity codes) per gene
    std::vector<double> m64((size_t)f.nmask_rows * 64, 0.0);
    for (Gene *gp : f.genes)
        for (int c = 0; c < gp->ncodes; c++)
            for (int k = 0; k < kS; k++)
                m64[(size_t)(gp->mask_base + c) * 64 + k] = gp->code_mask[(size_t)c * kS + k];
    HIPCHK(hipMalloc(&f.d_mask64, sizeof(double) * m64.size()));
    HIPCHK(hipMemcpy(f.d_mask64, m64.data(), sizeof(double) * m64.size(), hipMemcpyHostToDevice));
    // each tip's patterns in the order of its code (counting sort, stable),
    // the forest's padding patterns after them in place: a 64-pattern tile
    // of a tip's branch matrix then holds one or a few codons.  An entry is
    // the pattern (low 24 bits) and its code (top 8 bits).
    if (f.npad > (1 << 24)) return fail(BSA_EINVAL, "more than 2^24 site patterns");
    std::vector<int> tp((size_t)f.ns_total * f.npad);
    for (Gene *gp : f.genes) {
        const Gene &g = *gp;
        for (int v = 0; v < g.ns; v++) {
            int *row = tp.data() + (size_t)(g.tip_base + v) * f.npad;
            const unsigned char *zv = g.z.data() + (size_t)v * g.npatt;
            int start[257] = {0};
            for (int h = 0; h < g.npatt; h++) start[zv[h] + 1]++;
            for (int c = 0; c < 256; c++) start[c + 1] += start[c];
            for (int h = 0; h < g.npatt; h++) row[start[zv[h]]++] = h;
            for (int h = g.npatt; h < f.npad; h++) row[h] = h;
            for (int i = 0; i < f.npad; i++) {
                const int h = row[i];
                const unsigned code = h < g.npatt ? zv[h] : 0u;
                row[i] = (int)((code << 24) | (unsigned)h);
            }
        }
    }
    HIPCHK(hipMalloc(&f.d_tperm, sizeof(int) * tp.size()));
    HIPCHK(hipMemcpy(f.d_tperm, tp.data(), sizeof(int) * tp.size(), hipMemcpyHostToDevice));
    HIPCHK(hipHostMalloc((void **)&f.h_adjout, sizeof(double) * (nej + 3 * G * 4096),
                         hipHostMallocDefault));
    f.adj_ready = true;
    return BSA_OK;
}

// Every index a kernel will follow, checked against the buffer it indexes
// (forward schedule always; the adjoint schedule once planned).
int forest_check(const Forest &f)
{
    const long nnode = (long)f.ns_total + f.nint_total, npm = 3 * nnode, ntip_pm = 3L * f.ns_total;
    const long nasl = 4L * f.nint_total, nset = 3L * f.G;
    auto in = [](long v, long n) { return v >= 0 && v < n; };
    auto bad = [](const char *what, long i) {
        return fail(BSA_EINVAL, "forest schedule: %s of job %ld out of range", what, i);
    };
    if ((long)f.pjobs.size() == 0 || f.npad % kPadP != 0 || f.npad <= 0) return bad("size", -1);
    for (size_t i = 0; i < f.pjobs.size(); i++) {
        const PmatJob &p = f.pjobs[i];
        if (!in(p.node, nnode) || !in(p.pm, npm) || !in(p.set, nset) || p.pm / 3 != p.node)
            return bad("P(t) matrix", (long)i);
        if (p.tip >= 0 && (!in(p.tip, f.ns_total) || p.pm >= ntip_pm)) return bad("tip", (long)i);
        if (f.ncodes > kS && (p.amb < 0 || p.amb + (f.ncodes - kS) > f.namb_rows))
            return bad("ambiguity rows", (long)i);
    }
    for (size_t i = 0; i < f.prune.size(); i++) {
        const PruneJob &j = f.prune[i];
        if (!in(j.out_slot, f.nslot) || j.nch < 1 || j.nch > kMaxCh) return bad("pruning", (long)i);
        for (int k = 0; k < j.nch; k++) {
            if (!in(j.ch_pmat[k], npm)) return bad("child P(t)", (long)i);
            if (j.ch_tip[k] >= 0) {
                if (!in(j.ch_tip[k], f.ns_total) || j.ch_pmat[k] >= ntip_pm)
                    return bad("tip child", (long)i);
            } else if (!in(j.ch_slot[k], f.nslot) || !in(j.ch_cslot[k], nasl)) {
                return bad("interior child", (long)i);
            }
        }
    }
    // persistent-pruning lists: each gene's jobs, in level order, exactly
    // once; items within the forest's tiles
    if ((long)f.woff.size() != f.G + 1 || f.woff[0] != 0 ||
        f.woff.back() != (int)f.wjobs.size() || f.wjobs.size() != f.prune.size())
        return bad("walk lists", -1);
    for (const Gene *gp : f.genes) {
        const Gene &g = *gp;
        int q = f.woff[g.gi];
        for (const auto &rg : g.prune_rng)
            for (int i = rg[0]; i < rg[1]; i++, q++)
                if (q >= f.woff[g.gi + 1] || f.wjobs[q] != i) return bad("walk jobs", g.gi);
        if (q != f.woff[g.gi + 1]) return bad("walk jobs", g.gi);
    }
    if (f.witems.size() != (size_t)f.G * (f.npad / kPadP)) return bad("walk items", -1);
    for (int v : f.witems)
        if (!in(v, (long)f.G * (f.npad / kPadP))) return bad("walk item", v);
    for (const Gene *gp : f.genes) {
        const Gene &g = *gp;
        for (int c = 0; c < 4; c++)
            if (!in(g.slot_base + g.root_slot[c], f.nslot)) return bad("root slot", g.gi);
        if (g.int_base + (g.root - g.ns) >= f.nint_total || g.npatt > f.npad)
            return bad("root", g.gi);
        if (g.mask_base + g.ncodes > f.nmask_rows) return bad("code rows", g.gi);
    }
    if (f.adj.empty()) return BSA_OK;
    for (size_t i = 0; i < f.adj.size(); i++) {
        const AdjJob &j = f.adj[i];
        if (j.m_index != (int)i || !in(j.v_pmat, npm) || j.nterm < 1 || j.nterm > 2 ||
            j.nsib < 0 || j.nsib > kMaxCh - 1)
            return bad("adjoint", (long)i);
        if (j.v_tip >= 0) {
            if (!in(j.v_tip, f.ns_total) || j.v_pmat >= ntip_pm || j.v_aslot >= 0)
                return bad("adjoint tip", (long)i);
        } else if (!in(j.v_slot, f.nslot) || !in(j.v_aslot, nasl)) {
            return bad("adjoint child", (long)i);
        }
        if (j.v_mask < 0 || j.v_mask >= f.nmask_rows) return bad("code rows", (long)i);
        for (int t = 0; t < j.nterm; t++) {
            if (!in(j.a_aslot[t], nasl)) return bad("parent adjoint", (long)i);
            for (int q = 0; q < j.nsib; q++) {
                if (!in(j.sib_pmat[t][q], npm)) return bad("sibling P(t)", (long)i);
                if (j.sib_tip[q] >= 0) {
                    if (!in(j.sib_tip[q], f.ns_total) || j.sib_pmat[t][q] >= ntip_pm)
                        return bad("tip sibling", (long)i);
                } else if (!in(j.sib_slot[t][q], f.nslot) || !in(j.sib_cslot[t][q], nasl)) {
                    return bad("interior sibling", (long)i);
                }
            }
        }
    }
    for (size_t d = 0; d + 1 < f.adj_off.size(); d++) {
        const int a = f.adj_off[d], b = f.adj_off[d + 1], ni = f.adj_nint[d];
        if (a > b || ni > b - a) return bad("depth", (long)d);
        for (int i = a; i < b; i++)
            if ((f.adj[i].v_tip < 0) != (i < a + ni)) return bad("depth order", (long)i);
    }
    for (size_t i = 0; i < f.ejobs.size(); i++)
        if (!in(f.ejobs[i].m_index, f.nmrow) || !in(f.ejobs[i].set, nset))
            return bad("eigen-basis", (long)i);
    if ((long)f.set_off.size() != nset + 1 || f.set_off.back() != (int)f.ejobs.size())
        return bad("set lists", -1);
    for (int v : f.set_list)
        if (!in(v, (long)f.ejobs.size())) return bad("set list", v);
    if (f.tpc < 1 || f.nchunk < 1 || (long)f.nchunk * f.tpc * 64 < f.npad) return bad("chunks", -1);
    return BSA_OK;
}

int forest_build(Forest &f, const std::vector<Gene *> &genes, int device, int precision,
                 int kernel, int tile)
{
    int rc = forest_plan(f, genes, precision, kernel, tile);
    if (rc) return rc;
    if ((rc = forest_check(f))) return rc;
    return forest_upload(f, device);
}

// ---------------------------------------------------------------- evaluation

// Job lists of one round (device pointers into Forest::d_sel and counts);
// a null list means every job of that launch, in order.
struct RoundSel {
    const int *pj = nullptr;
    int npj = 0;
    std::vector<const int *> lev;       // per pruning level
    std::vector<int> nlev;
    const int *genes = nullptr;
    int ngenes = 0;
    std::vector<const int *> adl;       // per adjoint depth, interior children
    std::vector<int> nadl;
    const int *adm = nullptr;           // every adjoint job (branch matrices)
    int nadm = 0;
    const int *ej = nullptr;
    int nej = 0;
    const int *sets = nullptr;
    int nsets = 0;
    const int *witems = nullptr;        // persistent-pruning items of the round's genes
    int nwitems = 0;
};

// P(t), pruning and root pass; with `adjoint`, the contributions and the
// root's adjoint seeds that the sweep will need are stored too
template <typename T>
static int enqueue_forward(Forest &f, hipStream_t s, bool adjoint, bool events,
                           const RoundSel *rs = nullptr)
{
#define REC(i) \
    if (events) HIPCHK(hipEventRecord(f.ev[i], s))
    REC(0);
    launch_pmat(s, f.d_uvr, f.d_pjobs, rs ? rs->pj : nullptr, rs ? rs->npj : (int)f.pjobs.size(),
                (double *)f.d_PT, f.d_mask_amb, f.ncodes, (double *)f.d_AT);
    HIPCHK(hipGetLastError());
    REC(1);
    // BSA_KERNEL_FP32MC: the pruning's products on the FP32 matrix cores
    const int pk = f.kernel == BSA_KERNEL_VALU ? 1 : f.kernel == BSA_KERNEL_FP32MC ? 2 : 0;
    if (f.prune_walk) {
        // every level in one persistent launch (FP64 matrix-core products)
        launch_prune_walk<T>(s, f.d_prune, f.d_wjobs, f.d_woff, rs ? rs->witems : f.d_witems,
                             rs ? rs->nwitems : (int)f.witems.size(), f.walk_nblk, f.npad, f.d_z,
                             (const double *)f.d_PT, (const double *)f.d_AT, f.ncodes,
                             (T *)f.d_L, f.d_E, adjoint ? (T *)f.d_C : nullptr);
    }
    for (size_t l = 0; !f.prune_walk && l + 1 < f.level_off.size(); l++) {
        const int a = f.level_off[l], b = f.level_off[l + 1];
        launch_prune<T>(s, pk, f.tile,
                        rs ? f.d_prune : f.d_prune + a, rs ? rs->lev[l] : nullptr,
                        rs ? rs->nlev[l] : b - a, f.level_maxint[l], f.npad, f.d_z,
                        (const double *)f.d_PT, (const double *)f.d_AT, f.ncodes, (T *)f.d_L,
                        f.d_E, adjoint ? (T *)f.d_C : nullptr);
    }
    HIPCHK(hipGetLastError());
    REC(2);
    launch_root<T>(s, f.d_rargs + (adjoint ? f.G : 0), rs ? rs->genes : nullptr,
                   rs ? rs->ngenes : f.G, (const T *)f.d_L, f.d_E, f.d_pi, f.d_uvr, f.d_fpatt,
                   f.d_fh, f.d_site, f.d_partial, f.nroot_blk, f.d_out, (T *)f.d_A, f.d_EA);
    HIPCHK(hipGetLastError());
    REC(3);
#undef REC
    return BSA_OK;
}

// The adjoint sweep (child adjoints per depth, branch matrices, eigen-basis
// finish) of the last forward pass run with `adjoint`
template <typename T>
static int enqueue_adjoint(Forest &f, hipStream_t s, bool events,
                           const RoundSel *rs = nullptr)
{
#define REC(i) \
    if (events) HIPCHK(hipEventRecord(f.ev[i], s))
    // BSA_KERNEL_MFMA32 and _FP32MC: the child adjoints and the interior
    // children's branch matrices on the FP32 matrix cores
    const bool mfma32 = sizeof(T) == 4 &&
                        (f.kernel == BSA_KERNEL_MFMA32 || f.kernel == BSA_KERNEL_FP32MC);
    for (size_t d = 0; d + 1 < f.adj_off.size(); d++) {
        if (mfma32)
            launch_adjoint_level32(s, rs ? f.d_adj : f.d_adj + f.adj_off[d],
                                   rs ? rs->adl[d] : nullptr,
                                   rs ? rs->nadl[d] : f.adj_nint[d], f.npad, f.d_z,
                                   (const double *)f.d_PT, (const double *)f.d_AT, f.ncodes,
                                   f.d_E, (const float *)f.d_C, (float *)f.d_A, f.d_EA,
                                   f.tpb_set);
        else
            launch_adjoint_level<T>(s, rs ? f.d_adj : f.d_adj + f.adj_off[d],
                                    rs ? rs->adl[d] : nullptr,
                                    rs ? rs->nadl[d] : f.adj_nint[d], f.npad, f.d_z,
                                    (const double *)f.d_PT, (const double *)f.d_AT, f.ncodes,
                                    f.d_E, (const T *)f.d_C, (T *)f.d_A, f.d_EA, f.tpb_set);
    }
    REC(6);
    launch_adjoint_m<T>(s, f.d_adj, rs ? rs->adm : nullptr, rs ? rs->nadm : (int)f.adj.size(),
                        f.npad, f.nchunk, f.tpc, mfma32, f.d_z, (const double *)f.d_PT,
                        (const double *)f.d_AT, f.ncodes, f.d_mask64, f.d_tperm,
                        (const T *)f.d_L, f.d_E, (const T *)f.d_C, (const T *)f.d_A, f.d_EA,
                        f.d_Mpart);
    REC(4);
    launch_adjoint_finish(s, f.d_Mpart, f.nchunk, f.d_ejobs, rs ? rs->ej : nullptr,
                          rs ? rs->nej : (int)f.ejobs.size(), rs ? rs->sets : nullptr,
                          rs ? rs->nsets : 3 * f.G, f.d_set_off, f.d_set_list, f.d_uvr, f.d_G,
                          f.d_dtau, f.d_W, f.d_Z);
    HIPCHK(hipGetLastError());
    REC(5);
#undef REC
    return BSA_OK;
}

template <typename T>
static int enqueue_device(Forest &f, hipStream_t s, bool adjoint, bool events,
                          const RoundSel *rs = nullptr)
{
    int rc = enqueue_forward<T>(f, s, adjoint, events, rs);
    if (!rc && adjoint) rc = enqueue_adjoint<T>(f, s, events, rs);
    return rc;
}

// device times of the phases of the last evaluation from the stream's events
// (forward: whether the pruning ran, for the launch count)
static int read_timing(Forest &f, bool adjoint, int nadj_launch, bool forward = true)
{
    float ms;
    HIPCHK(hipEventElapsedTime(&ms, f.ev[0], f.ev[1]));
    f.timing[0] = ms;
    HIPCHK(hipEventElapsedTime(&ms, f.ev[1], f.ev[2]));
    f.timing[1] = ms;
    HIPCHK(hipEventElapsedTime(&ms, f.ev[2], f.ev[3]));
    f.timing[2] = ms;
    f.timing[3] = f.timing[4] = f.timing[5] = 0;
    if (adjoint) {
        HIPCHK(hipEventElapsedTime(&ms, f.ev[3], f.ev[6]));
        f.timing[3] = ms;
        HIPCHK(hipEventElapsedTime(&ms, f.ev[6], f.ev[4]));
        f.timing[4] = ms;
        HIPCHK(hipEventElapsedTime(&ms, f.ev[4], f.ev[5]));
        f.timing[5] = ms;
        f.n_adj_launch += nadj_launch;
    }
    for (int i = 0; i < 6; i++) f.timing_sum[i] += f.timing[i];
    if (forward) f.n_launch += f.prune_walk ? 1.0 : (double)(f.level_off.size() - 1);
    return BSA_OK;
}

template <typename T>
static int run_device(Forest &f, bool adjoint)
{
    hipStream_t s = f.stream;
    HIPCHK(hipMemcpyAsync(f.d_uvr, f.h_uvr, sizeof(double) * f.model_block(), hipMemcpyHostToDevice,
                          s));
    HIPCHK(hipMemcpyAsync(f.d_pjobs, f.h_pjobs, sizeof(PmatJob) * f.pjobs.size(),
                          hipMemcpyHostToDevice, s));
    if (adjoint)
        HIPCHK(hipMemcpyAsync(f.d_ejobs, f.h_ejobs, sizeof(EigJob) * f.ejobs.size(),
                              hipMemcpyHostToDevice, s));
    if (f.use_graph) {
        // the launch sequence depends only on the trees and the data, so it
        // is captured once and replayed; per-evaluation inputs live in device
        // buffers refreshed by the copies above
        hipGraphExec_t gr = f.graph[adjoint ? 1 : 0];
        if (!gr) return fail(BSA_EDEVICE, "evaluation graph not captured");
        HIPCHK(hipEventRecord(f.ev[0], s));
        HIPCHK(hipGraphLaunch(gr, s));
        HIPCHK(hipEventRecord(f.ev[5], s));
    } else {
        const int rc = enqueue_device<T>(f, s, adjoint, true);
        if (rc) return rc;
    }
    if (adjoint) {
        HIPCHK(hipMemcpyAsync(f.h_adjout, f.d_dtau, sizeof(double) * f.ejobs.size(),
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(f.h_adjout + f.ejobs.size(), f.d_Z, sizeof(double) * 3 * f.G * 4096,
                              hipMemcpyDeviceToHost, s));
    }
    HIPCHK(hipMemcpyAsync(f.h_out, f.d_out, sizeof(double) * kRootSums * f.G, hipMemcpyDeviceToHost,
                          s));
    HIPCHK(hipStreamSynchronize(s));
    if (f.use_graph) {
        // one device time for the whole evaluation (out[0] of bsa_timing)
        float ms;
        HIPCHK(hipEventElapsedTime(&ms, f.ev[0], f.ev[5]));
        for (int i = 0; i < 6; i++) f.timing[i] = 0;
        f.timing[0] = ms;
        for (int i = 0; i < 6; i++) f.timing_sum[i] += f.timing[i];
        return BSA_OK;
    }
    int nl = 1;   // the branch-matrix launch, plus one per depth with interior children
    for (int v : f.adj_nint) nl += v > 0;
    return read_timing(f, adjoint, nl);
}

// The two parts of an evaluation with gradient, run one after the other:
// part 1 the forward pass that keeps what the sweep needs (and the staged
// inputs of every phase), part 2 the adjoint sweep of the last part 1.  The
// device work and its order are those of run_device(f, true), so the
// results are the same bitwise.
template <typename T>
static int run_part(Forest &f, int part)
{
    hipStream_t s = f.stream;
    if (part == 1) {
        HIPCHK(hipMemcpyAsync(f.d_uvr, f.h_uvr, sizeof(double) * f.model_block(),
                              hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(f.d_pjobs, f.h_pjobs, sizeof(PmatJob) * f.pjobs.size(),
                              hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(f.d_ejobs, f.h_ejobs, sizeof(EigJob) * f.ejobs.size(),
                              hipMemcpyHostToDevice, s));
        int rc = enqueue_forward<T>(f, s, true, true);
        if (rc) return rc;
        HIPCHK(hipMemcpyAsync(f.h_out, f.d_out, sizeof(double) * kRootSums * f.G,
                              hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        // the forward phases of this evaluation; the sweep's are added by part 2
        rc = read_timing(f, false, 0);
        return rc;
    }
    HIPCHK(hipEventRecord(f.ev[3], s));
    int rc = enqueue_adjoint<T>(f, s, true);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(f.h_adjout, f.d_dtau, sizeof(double) * f.ejobs.size(),
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(f.h_adjout + f.ejobs.size(), f.d_Z, sizeof(double) * 3 * f.G * 4096,
                          hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int nl = 1;
    for (int v : f.adj_nint) nl += v > 0;
    float ms;
    HIPCHK(hipEventElapsedTime(&ms, f.ev[3], f.ev[6]));
    f.timing[3] = ms;
    HIPCHK(hipEventElapsedTime(&ms, f.ev[6], f.ev[4]));
    f.timing[4] = ms;
    HIPCHK(hipEventElapsedTime(&ms, f.ev[4], f.ev[5]));
    f.timing[5] = ms;
    for (int i = 3; i < 6; i++) f.timing_sum[i] += f.timing[i];
    f.n_adj_launch += nl;
    return BSA_OK;
}

int forest_run_part(Forest &f, int part)
{
    if (part != 1 && part != 2) return fail(BSA_EINVAL, "evaluation part %d: 1 or 2", part);
    if (!f.adj_ready) return fail(BSA_EINVAL, "adjoint not set up");
    if (f.use_graph) return fail(BSA_EINVAL, "split evaluations need eager launches (no graph)");
    HIPCHK(hipSetDevice(f.dev));
    return f.precision == BSA_FP32 ? run_part<float>(f, part) : run_part<double>(f, part);
}

// copies of the per-gene ranges [base(g), base(g) + len(g)) of the listed
// genes (ascending), runs of consecutive genes merged into one copy; all
// byte offsets and lengths are multiples of `unit`
template <class Base, class Len>
static int copy_runs(hipStream_t s, void *dst, const void *src, size_t unit,
                     const std::vector<int> &genes, Base base, Len len, hipMemcpyKind kind)
{
    size_t i = 0;
    while (i < genes.size()) {
        size_t j = i;
        while (j + 1 < genes.size() && genes[j + 1] == genes[j] + 1) j++;
        const size_t a = base(genes[i]), b = base(genes[j]) + len(genes[j]);
        if (b > a)
            HIPCHK(hipMemcpyAsync((char *)dst + a * unit, (const char *)src + a * unit,
                                  (b - a) * unit, kind, s));
        i = j + 1;
    }
    return BSA_OK;
}

// A round: the forward passes of the genes gl (with keep, storing what a
// sweep needs) and the adjoint sweeps of the genes ga, each of the last
// forward pass run with keep for that gene (this round's where the gene is
// in both lists).  Gene regions of the device buffers are disjoint, so a
// gene's kept forward pass survives other genes' rounds.
template <typename T>
static int run_round(Forest &f, const std::vector<int> &gl, const std::vector<int> &ga, bool keep)
{
    hipStream_t s = f.stream;
    const bool adjoint = !ga.empty();
    const size_t nlev = f.level_off.size() - 1, nd = adjoint ? f.adj_off.size() - 1 : 0;
    // job lists: P(t), pruning levels, genes, child adjoints per depth,
    // branch matrices, eigen-basis jobs, eigen sets
    size_t cap = f.pjobs.size() + f.prune.size() + f.G + 2 * f.adj.size() + f.ejobs.size() +
                 3 * (size_t)f.G + f.witems.size();
    if (cap > f.sel_cap) {
        if (f.h_sel) HIPCHK(hipHostFree(f.h_sel));
        if (f.d_sel) HIPCHK(hipFree(f.d_sel));
        f.h_sel = nullptr;
        f.d_sel = nullptr;
        f.sel_cap = 0;
        HIPCHK(hipHostMalloc((void **)&f.h_sel, sizeof(int) * cap, hipHostMallocDefault));
        HIPCHK(hipMalloc(&f.d_sel, sizeof(int) * cap));
        f.sel_cap = cap;
    }
    int *h = f.h_sel;
    size_t n = 0;
    auto put_range = [&](int a, int b) {
        for (int i = a; i < b; i++) h[n++] = i;
    };
    RoundSel rs;
    std::vector<size_t> off;   // start of each list in h_sel, in launch order
    off.push_back(n);
    for (int g : gl) {
        const Gene &ge = *f.genes[g];
        put_range(ge.pj_base, ge.pj_base + (int)ge.pjobs.size());
    }
    rs.npj = (int)(n - off.back());
    for (size_t l = 0; l < nlev; l++) {
        off.push_back(n);
        for (int g : gl) put_range(f.genes[g]->prune_rng[l][0], f.genes[g]->prune_rng[l][1]);
        rs.nlev.push_back((int)(n - off.back()));
    }
    off.push_back(n);
    for (int g : gl) h[n++] = g;
    rs.ngenes = (int)gl.size();
    off.push_back(n);
    if (f.prune_walk) {
        // the round's genes in the forest's walk order, every tile
        std::vector<char> in_round(f.G, 0);
        for (int g : gl) in_round[g] = 1;
        const int wt = f.npad / kPadP;
        for (int g : f.walk_order)
            if (in_round[g])
                for (int t = 0; t < wt; t++) h[n++] = g * wt + t;
    }
    rs.nwitems = (int)(n - off.back());
    if (adjoint) {
        for (size_t d = 0; d < nd; d++) {
            off.push_back(n);
            for (int g : ga) put_range(f.genes[g]->adj_int_rng[d][0], f.genes[g]->adj_int_rng[d][1]);
            rs.nadl.push_back((int)(n - off.back()));
        }
        off.push_back(n);
        for (size_t d = 0; d < nd; d++)
            for (int g : ga) {
                const Gene &ge = *f.genes[g];
                put_range(ge.adj_int_rng[d][0], ge.adj_int_rng[d][1]);
                put_range(ge.adj_tip_rng[d][0], ge.adj_tip_rng[d][1]);
            }
        rs.nadm = (int)(n - off.back());
        off.push_back(n);
        for (int g : ga) {
            const Gene &ge = *f.genes[g];
            put_range(ge.ej_base, ge.ej_base + (int)ge.ejobs.size());
        }
        rs.nej = (int)(n - off.back());
        off.push_back(n);
        for (int g : ga)
            for (int k = 0; k < 3; k++) h[n++] = 3 * g + k;
        rs.nsets = 3 * (int)ga.size();
    }
    if (n > f.sel_cap) return fail(BSA_EINVAL, "round job lists exceed their buffer");
    HIPCHK(hipMemcpyAsync(f.d_sel, h, sizeof(int) * n, hipMemcpyHostToDevice, s));
    size_t k = 0;
    rs.pj = f.d_sel + off[k++];
    for (size_t l = 0; l < nlev; l++) rs.lev.push_back(f.d_sel + off[k++]);
    rs.genes = f.d_sel + off[k++];
    rs.witems = f.d_sel + off[k++];
    if (adjoint) {
        for (size_t d = 0; d < nd; d++) rs.adl.push_back(f.d_sel + off[k++]);
        rs.adm = f.d_sel + off[k++];
        rs.ej = f.d_sel + off[k++];
        rs.sets = f.d_sel + off[k++];
    }
    // Inputs and outputs.  The per-gene parameter blocks and Z matrices (about
    // 200 and 100 KB a gene) cross for the listed genes only, one copy per run
    // of consecutive genes.  The small per-gene lists (P(t) jobs and
    // eigen-basis jobs in, lnL sums and d lnL/d tau out) cross whole, one
    // copy each: a round of scattered genes would otherwise take a copy per
    // gene and list (80 a round in a scan).  That is safe: kernels read only
    // the listed genes' entries; a gene awaiting the sweep of a kept forward
    // pass does not stage (its host lists are those of that pass); and a
    // gene's device outputs change only in rounds that list it, which copy
    // them back, so the other genes' host values are rewritten unchanged.
    int rc = copy_runs(s, f.d_uvr, f.h_uvr, sizeof(double), gl,
                       [](int g) { return (size_t)g * kGeneBlock; },
                       [](int) { return (size_t)kGeneBlock; }, hipMemcpyHostToDevice);
    if (rc) return rc;
    if (!gl.empty())
        HIPCHK(hipMemcpyAsync(f.d_pjobs, f.h_pjobs, sizeof(PmatJob) * f.pjobs.size(),
                              hipMemcpyHostToDevice, s));
    if (adjoint)
        HIPCHK(hipMemcpyAsync(f.d_ejobs, f.h_ejobs, sizeof(EigJob) * f.ejobs.size(),
                              hipMemcpyHostToDevice, s));
    if ((rc = enqueue_forward<T>(f, s, keep, true, &rs))) return rc;
    if (adjoint && (rc = enqueue_adjoint<T>(f, s, true, &rs))) return rc;
    if (adjoint) {
        HIPCHK(hipMemcpyAsync(f.h_adjout, f.d_dtau, sizeof(double) * f.ejobs.size(),
                              hipMemcpyDeviceToHost, s));
        rc = copy_runs(s, f.h_adjout + f.ejobs.size(), f.d_Z, sizeof(double), ga,
                       [](int g) { return (size_t)g * 3 * 4096; },
                       [](int) { return (size_t)3 * 4096; }, hipMemcpyDeviceToHost);
        if (rc) return rc;
    }
    if (!gl.empty())
        HIPCHK(hipMemcpyAsync(f.h_out, f.d_out, sizeof(double) * kRootSums * f.G,
                              hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int nl = 1;
    for (int v : rs.nadl) nl += v > 0;
    return read_timing(f, adjoint, nl, !gl.empty());
}

int forest_run_round(Forest &f, const std::vector<int> &fwd, const std::vector<int> &adj,
                     bool keep)
{
    for (const std::vector<int> *v : {&fwd, &adj})
        for (size_t i = 0; i < v->size(); i++)
            if ((*v)[i] < 0 || (*v)[i] >= f.G || (i > 0 && (*v)[i] <= (*v)[i - 1]))
                return fail(BSA_EINVAL, "round genes must be ascending gene indices");
    if (fwd.empty() && adj.empty()) return BSA_OK;
    if ((keep || !adj.empty()) && !f.adj_ready) return fail(BSA_EINVAL, "adjoint not set up");
    HIPCHK(hipSetDevice(f.dev));
    return f.precision == BSA_FP32 ? run_round<float>(f, fwd, adj, keep)
                                   : run_round<double>(f, fwd, adj, keep);
}

int forest_run_round(Forest &f, const std::vector<int> &genes, bool adjoint)
{
    return forest_run_round(f, genes, adjoint ? genes : std::vector<int>(), adjoint);
}

int forest_run(Forest &f, bool adjoint)
{
    HIPCHK(hipSetDevice(f.dev));
    return f.precision == BSA_FP32 ? run_device<float>(f, adjoint) : run_device<double>(f, adjoint);
}

// ------------------------------------------- evaluation in the caller's memory

size_t forest_place_lnl(Forest &f, char *base)
{
    const size_t ts = tsize(f), np = f.npad, G = f.G;
    const size_t nnode_total = (size_t)f.ns_total + f.nint_total;
    f.nroot_blk = f.npad / 16;
    size_t off = 0;
    auto take = [&](auto *&ptr, size_t bytes) {
        off = (off + 255) & ~(size_t)255;
        if (base) ptr = reinterpret_cast<std::remove_reference_t<decltype(ptr)>>(base + off);
        off += bytes;
    };
    take(f.d_uvr, sizeof(double) * f.model_block());
    take(f.d_pjobs, sizeof(PmatJob) * f.pjobs.size());
    take(f.d_prune, sizeof(PruneJob) * std::max<size_t>(1, f.prune.size()));
    take(f.d_rargs, sizeof(RootArgs) * 2 * G);
    take(f.d_PT, sizeof(double) * 4096 * nnode_total * 3);
    take(f.d_AT, sizeof(double) * 64 * (size_t)f.ncodes * f.ns_total * 3);
    if (f.namb_rows > 0) take(f.d_mask_amb, 64 * (size_t)f.namb_rows);
    take(f.d_L, ts * 64 * np * f.nslot);
    take(f.d_E, sizeof(int) * np * f.nslot);
    take(f.d_z, np * f.ns_total);
    take(f.d_pi, sizeof(double) * 64 * G);
    take(f.d_fpatt, sizeof(double) * np * G);
    take(f.d_fh, sizeof(double) * 4 * np * G);
    take(f.d_site, sizeof(double) * np * G);
    take(f.d_partial, sizeof(double) * 2 * kRootSums * f.nroot_blk * G);
    take(f.d_out, sizeof(double) * kRootSums * G);
    return (off + 255) & ~(size_t)255;
}

template <typename T>
static int lnl_in_place(Forest &f, hipStream_t s, const double *x, double *lnl, double *fhK,
                        double *site)
{
    Gene &g = *f.genes[0];
    std::vector<double> h_uvr(f.model_block(), 0.0), h_out((size_t)kRootSums * f.G, 0.0);
    std::vector<PmatJob> h_pj(f.pjobs);
    f.h_uvr = h_uvr.data();
    f.h_out = h_out.data();
    f.h_pjobs = h_pj.data();
    int rc = gene_stage_x(g, f, x, false);
    if (rc) return rc;
    StaticInputs si;
    static_inputs(f, si);
    const size_t nnode_total = (size_t)f.ns_total + f.nint_total;
    // the array keeps the last call's contents: P(t) - I must start at zero
    // (the pad rows and columns are never written)
    HIPCHK(hipMemsetAsync(f.d_PT, 0, sizeof(double) * 4096 * nnode_total * 3, s));
    HIPCHK(hipMemcpyAsync(f.d_z, si.z.data(), si.z.size(), hipMemcpyHostToDevice, s));
    if (f.namb_rows > 0)
        HIPCHK(hipMemcpyAsync(f.d_mask_amb, si.m.data(), si.m.size(), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f.d_pi, si.pi64.data(), sizeof(double) * si.pi64.size(),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f.d_fpatt, si.fp.data(), sizeof(double) * si.fp.size(),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f.d_rargs, si.ra.data(), sizeof(RootArgs) * si.ra.size(),
                          hipMemcpyHostToDevice, s));
    if (!f.prune.empty())
        HIPCHK(hipMemcpyAsync(f.d_prune, f.prune.data(), sizeof(PruneJob) * f.prune.size(),
                              hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f.d_uvr, f.h_uvr, sizeof(double) * f.model_block(),
                          hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(f.d_pjobs, f.h_pjobs, sizeof(PmatJob) * f.pjobs.size(),
                          hipMemcpyHostToDevice, s));
    if ((rc = enqueue_device<T>(f, s, false, false))) return rc;
    HIPCHK(hipMemcpyAsync(f.h_out, f.d_out, sizeof(double) * kRootSums * f.G,
                          hipMemcpyDeviceToHost, s));
    if (fhK)
        for (int c = 0; c < 4; c++)
            HIPCHK(hipMemcpyAsync(fhK + (size_t)c * g.npatt, f.d_fh + (size_t)c * f.npad,
                                  sizeof(double) * g.npatt, hipMemcpyDeviceToHost, s));
    if (site)
        HIPCHK(hipMemcpyAsync(site, f.d_site, sizeof(double) * g.npatt, hipMemcpyDeviceToHost, s));
    // the host images above are freed on return
    HIPCHK(hipStreamSynchronize(s));
    *lnl = gene_collect(g, f);
    return BSA_OK;
}

int forest_lnl_in_place(Forest &f, char *base, size_t cap, hipStream_t s, const double *x,
                        double *lnl, double *fhK, double *site)
{
    if (f.G != 1) return fail(BSA_EINVAL, "one gene per evaluation in given memory");
    const size_t need = forest_place_lnl(f, nullptr);
    if (need > cap)
        return fail(BSA_EINVAL,
                    "the gene needs %zu bytes of device scratch (%d taxa, %d patterns padded to %d, "
                    "%s partials); the limit is %zu bytes",
                    need, f.genes[0]->ns, f.genes[0]->npatt, f.npad,
                    f.precision == BSA_FP32 ? "FP32" : "FP64", cap);
    forest_place_lnl(f, base);
    f.stream = s;
    const int rc = f.precision == BSA_FP32 ? lnl_in_place<float>(f, s, x, lnl, fhK, site)
                                           : lnl_in_place<double>(f, s, x, lnl, fhK, site);
    // nothing here is the forest's to release
    Forest detached;
    detached.precision = f.precision;
    f = detached;
    return rc;
}

// Capture the device work of one evaluation as a graph (the launch sequence
// depends only on the trees and the data; per-evaluation inputs are copied
// to device buffers before each replay).
template <typename T>
static int capture_graph(Forest &f, bool adjoint)
{
    hipStream_t s = f.stream;
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_device<T>(f, s, adjoint, false);
    const hipError_t ec = hipStreamEndCapture(s, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ec != hipSuccess) return fail(BSA_EDEVICE, "graph capture: %s", hipGetErrorString(ec));
    hipGraphExec_t gr = nullptr;
    const hipError_t ei = hipGraphInstantiate(&gr, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(BSA_EDEVICE, "graph instantiation: %s", hipGetErrorString(ei));
    f.graph[adjoint ? 1 : 0] = gr;
    return BSA_OK;
}

int forest_prepare_graph(Forest &f, bool adjoint)
{
    if (!f.use_graph || f.graph[adjoint ? 1 : 0]) return BSA_OK;
    std::unique_lock<std::shared_mutex> lock(g_hip_mu);
    if (f.graph[adjoint ? 1 : 0]) return BSA_OK;
    HIPCHK(hipSetDevice(f.dev));
    if (adjoint) {
        const int rc = forest_setup_adjoint(f);
        if (rc) return rc;
    }
    return f.precision == BSA_FP32 ? capture_graph<float>(f, adjoint)
                                   : capture_graph<double>(f, adjoint);
}

int gene_stage(Gene &g, Forest &f, const double *x, double kappa, double p0, double p1, double w0,
               double w2, const double *qf_fixed, bool adjoint)
{
    for (int b = 0; b < g.nbranch; b++)
        if (!std::isfinite(x[b])) return fail(BSA_EINVAL, "branch length %d not finite", b);
    if (!(kappa > 0) || !(w0 >= 0) || !(w2 >= 0) || !(p0 >= 0) || !(p1 >= 0) || !(p0 + p1 > 0) ||
        p0 + p1 > 1 + 1e-12)
        return fail(BSA_EINVAL, "parameters out of range: kappa=%g p0=%g p1=%g w0=%g w2=%g", kappa,
                    p0, p1, w0, w2);
    double *uvr = f.h_uvr + (size_t)g.gi * kGeneBlock;
    const bool same = g.have_eigen && kappa == g.cache_kappa && w0 == g.cache_w0 &&
                      w2 == g.cache_w2;
    if (same) {
        // eigen systems unchanged: only proportions and Qfactors are recomputed
        if (!setup_model_a(kappa, p0, p1, w0, w2, g.pi.data(), &g.model, false))
            return fail(BSA_ENUMERIC, "model set-up failed");
    } else {
        if (!setup_model_a(kappa, p0, p1, w0, w2, g.pi.data(), &g.model, true, f.G == 1)) {
            g.have_eigen = false;
            return fail(BSA_ENUMERIC, "eigen decomposition failed");
        }
        g.have_eigen = true;
        g.cache_kappa = kappa;
        g.cache_w0 = w0;
        g.cache_w2 = w2;
        // padded U, V, root per eigen set
        for (int s = 0; s < 3; s++) {
            double *U = uvr + s * kUVRSet, *V = U + 4096, *R = U + 8192;
            std::memset(U, 0, sizeof(double) * kUVRSet);
            for (int i = 0; i < kS; i++)
                for (int k = 0; k < kS; k++) {
                    U[i * 64 + k] = g.model.U[s][i * kS + k];
                    V[i * 64 + k] = g.model.V[s][i * kS + k];
                }
            for (int k = 0; k < kS; k++) R[k] = g.model.root[s][k];
        }
    }
    if (qf_fixed) {
        g.model.Qfactor[0] = qf_fixed[0];
        g.model.Qfactor[1] = qf_fixed[1];
    }
    double *fk = uvr + 3 * kUVRSet;
    for (int c = 0; c < 4; c++) fk[c] = g.model.freqK[c];
    PmatJob *pj = f.h_pjobs + g.pj_base;
    for (size_t i = 0; i < g.pjobs.size(); i++) {
        const int v = g.pjobs[i].node;
        pj[i].t = x[g.ibranch[v]] * g.model.Qfactor[g.label[v]];
    }
    if (adjoint) {
        EigJob *ej = f.h_ejobs + g.ej_base;
        for (size_t i = 0; i < g.ejobs.size(); i++) {
            const int v = g.ejob_node[i];
            ej[i].tau = x[g.ibranch[v]] * g.model.Qfactor[g.label[v]];
        }
    }
    return BSA_OK;
}

int gene_stage_x(Gene &g, Forest &f, const double *x, bool adjoint)
{
    double kappa, p0, p1, w0, w2;
    natural_params(g, x, &kappa, &p0, &p1, &w0, &w2);
    return gene_stage(g, f, x, kappa, p0, p1, w0, w2, nullptr, adjoint);
}

double gene_collect(Gene &g, const Forest &f)
{
    g.last_lnl = f.h_out[(size_t)g.gi * kRootSums];
    g.evaluated = true;
    return g.last_lnl;
}

void gene_chain_rule(const Gene &g, const Forest &f, const double *x, double *grad)
{
    const size_t nej = g.ejobs.size();
    const double *dtau = f.h_adjout + g.ej_base;
    const double *Z = f.h_adjout + f.ejobs.size() + (size_t)3 * g.gi * 4096;
    std::vector<double> dt_node(g.nnode, 0.0);
    for (size_t i = 0; i < nej; i++) dt_node[g.ejob_node[i]] += dtau[i];
    double dQf[2] = {0, 0};
    const int nb = g.nbranch;
    for (int v = 0; v < g.nnode; v++) {
        if (v == g.root) continue;
        const int b = g.ibranch[v], l = g.label[v];
        dQf[l] += x[b] * dt_node[v];
        grad[b] = -g.model.Qfactor[l] * dt_node[v];
    }
    const double *o = f.h_out + (size_t)g.gi * kRootSums;
    const double Sf[4] = {o[1], o[2], o[3], o[4]};
    ModelGrad mg;
    model_a_param_grad(g.model, g.pi.data(), Z, dQf, Sf, &mg);
    grad[nb] = -mg.kappa;
    if (g.opts.p_transformed) {
        const double p0 = g.model.p0, p1 = g.model.p1;
        grad[nb + 1] = -(mg.p0 * p0 * (1 - p0) - mg.p1 * p0 * p1);
        grad[nb + 2] = -(-mg.p0 * p0 * p1 + mg.p1 * p1 * (1 - p1));
    } else {
        grad[nb + 1] = -mg.p0;
        grad[nb + 2] = -mg.p1;
    }
    grad[nb + 3] = -mg.w0;
    if (!g.opts.fix_omega) grad[nb + 4] = -mg.w2;
}

}  // namespace bsa

namespace bsa {

int forest_set_shapes(Forest &f, int tile, int tpb, int tpc)
{
    HIP_SHARED;
    if (tile != 64 && tile != 128) return fail(BSA_EINVAL, "tile %d: 64 or 128 patterns per block", tile);
    if (tile == 128 && f.kernel == BSA_KERNEL_VALU)
        return fail(BSA_EINVAL, "128-pattern tiles need the MFMA kernel");
    if (tpb < 0 || tpb > 8) return fail(BSA_EINVAL, "tiles per child-adjoint block %d: 0..8", tpb);
    if (tpc < 0 || tpc > kAdjTilesPerChunk)
        return fail(BSA_EINVAL, "tiles per branch-matrix chunk %d: 0..%d", tpc, kAdjTilesPerChunk);
    for (auto &gr : f.graph)
        if (gr) {
            HIPCHK(hipGraphExecDestroy(gr));
            gr = nullptr;
        }
    f.tile = tile;
    f.tpb_set = tpb;
    f.tpc_set = tpc;
    if (!f.adj_ready) return BSA_OK;   // the chunking is set when the adjoint is
    const int ntile = f.npad / 64;
    f.tpc = tpc > 0 ? tpc
                    : std::max(1, std::min(kAdjTilesPerChunkAuto,
                                           ntile * (int)f.adj.size() / 1024));
    f.nchunk = (ntile + f.tpc - 1) / f.tpc;
    if (f.nchunk > f.mpart_chunks) {
        HIPCHK(hipSetDevice(f.dev));
        HIPCHK(hipStreamSynchronize(f.stream));
        HIPCHK(hipFree(f.d_Mpart));
        f.d_Mpart = nullptr;
        f.mpart_chunks = 0;
        HIPCHK(hipMalloc(&f.d_Mpart, sizeof(double) * 4096 * (size_t)f.nmrow * f.nchunk));
        f.mpart_chunks = f.nchunk;
    }
    return forest_check(f);
}

int forest_set_prune(Forest &f, int mode)
{
    HIP_SHARED;
    if (mode != BSA_PRUNE_LEVELS && mode != BSA_PRUNE_WALK)
        return fail(BSA_EINVAL, "pruning mode %d: BSA_PRUNE_LEVELS or BSA_PRUNE_WALK", mode);
    if (mode == BSA_PRUNE_WALK && f.kernel != BSA_KERNEL_MFMA && f.kernel != BSA_KERNEL_MFMA32)
        return fail(BSA_EINVAL, "the persistent pruning needs a kernel with FP64 matrix-core "
                                "products (BSA_KERNEL_MFMA or BSA_KERNEL_MFMA32)");
    if (mode != f.prune_walk)
        for (auto &gr : f.graph)
            if (gr) {
                HIPCHK(hipGraphExecDestroy(gr));
                gr = nullptr;
            }
    f.prune_walk = mode;
    return BSA_OK;
}

int forest_tune(Forest &f, const std::function<int()> &eval, int reps, double *out)
{
    if (reps < 1) reps = 3;
    auto device_ms = [&]() {
        double t = 0;
        for (int i = 0; i < 6; i++) t += f.timing[i];
        return t;
    };
    auto measure = [&](double *ms) {
        double best = 1e300;
        for (int r = 0; r < reps; r++) {
            const int rc = eval();
            if (rc) return rc;
            best = std::min(best, device_ms());
        }
        *ms = best;
        return BSA_OK;
    };
    int cur[3] = {f.tile, f.tpb_set, f.tpc_set};
    int rc = eval();   // first call: adjoint set-up, graph capture
    if (rc) return rc;
    double best;
    if ((rc = measure(&best))) return rc;
    const double start = best;
    // pruning by level launches or by the persistent walk (FP64-product
    // kernels; the walk pays where many genes' tiles fill the GPU)
    if (f.kernel == BSA_KERNEL_MFMA || f.kernel == BSA_KERNEL_MFMA32) {
        const int keep = f.prune_walk;
        if ((rc = forest_set_prune(f, 1 - keep))) return rc;
        double ms;
        if ((rc = measure(&ms))) return rc;
        if (ms < best)
            best = ms;
        else if ((rc = forest_set_prune(f, keep)))
            return rc;
    }
    const std::vector<int> cand[3] = {
        f.kernel != BSA_KERNEL_VALU ? std::vector<int>{64, 128} : std::vector<int>{64},
        {0, 1, 2, 4, 8},
        {0, 4, 8, 16, 32, 40, 64}};
    for (int k = 0; k < 3; k++) {
        int keep = cur[k];
        for (int v : cand[k]) {
            if (v == keep) continue;
            int shp[3] = {cur[0], cur[1], cur[2]};
            shp[k] = v;
            if ((rc = forest_set_shapes(f, shp[0], shp[1], shp[2]))) return rc;
            double ms;
            if ((rc = measure(&ms))) return rc;
            if (ms < best) {
                best = ms;
                cur[k] = v;
            }
        }
        if ((rc = forest_set_shapes(f, cur[0], cur[1], cur[2]))) return rc;
    }
    if (out) {
        out[0] = f.tile;
        out[1] = f.tpb_set;
        out[2] = f.tpc;
        out[3] = f.nchunk;
        out[4] = best;
        out[5] = start;
        out[6] = f.prune_walk;
    }
    return BSA_OK;
}

}  // namespace bsa
