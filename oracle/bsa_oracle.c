/*
 * ORACLE — test infrastructure only.  Never linked into, loaded by or called
 * from the product library; only tests/, the smoke check and the CPU-baseline
 * leg of bench.py use it, as the checker.
 *
 * A plain-C, double-precision restatement of the reference's likelihood of
 * branch-site model A (codeml model=2, NSsites=2), written from the
 * algorithm's description in the reference sources:
 *
 *   orc_q          eigenQcodon Q construction      codeml.c:3147-3186
 *   orc_eigen      eigenQREV (S = Pi^1/2 Q Pi^-1/2, codons with pi=0 split
 *                  off, U = Pi^-1/2 R, V = R' Pi^1/2)  tools.c:4850-4937;
 *                  the symmetric eigensolver is a cyclic Jacobi iteration
 *                  here instead of Householder + QL (tools.c:4943-5134):
 *                  both give the same decomposition to rounding
 *   orc_pmat       PMatUVRoot                      tools.c:515-545
 *   orc_model      SetParametersNSsites, model A   codeml.c:2470-2520
 *                  (class proportions, two Qfactors, three U/V/Root sets)
 *   orc_lnl        fx_r + ConditionalPNode + NodeScale + lfundG
 *   orc_lnl_qf     the same with the Qfactors given (BEB, treesub.c:7289-7301)
 *                  treesub.c:7485-7550, codeml.c:3402-3458,
 *                  treesub.c:6917-6968, treesub.c:7342-7394
 *   orc_scale_marks SetNodeScale (every=15 for codons) treesub.c:6917-6937
 *
 * Parity of every function here is pinned in tests/: against lnL and
 * per-site log f values written by the reference's own CPU codeml on
 * seeded synthetic data, against scipy/numpy (expm, eigh), and against
 * brute-force sums over ancestral states on small trees.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define N61 61

static const char AA64[] =
    "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
static int from61[N61];
static int tables_ready = 0;

static void init_tables(void)
{
    int i, k = 0;
    if (tables_ready) return;
    for (i = 0; i < 64; i++)
        if (AA64[i] != '*') from61[k++] = i;
    tables_ready = 1;
}

/* GY94 rate matrix and mean rate; codeml.c:3147-3186 (HKY form, universal code). */
void orc_q(double kappa, double omega, const double *pi, double *Q, double *mr)
{
    int i, j, k, n = N61;
    init_tables();
    memset(Q, 0, sizeof(double) * n * n);
    for (i = 1; i < n; i++) {
        int c1 = from61[i], f[3] = {c1 / 16, (c1 / 4) % 4, c1 % 4};
        for (j = 0; j < i; j++) {
            int c2 = from61[j], t[3] = {c2 / 16, (c2 / 4) % 4, c2 % 4};
            int nd = 0, pos = 0;
            double q = 1;
            for (k = 0; k < 3; k++)
                if (f[k] != t[k]) { nd++; pos = k; }
            if (nd != 1) continue;
            if (f[pos] + t[pos] == 1 || f[pos] + t[pos] == 5) q = kappa;
            if (AA64[c1] != AA64[c2]) q *= omega;
            Q[i * n + j] = q * pi[j];
            Q[j * n + i] = q * pi[i];
        }
    }
    *mr = 0;
    for (i = 0; i < n; i++) {
        double s = 0;
        for (j = 0; j < n; j++) s += Q[i * n + j];
        Q[i * n + i] = -s;
    }
    for (i = 0; i < n; i++) *mr -= pi[i] * Q[i * n + i];
}

/* cyclic Jacobi on a symmetric m x m matrix A (destroyed); eigenvectors in
   columns of R, eigenvalues in d, sorted descending. */
static void jacobi_sym(double *A, int m, double *d, double *R)
{
    int i, j, p, q, sweep;
    for (i = 0; i < m; i++)
        for (j = 0; j < m; j++) R[i * m + j] = (i == j);
    for (sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (p = 0; p < m; p++)
            for (q = 0; q < m; q++) {
                tot += A[p * m + q] * A[p * m + q];
                if (p != q) off += A[p * m + q] * A[p * m + q];
            }
        if (off <= 1e-34 * tot || off == 0) break;
        for (p = 0; p < m - 1; p++)
            for (q = p + 1; q < m; q++) {
                double apq = A[p * m + q], app, aqq, theta, t, c, s;
                if (fabs(apq) < 1e-300) continue;
                app = A[p * m + p];
                aqq = A[q * m + q];
                theta = (aqq - app) / (2 * apq);
                t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                c = 1 / sqrt(t * t + 1);
                s = t * c;
                for (i = 0; i < m; i++) {
                    double aip = A[i * m + p], aiq = A[i * m + q];
                    A[i * m + p] = c * aip - s * aiq;
                    A[i * m + q] = s * aip + c * aiq;
                }
                for (i = 0; i < m; i++) {
                    double api = A[p * m + i], aqi = A[q * m + i];
                    A[p * m + i] = c * api - s * aqi;
                    A[q * m + i] = s * api + c * aqi;
                }
                for (i = 0; i < m; i++) {
                    double rip = R[i * m + p], riq = R[i * m + q];
                    R[i * m + p] = c * rip - s * riq;
                    R[i * m + q] = s * rip + c * riq;
                }
            }
    }
    for (i = 0; i < m; i++) d[i] = A[i * m + i];
    /* selection sort, descending (EigenSort, tools.c:4963) */
    for (i = 0; i < m - 1; i++) {
        int kmax = i;
        for (j = i + 1; j < m; j++)
            if (d[j] > d[kmax]) kmax = j;
        if (kmax != i) {
            double t = d[i];
            d[i] = d[kmax];
            d[kmax] = t;
            for (j = 0; j < m; j++) {
                t = R[j * m + i];
                R[j * m + i] = R[j * m + kmax];
                R[j * m + kmax] = t;
            }
        }
    }
}

/* eigenQREV restated: Q = U diag(Root) V with UV = I; tools.c:4850-4937.
   U, V row-major n x n (stride 61). */
void orc_eigen(const double *Q, const double *pi, double *Root, double *U, double *V)
{
    int n = N61, i, j, k, m = 0, idx[N61];
    double sq[N61], A[N61 * N61], R[N61 * N61], d[N61];
    for (i = 0; i < n; i++)
        if (pi[i] > 1e-100) { idx[m] = i; sq[m] = sqrt(pi[i]); m++; }
    for (i = 0; i < m; i++)
        for (j = 0; j < m; j++)
            A[i * m + j] = (i == j) ? Q[idx[i] * n + idx[i]]
                         : Q[idx[i] * n + idx[j]] * sq[i] / sq[j];
    /* symmetrise exactly as the reference does: lower triangle wins */
    for (i = 0; i < m; i++)
        for (j = 0; j < i; j++) A[j * m + i] = A[i * m + j];
    jacobi_sym(A, m, d, R);
    for (i = 0; i < n * n; i++) { U[i] = 0; V[i] = 0; }
    for (i = 0; i < n; i++) Root[i] = 0;
    /* reduced-space eigenpairs go to the positions of the kept codons, in
       descending order; dropped codons keep identity rows and Root 0. */
    for (k = 0; k < m; k++) Root[idx[k]] = d[k];
    for (i = 0; i < n; i++)
        if (pi[i] <= 1e-100) { U[i * n + i] = 1; V[i * n + i] = 1; }
    for (i = 0; i < m; i++)
        for (k = 0; k < m; k++) {
            U[idx[i] * n + idx[k]] = R[i * m + k] / sq[i];
            V[idx[k] * n + idx[i]] = R[i * m + k] * sq[i];
        }
}

/* P(t) = U diag(exp(Root t)) V, negatives to 0, identity for t < 1e-100
   (PMatUVRoot, tools.c:515-545). */
void orc_pmat(const double *U, const double *V, const double *Root, double t, double *P)
{
    int n = N61, i, j, k;
    if (t < 1e-100) {
        for (i = 0; i < n * n; i++) P[i] = 0;
        for (i = 0; i < n; i++) P[i * n + i] = 1;
        return;
    }
    for (i = 0; i < n * n; i++) P[i] = 0;
    for (k = 0; k < n; k++) {
        double e = exp(t * Root[k]);
        for (i = 0; i < n; i++) {
            double ue = U[i * n + k] * e;
            for (j = 0; j < n; j++) P[i * n + j] += ue * V[k * n + j];
        }
    }
    for (i = 0; i < n * n; i++)
        if (P[i] < 0) P[i] = 0;
}

/* Model A set-up (SetParametersNSsites, codeml.c:2470-2520).
   freqK[4], Qfactor[2] (background, foreground), UVR sets for w0, 1, w2. */
void orc_model(double kappa, double p0, double p1, double w0, double w2, const double *pi,
               double *freqK, double *Qfactor, double *Root3, double *U3, double *V3)
{
    double Q[N61 * N61], mr, t = p0 + p1, w[3];
    int i;
    freqK[0] = p0;
    freqK[1] = p1;
    freqK[2] = (1 - t) * p0 / t;
    freqK[3] = (1 - t) * p1 / t;
    orc_q(kappa, (p0 / t) * w0 + (p1 / t) * 1.0, pi, Q, &mr);
    Qfactor[0] = 1 / mr;
    orc_q(kappa, p0 * w0 + p1 * 1.0 + (1 - t) * w2, pi, Q, &mr);
    Qfactor[1] = 1 / mr;
    w[0] = w0; w[1] = 1; w[2] = w2;
    for (i = 0; i < 3; i++) {
        orc_q(kappa, w[i], pi, Q, &mr);
        orc_eigen(Q, pi, Root3 + i * N61, U3 + i * N61 * N61, V3 + i * N61 * N61);
    }
}

/* class -> U/V/Root set (Set_UVR_BranchSite, codeml.c:2555-2558) */
static int iuvr(int iclass, int label)
{
    if (label == 0) return iclass % 2;
    return iclass <= 1 ? iclass : 2;
}

/* SetNodeScale: mark a node when more than 15 tips (since the last marked
   node below) hang under it; root never marked.  Returns tip count. */
static int mark_scale(int v, int root, const int *nson, const int *sons, int maxson,
                      char *mark, int *nmark)
{
    int i, d = 0;
    for (i = 0; i < nson[v]; i++) {
        int s = sons[v * maxson + i];
        d += nson[s] ? mark_scale(s, root, nson, sons, maxson, mark, nmark) : 1;
    }
    if (v != root && d > 15) { mark[v] = 1; d = 1; (*nmark)++; }
    return d;
}

int orc_scale_marks(int nnode, int root, const int *nson, const int *sons, int maxson,
                    char *mark)
{
    int nm = 0;
    memset(mark, 0, nnode);
    mark_scale(root, root, nson, sons, maxson, mark, &nm);
    return nm;
}

typedef struct {
    int ns, nnode, root, maxson, npatt, ncodes;
    const int *nson, *sons, *label;
    const double *branch;            /* per node */
    const unsigned char *z;          /* ns x npatt */
    const double *codemask;          /* ncodes x 61 */
    double *conP;                    /* nnode x npatt x 61 */
    double *P;                       /* scratch 61 x 61 */
    const char *mark;
    double *scaleF;                  /* nscale x npatt */
    const int *scale_k;
    const double *Qfactor, *U3, *V3, *Root3;
    int iclass;
} orc_ctx;

/* ConditionalPNode restated (codeml.c:3402-3458) for cleandata=0, followed
   by NodeScale (treesub.c:6940-6968) at marked nodes. */
static void cond_node(orc_ctx *c, int v)
{
    int i, j, k, h, n = N61, np = c->npatt;
    double *L = c->conP + (size_t)v * np * n;
    for (i = 0; i < c->nson[v]; i++) {
        int s = c->sons[v * c->maxson + i];
        if (c->nson[s] > 0) cond_node(c, s);
    }
    for (h = 0; h < np * n; h++) L[h] = 1;
    for (i = 0; i < c->nson[v]; i++) {
        int s = c->sons[v * c->maxson + i], lab = c->label[s];
        int set = iuvr(c->iclass, lab);
        double t = c->branch[s] * c->Qfactor[lab];
        orc_pmat(c->U3 + set * n * n, c->V3 + set * n * n, c->Root3 + set * n, t, c->P);
        if (c->nson[s] < 1) {
            for (h = 0; h < np; h++) {
                int code = c->z[(size_t)s * np + h];
                const double *mk = c->codemask + (size_t)code * n;
                for (j = 0; j < n; j++) {
                    double acc = 0;
                    for (k = 0; k < n; k++)
                        if (mk[k] != 0) acc += c->P[j * n + k];
                    L[h * n + j] *= acc;
                }
            }
        } else {
            const double *Ls = c->conP + (size_t)s * np * n;
            for (h = 0; h < np; h++)
                for (j = 0; j < n; j++) {
                    double acc = 0;
                    for (k = 0; k < n; k++) acc += c->P[j * n + k] * Ls[h * n + k];
                    L[h * n + j] *= acc;
                }
        }
    }
    if (c->mark[v]) {
        int kk = c->scale_k[v];
        for (h = 0; h < np; h++) {
            double t = 0;
            for (j = 0; j < n; j++)
                if (L[h * n + j] > t) t = L[h * n + j];
            if (t < 1e-300) {
                for (j = 0; j < n; j++) L[h * n + j] = 1;
                c->scaleF[(size_t)kk * np + h] = -800;
            } else {
                for (j = 0; j < n; j++) L[h * n + j] /= t;
                c->scaleF[(size_t)kk * np + h] = log(t);
            }
        }
    }
}

/*
 * Full model-A log-likelihood at fixed parameters.
 * Tree: nson[nnode], sons[nnode*maxson], label[nnode] (0/1), branch[nnode]
 * (length of the edge above each node; root ignored); tips are 0..ns-1.
 * Data: z[ns*npatt] codes (0..60 sense, >=61 ambiguity rows of codemask),
 * fpatt[npatt], pi[61].  Outputs (optional): fhK[4*npatt] per-class
 * log f(x_h | class) and site_lnf[npatt] mixture log f_h.  Returns lnL.
 */
double orc_lnl_qf(int ns, int nnode, int root, int maxson, const int *nson, const int *sons,
                  const int *label, const double *branch, int npatt, const unsigned char *z,
                  int ncodes, const double *codemask, const double *fpatt, const double *pi,
                  double kappa, double p0, double p1, double w0, double w2, int use_scaling,
                  const double *qf_fixed, double *fhK, double *site_lnf);

double orc_lnl(int ns, int nnode, int root, int maxson, const int *nson, const int *sons,
               const int *label, const double *branch, int npatt, const unsigned char *z,
               int ncodes, const double *codemask, const double *fpatt, const double *pi,
               double kappa, double p0, double p1, double w0, double w2, int use_scaling,
               double *fhK, double *site_lnf)
{
    return orc_lnl_qf(ns, nnode, root, maxson, nson, sons, label, branch, npatt, z, ncodes,
                      codemask, fpatt, pi, kappa, p0, p1, w0, w2, use_scaling, NULL, fhK,
                      site_lnf);
}

/* As orc_lnl; qf_fixed (if not NULL) replaces the two Qfactors, as the
 * reference's BEB does with the MLE's (GetPMatBranch, treesub.c:7289-7301). */
double orc_lnl_qf(int ns, int nnode, int root, int maxson, const int *nson, const int *sons,
                  const int *label, const double *branch, int npatt, const unsigned char *z,
                  int ncodes, const double *codemask, const double *fpatt, const double *pi,
                  double kappa, double p0, double p1, double w0, double w2, int use_scaling,
                  const double *qf_fixed, double *fhK, double *site_lnf)
{
    orc_ctx c;
    double freqK[4], Qf[2], *U3, *V3, R3[3 * N61], lnL = 0;
    double *conP, *scaleF, *fh;
    char *mark;
    int *scale_k, nscale = 0, i, ir, h, v, n = N61;

    U3 = malloc(sizeof(double) * 3 * n * n);
    V3 = malloc(sizeof(double) * 3 * n * n);
    orc_model(kappa, p0, p1, w0, w2, pi, freqK, Qf, R3, U3, V3);
    if (qf_fixed) {
        Qf[0] = qf_fixed[0];
        Qf[1] = qf_fixed[1];
    }

    mark = calloc(nnode, 1);
    scale_k = calloc(nnode, sizeof(int));
    if (use_scaling) nscale = orc_scale_marks(nnode, root, nson, sons, maxson, mark);
    for (v = 0, i = 0; v < nnode; v++)
        if (mark[v]) scale_k[v] = i++;
    conP = malloc(sizeof(double) * (size_t)nnode * npatt * n);
    scaleF = calloc((size_t)(nscale ? nscale : 1) * npatt, sizeof(double));
    fh = malloc(sizeof(double) * 4 * npatt);

    c.ns = ns; c.nnode = nnode; c.root = root; c.maxson = maxson; c.npatt = npatt;
    c.ncodes = ncodes; c.nson = nson; c.sons = sons; c.label = label; c.branch = branch;
    c.z = z; c.codemask = codemask; c.conP = conP; c.P = malloc(sizeof(double) * n * n);
    c.mark = mark; c.scaleF = scaleF; c.scale_k = scale_k; c.Qfactor = Qf;
    c.U3 = U3; c.V3 = V3; c.Root3 = R3;

    for (ir = 0; ir < 4; ir++) {
        const double *Lr = conP + (size_t)root * npatt * n;
        c.iclass = ir;
        cond_node(&c, root);
        for (h = 0; h < npatt; h++) {
            double f = 0;
            int k;
            for (i = 0; i < n; i++) f += pi[i] * Lr[h * n + i];
            if (f <= 0) f = 1e-300;
            if (nscale) {
                f = log(f);
                for (k = 0; k < nscale; k++) f += scaleF[(size_t)k * npatt + h];
            }
            fh[ir * npatt + h] = f;
        }
    }
    /* mixture over classes (lfundG, treesub.c:7364-7393) */
    for (h = 0; h < npatt; h++) {
        double f = 0, l;
        if (nscale) {
            int it = 0;
            for (ir = 1; ir < 4; ir++)
                if (fh[ir * npatt + h] > fh[it * npatt + h]) it = ir;
            for (ir = 0; ir < 4; ir++) f += freqK[ir] * exp(fh[ir * npatt + h] - fh[it * npatt + h]);
            l = fh[it * npatt + h] + log(f);
        } else {
            for (ir = 0; ir < 4; ir++) f += freqK[ir] * fh[ir * npatt + h];
            if (f <= 0) f = 1e-300;
            l = log(f);
        }
        if (site_lnf) site_lnf[h] = l;
        if (fpatt[h] > 0) lnL += l * fpatt[h];
    }
    if (fhK) {
        for (ir = 0; ir < 4; ir++)
            for (h = 0; h < npatt; h++)
                fhK[ir * npatt + h] = nscale ? fh[ir * npatt + h] : log(fh[ir * npatt + h]);
    }
    free(U3); free(V3); free(mark); free(scale_k); free(conP); free(scaleF); free(fh); free(c.P);
    return lnL;
}
