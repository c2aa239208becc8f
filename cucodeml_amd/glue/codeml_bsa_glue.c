/*
 * Reference-side binding of the MI355X engine: the single-GPU replacement of
 * cuda-codeml.cu for the reference codeml host program.  It is compiled
 * together with the reference's own codeml.c and tools.c (built with -DCUDA,
 * which turns on the CUDA_* hooks at codeml.c:683-696 and treesub.c:7396-7594)
 * and linked against libbsa.so; the reference host code is unchanged.
 *
 * Entry points replaced (declared in the reference's codeml.h:67-85):
 *   CUDA_Init(x)                       cuda-codeml.cu:1106
 *   CUDA_FreeAll()                     cuda-codeml.cu:1180
 *   CUDA_ScheduleConditionalPNode(ir, ig, x)   cuda-codeml.cu:1336
 * CUDA_lfundG / CUDA_fx_r stay the reference's own (treesub.c:7397, 7554):
 * they call CUDA_ScheduleConditionalPNode once per site class and finish the
 * mixture on the host from com.fhK.  The engine evaluates all four classes
 * in one pass at the first class and hands the rows out from there.
 *
 * CUDA_dfun is the analytic gradient for ming2's dfun slot (tools.c:6361,
 * 6408, 6461); the reference passes NULL there (codeml.c:688).  When this
 * file is linked with -Wl,--wrap=ming2 and BSA_DFUN=1 is set in the
 * environment, __wrap_ming2 inserts it without touching the reference code.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "paml.h"
#include "codeml.h"
#include "bsa.h"

extern int LASTROUND;

static bsa_engine *g_eng = NULL;
static double *g_fh = NULL;      /* 4 x npatt per-class log f */

static void die(const char *what)
{
    fprintf(stderr, "\nbsa: %s: %s\n", what, bsa_last_error());
    exit(-1);
}

/* a control-file setting the engine does not compute: stop before any
 * device call, naming the key */
static void refuse(const char *key, int value, const char *why)
{
    fprintf(stderr, "\nbsa: %s = %d is not offloaded: %s\n", key, value, why);
    exit(-1);
}

static void sync_opts(void)
{
    bsa_model_opts o;
    o.fix_omega = com.fix_omega;
    o.omega_fix = com.omega_fix;
    o.p_transformed = !LASTROUND;
    if (bsa_set_model(g_eng, &o)) die("bsa_set_model");
}

void CUDA_Init(double *x)
{
    int i, j, k, h, ncodes = 61;
    int nn = tree.nnode, ns = com.ns, npatt = com.npatt;
    int *father = malloc(sizeof(int) * nn), *label = malloc(sizeof(int) * nn);
    int *ibranch = malloc(sizeof(int) * nn);
    unsigned char *z, *mask;
    bsa_problem p;
    const char *prec = getenv("BSA_PRECISION");
    (void)x;
    if (com.seqtype != 1 || com.model != 2 || com.NSsites != 2)
        die("only branch-site model A (seqtype=1, model=2, NSsites=2) is offloaded");
    /* the engine's states are the 61 sense codons of the universal code in
     * TCAG order; another code numbers its codons differently */
    if (com.icode != 0)
        refuse("icode", com.icode, "only the universal genetic code (icode = 0) is implemented");
    if (com.ncode != 61) refuse("ncode", com.ncode, "the engine has 61 sense codons");
    if (com.ngene > 1) refuse("ngene", com.ngene, "one partition (Mgene = 0, no G option)");
    if (com.clock) refuse("clock", com.clock, "only clock = 0 is implemented");
    if (com.fix_blength == 2)
        refuse("fix_blength", 2, "x carries no branch lengths; the engine expects them");
    z = malloc((size_t)ns * npatt);
    for (i = 0; i < nn; i++) {
        father[i] = nodes[i].father;
        label[i] = i == tree.root ? 0 : (int)nodes[i].label;
        ibranch[i] = nodes[i].ibranch;
    }
    for (j = 0; j < ns; j++)
        for (h = 0; h < npatt; h++) {
            z[(size_t)j * npatt + h] = com.z[j][h];
            if (com.z[j][h] + 1 > ncodes) ncodes = com.z[j][h] + 1;
        }
    /* ambiguity codes resolve through CharaMap (SetMapAmbiguity, treesub.c:1146) */
    mask = calloc((size_t)ncodes * 61, 1);
    for (i = 0; i < ncodes; i++) {
        if (i < 61) { mask[i * 61 + i] = 1; continue; }
        for (k = 0; k < nChara[i]; k++) mask[i * 61 + CharaMap[i][k]] = 1;
    }
    memset(&p, 0, sizeof p);
    p.ns = ns;
    p.nnode = nn;
    p.root = tree.root;
    p.father = father;
    p.label = label;
    p.ibranch = ibranch;
    p.npatt = npatt;
    p.z = z;
    p.ncodes = ncodes;
    p.code_mask = mask;
    p.fpatt = com.fpatt;
    p.pi = com.pi;
    p.precision = (prec && !strcmp(prec, "fp32")) ? BSA_FP32 : BSA_FP64;
    p.kernel = BSA_KERNEL_MFMA;
    p.device = 0;
    if (bsa_create(&p, &g_eng)) die("bsa_create");
    free(father); free(label); free(ibranch); free(z); free(mask);
    g_fh = malloc(sizeof(double) * 4 * npatt);
    sync_opts();
}

void CUDA_FreeAll(void)
{
    bsa_destroy(g_eng);
    g_eng = NULL;
    free(g_fh);
    g_fh = NULL;
}

/* com.fhK[ir][h]: log f(x_h | ir) when node scaling is on, f itself when off
 * (fx_r, treesub.c:7518-7538) */
void CUDA_ScheduleConditionalPNode(int ir, int ig, double *x)
{
    int h, npatt = com.npatt;
    double negl;
    (void)ig;
    if (ir == 0) {
        sync_opts();
        if (bsa_lnl(g_eng, x, com.np, &negl)) die("bsa_lnl");
        if (bsa_class_lnl(g_eng, g_fh, NULL)) die("bsa_class_lnl");
    }
    for (h = 0; h < npatt; h++) {
        double v = g_fh[ir * npatt + h];
        com.fhK[ir * npatt + h] = com.NnodeScale ? v : exp(v);
    }
}

/* ming2's dfun: f = -lnL and dx = d(-lnL)/dx at x */
int CUDA_dfun(double x[], double *f, double dx[], int n)
{
    sync_opts();
    if (bsa_lnl_grad(g_eng, x, n, f, dx)) die("bsa_lnl_grad");
    return 0;
}

int __real_ming2(FILE *fout, double *f, double (*fun)(double x[], int n),
                 int (*dfun)(double x[], double *f, double dx[], int n), double x[],
                 double xb[][2], double space[], double e, int n);

int __wrap_ming2(FILE *fout, double *f, double (*fun)(double x[], int n),
                 int (*dfun)(double x[], double *f, double dx[], int n), double x[],
                 double xb[][2], double space[], double e, int n)
{
    const char *s = getenv("BSA_DFUN");
    if (!dfun && g_eng && s && s[0] == '1') dfun = CUDA_dfun;
    return __real_ming2(fout, f, fun, dfun, x, xb, space, e, n);
}
