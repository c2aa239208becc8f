"""codeml-compatible driver for branch-site model A on the engine.

Reads a codeml control file (the subset the model-A path uses: seqfile,
treefile, outfile, seqtype = 1, model = 2, NSsites = 2, CodonFreq = 2,
kappa, fix_omega, omega, cleandata = 0, ndata, getSE, Small_Diff,
fix_blength 0, 1 or 2; codeml.c:1430-1768), the PHYLIP codon alignment(s)
and the Newick tree with ``#1`` on the foreground branch, fits the model by
maximum likelihood on the GPU (bsa_fit, in place of ming2 at codeml.c:688)
and writes the estimates in the layout of codeml's ``out`` file (lnL line
codeml.c:738-741, parameter vector 751, site-class table 989-1001), the
per-pattern log f of the final evaluation in the layout of its ``lnf``
file (codeml.c:596-597, 613, 832-834; treesub.c:7332-7337) and, under H1,
the site posteriors in ``rst``.

With fix_blength = 2 the branch lengths are held at the tree file's values
and only kappa, the p's and the omegas are estimated (treesub.c:3729-3730,
codeml.c:617-627); 0 and 1 estimate them (the tree's lengths, 0.1 where
absent, are the start); -1 (random starts) and 3 are refused.

Starting values: an ``in.codeml`` file beside the control file is read as
codeml reads it (treesub.c:3925-3927: the parameter vector in codeml's
optimisation variables, or, after a leading -1, fixed values that are only
evaluated); otherwise branch lengths come from the tree (0.1 where absent),
p0 = 0.6, p1 = 0.2, w0 = 0.3, kappa and w2 from the control file.

    python -m cucodeml_amd.codeml codeml.ctl [--precision fp32] [--fd]

The likelihood-ratio test of H0 (w2 = 1 fixed) against H1 is ``lrt``:
2 dlnL with its p-values against chi2_1 and the 50:50 mixture of 0 and chi2_1.
"""
import argparse
import math
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import engine, formats


def read_ctl(path):
    """key = value pairs of a codeml control file (* starts a comment)."""
    out = {}
    with open(path) as f:
        for line in f:
            line = line.split("*", 1)[0].strip()
            if "=" not in line:
                continue
            k, v = line.split("=", 1)
            out[k.strip()] = v.strip().split()[0] if v.strip() else ""
    return out


def _num(ctl, key, default):
    v = ctl.get(key)
    return default if v is None or v == "" else float(v)


def check_ctl(ctl):
    """Reject settings outside the path the engine implements."""
    need = {"seqtype": 1, "model": 2, "NSsites": 2}
    for k, v in need.items():
        if int(_num(ctl, k, v)) != v:
            raise ValueError(f"{k} = {ctl[k]}: only branch-site model A (seqtype = 1, model = 2, "
                             f"NSsites = 2) is implemented")
    if int(_num(ctl, "CodonFreq", 2)) != 2:
        raise ValueError("only CodonFreq = 2 (F3x4) is implemented")
    if int(_num(ctl, "clock", 0)) != 0:
        raise ValueError("only clock = 0 is implemented")
    if int(_num(ctl, "cleandata", 0)) != 0:
        raise ValueError("only cleandata = 0 is implemented")
    if int(_num(ctl, "icode", 0)) != 0:
        raise ValueError("only the universal code (icode = 0) is implemented")
    if int(_num(ctl, "fix_kappa", 0)) != 0:
        raise ValueError("fix_kappa = 1 is not implemented")
    fb = int(_num(ctl, "fix_blength", 0))
    if fb not in (0, 1, 2):
        raise ValueError(f"fix_blength = {fb} is not implemented (0, 1 and 2 are)")


@dataclass
class Fit:
    hypothesis: str
    lnL: float
    x: np.ndarray                 # optimisation variables (logistic p's)
    kappa: float
    p0: float
    p1: float
    w0: float
    w2: float
    proportions: np.ndarray       # classes 0, 1, 2a, 2b
    branch: np.ndarray            # branch lengths in parameter order
    np: int
    iterations: int
    evaluations: int
    converged: bool
    seconds: float
    extra: dict = field(default_factory=dict)
    stop: str = ""                # why the optimiser stopped (engine.STOP_NAMES)
    fixed_branches: bool = False  # fix_blength = 2: branch lengths not estimated

    def x_untransformed(self):
        """x as codeml prints it: p0, p1 themselves in place of the logistic
        variables, without the branch lengths when they are fixed."""
        nb = len(self.branch)
        x = self.x.copy()
        x[nb + 1], x[nb + 2] = self.p0, self.p1
        return x[nb:] if self.fixed_branches else x


def default_start(tree, kappa=2.0, omega=1.5, fix_omega=False):
    t = tree.branch_vector().copy()
    t[t <= 0] = 0.1
    x = list(t) + [kappa, math.log(0.6 / 0.2), 0.0, 0.3]
    if not fix_omega:
        x.append(max(omega, 1.0))
    return np.array(x)


def _bounds(eng_bounds, x0, nb, fixed_branches):
    """codeml's bounds, with lo = hi = x0 on the branch lengths when they are
    held fixed (bsa_fit keeps a variable with lo == hi where it is)."""
    lo, hi = eng_bounds
    if fixed_branches:
        lo, hi = lo.copy(), hi.copy()
        lo[:nb] = hi[:nb] = x0[:nb]
    return lo, hi


# Newton iterations of the last phase of an FP64 finish
FINISH_NEWTON = 8
# Newton iterations of an FP64 finish on more free variables than
# DENSE_NEWTON_MAX, whose Newton steps come from preconditioned conjugate
# gradients (about 40 gradients each, where the Hessian costs one per
# variable): there Newton iterations alone are the cheaper finish (512 taxa
# x 50,000 codons: the FP32 estimates 0.78 below the maximum in lnL, along
# the p2 w2 ridge; 7 Newton iterations, 14 s, reach it)
FINISH_NEWTON_LARGE = 16
DENSE_NEWTON_MAX = 160


def finish_fp64(fit, x, lo, hi, e, gradient="adjoint"):
    """FP64 finish of a fit that ended at the FP32 rounding-noise floor:
    Newton iterations from there (bsa_fit_opts.newton: the Hessian from
    differences of the adjoint gradient, which lands on the maximum also
    along flat directions such as w2 on a ridge, where a BFGS restart from
    the identity would stall), and, where they do not converge, BFGS
    followed by Newton iterations, and where those do not converge either,
    FINISH_NEWTON more Newton iterations.  An FP32 chain can stop far along
    a curved, nearly flat ridge (p2 w2 ~ const: one scan gene stopped at
    w2 = 2.58 where the maximum has 1.91, 2e-4 below it in lnL); Newton
    steps follow such a ridge at a linear rate and BFGS crawls, so the
    last phase gives Newton the iterations to get there.  With more than
    DENSE_NEWTON_MAX free variables the first phase is FINISH_NEWTON_LARGE
    Newton iterations.  ``fit`` is an engine's fit method."""
    if gradient == "adjoint":
        large = np.count_nonzero(np.asarray(lo) < np.asarray(hi)) > DENSE_NEWTON_MAX
        x1, f1, i1 = fit(x, lo, hi, e=e, max_iter=-1,
                         newton=FINISH_NEWTON_LARGE if large else 3)
        if i1["converged"]:
            return x1, f1, i1
        x2, f2, i2 = fit(x1, lo, hi, e=e, newton=4)
        i2 = _merge_info(i1, i2)
        if i2["converged"]:
            return x2, f2, i2
        x3, f3, i3 = fit(x2, lo, hi, e=e, max_iter=-1, newton=FINISH_NEWTON)
        return x3, f3, _merge_info(i2, i3)
    return fit(x, lo, hi, e=e, gradient=gradient)


def finish_fp64_batch(b, xs, los, his, e):
    """``finish_fp64`` for every gene of an FP64 batch engine: Newton
    iterations for all genes, then BFGS and Newton for those that did not
    converge, then Newton alone for those that still did not, the others
    held where they are (lo = hi).  los / his: per-gene
    bounds (None entries: codeml's).  Returns (list of x, -lnL, info)."""
    n = len(xs)
    x1, f1, i1 = b.fit(xs, e=e, lo=los, hi=his, max_iter=-1, newton=3)
    redo = [i for i in range(n) if not i1[i]["converged"]]
    if not redo:
        return x1, f1, i1
    lo2 = [los[i] if i in redo else x1[i] for i in range(n)]
    hi2 = [his[i] if i in redo else x1[i] for i in range(n)]
    x2, f2, i2 = b.fit(x1, e=e, lo=lo2, hi=hi2, newton=4)
    for i in redo:
        x1[i], f1[i], i1[i] = x2[i], f2[i], _merge_info(i1[i], i2[i])
    redo = [i for i in redo if not i1[i]["converged"]]
    if not redo:
        return x1, f1, i1
    lo3 = [los[i] if i in redo else x1[i] for i in range(n)]
    hi3 = [his[i] if i in redo else x1[i] for i in range(n)]
    x3, f3, i3 = b.fit(x1, e=e, lo=lo3, hi=hi3, max_iter=-1, newton=FINISH_NEWTON)
    for i in redo:
        x1[i], f1[i], i1[i] = x3[i], f3[i], _merge_info(i1[i], i3[i])
    return x1, f1, i1


def _merge_info(a, b):
    """Info of an FP32 fit finished by an FP64 one: counts summed, the
    finishing fit's stop reason."""
    return {"iterations": a["iterations"] + b["iterations"],
            "evaluations": a["evaluations"] + b["evaluations"],
            "converged": b["converged"], "stop": b["stop"]}


def fit_model_a(tree, data, fix_omega=False, omega=1.0, start=None, precision="fp64",
                gradient="adjoint", e=1e-8, device=0, kappa=2.0, eng=None, polish=True,
                fixed_branches=False):
    """Maximum-likelihood fit of H1 (fix_omega False) or H0 (w2 fixed at
    omega).  ``start`` is in codeml's optimisation variables.  With
    ``fixed_branches`` (fix_blength = 2) the branch lengths stay at the
    start's values (the tree's, when start is None) and only kappa, the p's
    and the omegas are fitted.

    With FP32 partials the likelihood carries rounding noise of about 1e-9
    of |lnL| (stochastic rounding of every stored value), which stalls BFGS
    where the remaining gains are of that size; with ``polish`` the fit is
    then finished on an FP64 engine from the FP32 estimates, and the
    reported lnL is the FP64 one."""
    own = eng is None
    if own:
        eng = engine.BranchSiteEngine(tree, data, precision=precision, device=device,
                                      fix_omega=fix_omega, omega_fix=omega)
    else:
        eng.set_model(fix_omega, omega, True)
    nb = tree.nbranch
    try:
        if start is None:
            x0 = default_start(tree, kappa, omega if not fix_omega else 1.5, fix_omega)
            if fixed_branches:
                x0[:nb] = tree.branch_vector()
        else:
            x0 = np.asarray(start, dtype=np.float64)
        t0 = time.perf_counter()
        lo, hi = _bounds(eng.default_bounds(), x0, nb, fixed_branches)
        x, f, info = eng.fit(x0, lo, hi, e=e, gradient=gradient)
        extra = {}
        par = eng.last_params()
        if eng.precision == "fp32" and polish:
            extra = {"fp32_lnL": -f, "fp32_iterations": info["iterations"],
                     "fp32_evaluations": info["evaluations"], "fp32_stop": info["stop"]}
            e64 = engine.BranchSiteEngine(tree, data, precision="fp64", device=eng.device,
                                          fix_omega=fix_omega, omega_fix=omega)
            try:
                x, f, info2 = finish_fp64(e64.fit, x, lo, hi, e, gradient)
                par = e64.last_params()
            finally:
                e64.close()
            info = _merge_info(info, info2)
        sec = time.perf_counter() - t0
        npar = eng.np - (nb if fixed_branches else 0)
        return Fit("H0" if fix_omega else "H1", -f, x, par[0], par[1], par[2], par[3], par[4],
                   par[7:11].copy(), x[:nb].copy(), npar, info["iterations"],
                   info["evaluations"], info["converged"], sec, extra, info["stop"],
                   fixed_branches)
    finally:
        if own:
            eng.close()


def fit_model_a_batch(problems, fix_omega=False, omega=1.0, starts=None, precision="fp64",
                      e=1e-8, device=0, kappa=2.0, polish=True, fixed_branches=False):
    """``fit_model_a`` for several data sets at once, on one batch engine
    (adjoint gradient; every data set has its own tree and alignment).
    ``problems`` is a list of (tree, data); ``starts`` a list of starting
    vectors or None entries.  Same optimiser, starting values, fixed branch
    lengths and FP32 polish as the one-data-set fit, so each result equals
    that fit's."""
    n = len(problems)
    starts = [None] * n if starts is None else list(starts)
    x0 = []
    for (t, _), st in zip(problems, starts):
        if st is None:
            v = default_start(t, kappa, omega if not fix_omega else 1.5, fix_omega)
            if fixed_branches:
                v[:t.nbranch] = t.branch_vector()
        else:
            v = np.asarray(st, dtype=np.float64)
        x0.append(v)
    t0 = time.perf_counter()

    def one_pass(prec, xs, finish=False):
        b = engine.BranchSiteBatch(problems, precision=prec, device=device,
                                   fix_omega=fix_omega, omega_fix=omega)
        try:
            bounds = [_bounds(b.default_bounds(i), xs[i], problems[i][0].nbranch,
                              fixed_branches) for i in range(n)]
            los, his = [a for a, _ in bounds], [c for _, c in bounds]
            if finish:
                x, f, info = finish_fp64_batch(b, xs, los, his, e)
            else:
                x, f, info = b.fit(xs, e=e, lo=los, hi=his)
            b.lnl(x)   # every gene's last evaluation at its estimates
            return x, f, info, [b.last_params(i) for i in range(n)], b.np
        finally:
            b.close()

    x, f, info, par, nps = one_pass(precision, x0)
    extra = [{} for _ in range(n)]
    if precision == "fp32" and polish:
        extra = [{"fp32_lnL": -f[i], "fp32_iterations": info[i]["iterations"],
                  "fp32_evaluations": info[i]["evaluations"], "fp32_stop": info[i]["stop"]}
                 for i in range(n)]
        x2, f, info2, par, nps = one_pass("fp64", x, finish=True)
        info = [_merge_info(a, c) for a, c in zip(info, info2)]
        x = x2
    sec = time.perf_counter() - t0
    out = []
    for i, (tree, _) in enumerate(problems):
        nb = tree.nbranch
        p = par[i]
        out.append(Fit("H0" if fix_omega else "H1", -f[i], x[i], p[0], p[1], p[2], p[3], p[4],
                       p[7:11].copy(), x[i][:nb].copy(), nps[i] - (nb if fixed_branches else 0),
                       info[i]["iterations"], info[i]["evaluations"], info[i]["converged"], sec,
                       extra[i], info[i]["stop"], fixed_branches))
    return out


def standard_errors(eng, x, small_diff=1e-6, free=None):
    """SEs of the parameters at x in codeml's printed form (p0, p1
    untransformed) from the Hessian approximation of Seo et al. 2004, as the
    reference computes it with getSE = 1 (HessianSKT2004,
    treesub.c:6979-7048): per-site log f at x +- 2 Small_Diff (|x_i| + 1) for
    every parameter (2 np GPU evaluations), H = -sum_h fpatt_h d_i d_j,
    SE = sqrt(diag((-H)^-1)), -1 where not positive (codeml.c:779-784).
    ``free`` (indices into x) limits the parameters to those estimated (the
    non-branch ones under fix_blength = 2, where codeml's np excludes the
    branch lengths).  The engine is left set to untransformed p's."""
    x = np.asarray(x, dtype=np.float64)
    eng.set_model(eng.fix_omega, eng.omega_fix, p_transformed=False)
    free = list(range(len(x))) if free is None else list(free)
    D = np.zeros((len(free), eng.data.npatt))
    for r, i in enumerate(free):
        eh = 2 * small_diff * (abs(x[i]) + 1)
        xs = x.copy()
        xs[i] = x[i] + eh
        eng.lnl(xs)
        up = eng.fhK()[1]
        xs[i] = x[i] - eh
        eng.lnl(xs)
        D[r] = (up - eng.fhK()[1]) / (2 * eh)
    H = -(D * eng.data.fpatt[None, :]) @ D.T
    d = np.diag(np.linalg.inv(-H))
    return np.where(d > 0, np.sqrt(np.abs(d)), -1.0)


def lrt(h0, h1):
    """2 dlnL of H1 over H0 and its p-values: chi2 with 1 df, and the 50:50
    mixture of a point mass at 0 and chi2_1 (the boundary null of w2 = 1)."""
    stat = max(0.0, 2.0 * (h1.lnL - h0.lnL))
    p1 = math.erfc(math.sqrt(stat / 2.0))
    return {"stat": stat, "p_chi2_1": p1, "p_mixture": 0.5 * p1}


def format_out(tree, data, fit, names):
    """Text in the layout of codeml's out file for the fitted model."""
    nb = tree.nbranch
    t = tree.branch.copy()
    for v in range(tree.nnode):
        if v != tree.root:
            t[v] = fit.branch[tree.ibranch[v]]
    tr = formats.Tree(tree.ns, tree.nnode, tree.root, tree.father, tree.sons, t, tree.label,
                      tree.ibranch, names)
    lines = [f"Model A: branch-site (model = 2, NSsites = 2), {fit.hypothesis}",
             f"ns = {data.ns}  ls = {data.ls}  patterns = {data.npatt}", ""]
    ntime = 0 if fit.fixed_branches else nb
    lines.append(f"lnL(ntime: {ntime}  np: {fit.np}):  {fit.lnL:12.6f}      +0.000000")
    lines.append("  ".join(f"{v:8.6f}" for v in fit.x_untransformed()))
    lines.append("")
    lines.append(tr.newick(lengths=True, labels=True, names=names, digits=6))
    lines.append("")
    lines.append(f"kappa (ts/tv) = {fit.kappa:8.5f}")
    lines.append("")
    lines.append("MLEs of dN/dS (w) for site classes (K=4)")
    lines.append("")
    lines.append("site class             0        1       2a       2b")
    pr = fit.proportions
    lines.append("proportion       " + " ".join(f"{v:8.5f}" for v in pr))
    lines.append("background w     " + " ".join(f"{v:8.5f}" for v in (fit.w0, 1.0, fit.w0, 1.0)))
    lines.append("foreground w     " + " ".join(f"{v:8.5f}" for v in (fit.w0, 1.0, fit.w2, fit.w2)))
    lines.append("")
    lines.append(f"iterations {fit.iterations}  evaluations {fit.evaluations}  "
                 f"converged {int(fit.converged)} ({fit.stop or 'fixed values'})  "
                 f"time {fit.seconds:.3f} s")
    return "\n".join(lines) + "\n"


def format_site_table(title, post, data, names):
    """Per-site class posteriors in the layout of the reference's rst file
    (PrintProbNSsites, codeml.c:4975-5016)."""
    lines = [title, f"(amino acids refer to 1st sequence: {names[0]})", ""]
    for i in range(data.ls):
        h = data.pose[i]
        p = post[:, h]
        best = 0
        for k in range(1, p.shape[0]):
            if p[k] > p[best]:
                best = k
        lines.append(f"{i + 1:4d} {formats.site_amino_acid(data, 0, i)}  "
                     + "".join(f" {v:5.5f}" for v in p) + f" ({best + 1:2d})")
    return "\n".join(lines) + "\n"


def positive_sites(post, data, cutoff=0.5):
    """(site, amino acid, P(w2 > 1), stars) for sites with P(class 2a or 2b)
    above the cutoff (codeml.c:6786-6799): * above 0.95, ** above 0.99."""
    out = []
    for i in range(data.ls):
        t = post[2, data.pose[i]] + post[3, data.pose[i]]
        if t > cutoff:
            out.append((i + 1, formats.site_amino_acid(data, 0, i), t,
                        "**" if t > 0.99 else ("*" if t > 0.95 else "")))
    return out


def format_beb_out(beb, data):
    """BEB part of the out file (codeml.c:6782-6822)."""
    lines = ["", "Bayes Empirical Bayes (BEB) analysis (Yang, Wong & Nielsen 2005. "
             "Mol. Biol. Evol. 22:1107-1118)", "Positive sites for foreground lineages Prob(w>1):"]
    for site, aa, t, sig in positive_sites(beb["post"], data):
        lines.append(f"{site:6d} {aa} {t:.3f}{sig}")
    grid = {"w0": (np.arange(10) + 0.5) / 10, "w2": 1 + (np.arange(10) + 0.5)}
    lines += ["", "", "The grid (see ternary graph for p0-p1)", ""]
    for k in ("w0", "w2"):
        lines.append(f"{k}: " + "".join(f" {v:6.3f}" for v in grid[k]))
    lines += ["", "", "Posterior on the grid", ""]
    for k in ("w0", "w2"):
        lines.append(f"{k}: " + "".join(f" {v:6.3f}" for v in beb["post_" + k]))
    lines += ["", "Posterior for p0-p1 (see the ternary graph) (YWN2015, fig. 1)", ""]
    row = []
    for k, v in enumerate(beb["post_p0p1"]):
        row.append(f" {v:5.3f}")
        if int(math.isqrt(k + 1)) ** 2 == k + 1:
            lines.append("".join(row))
            row = []
    lines.append(f"sum of density on p0-p1 = {np.sum(beb['post_p0p1']):10.6f}")
    lines.append(f"log(fX) = {beb['log_fX']:12.6f}")
    return "\n".join(lines) + "\n"


def read_in_codeml(path):
    """(fixed, values) from an in.codeml file (treesub.c:3925-3927)."""
    with open(path) as f:
        vals = [float(v) for v in f.read().split()]
    if vals and vals[0] == -1:
        return True, np.array(vals[1:])
    return False, np.array(vals)


def read_tree_text(path):
    """The first tree of a codeml tree file (which may start with "ns ntree")."""
    with open(path) as f:
        txt = f.read()
    first = txt.lstrip().split("\n", 1)
    if len(first) == 2 and all(tok.isdigit() for tok in first[0].split()):
        txt = first[1]
    return txt


def load_problems(ctl, base):
    """(tree, data, names) for each of the ``ndata`` alignments of the
    sequence file, all on the tree of the tree file (codeml.c:315-343 reads
    the data sets one after another from the open sequence file; Forestry
    reads the same tree file for each)."""
    ndata = max(1, int(_num(ctl, "ndata", 1)))
    sets = formats.read_phylip_sets(os.path.join(base, ctl["seqfile"]), ndata)
    txt = read_tree_text(os.path.join(base, ctl["treefile"]))
    out = []
    if int(_num(ctl, "fix_blength", 0)) == 2 and ":" not in txt:
        raise ValueError("fix_blength = 2: no branch lengths in tree")   # codeml.c:617
    for names, seqs in sets:
        tree = formats.parse_newick(txt, names)
        if len(tree.sons[tree.root]) < 3:
            raise ValueError("the tree is rooted; give it unrooted (root with 3 sons), as codeml "
                             "uses with clock = 0")
        if not np.any(tree.label == 1):
            raise ValueError("no foreground branch (#1) in the tree")
        out.append((tree, formats.load_codon_data(names, seqs), names))
    return out


def _fixed_fit(eng, tree, vals, fix_omega, omega, fixed_branches=False):
    """A Fit holding the likelihood at the fixed values of an in.codeml file
    that starts with -1 (p0, p1 untransformed; without branch lengths under
    fix_blength = 2, which come from the tree); no optimisation."""
    eng.set_model(fix_omega, omega, p_transformed=False)
    nb = tree.nbranch
    if fixed_branches:
        vals = np.concatenate([tree.branch_vector(), vals])
    negl = eng.lnl(vals[:eng.np])
    par = eng.last_params()
    fit = Fit("H0" if fix_omega else "H1", -negl, vals[:eng.np].copy(), par[0], par[1], par[2],
              par[3], par[4], par[7:11].copy(), vals[:nb].copy(),
              eng.np - (nb if fixed_branches else 0), 0, 1, True, 0.0, {}, "",
              fixed_branches)
    fit.x = engine.x_from_params(vals[:nb], par[0], par[1], par[2], par[3],
                                 None if fix_omega else par[4])
    return fit, vals[:eng.np].copy()


def reference_pattern_order(data):
    """Our pattern indices in the order codeml numbers its patterns:
    PatternWeight (treesub.c:1302-1435) keeps them sorted by the text of the
    column, the species' codon triplets one after another (strcmp order)."""
    text = [formats.code_triplet(data, c) for c in range(data.ncodes)]
    keys = ["".join(text[int(c)] for c in data.z[:, h]) for h in range(data.npatt)]
    return sorted(range(data.npatt), key=lambda h: keys[h])


def format_lnf(data, site_lnf, ntree=1, itree=1):
    """Text of codeml's lnf file for one tree: the header (codeml.c:597),
    the tree number (codeml.c:613) and, from the evaluation at the
    estimates (codeml.c:832-834), one line per site pattern in codeml's
    pattern order (print_lnf_site, treesub.c:7332-7337: pattern, count,
    log f, f, ls f, then each species' codon and amino acid as print1site,
    treesub.c:7647-7669)."""
    out = [f"{ntree:6d} {data.ls:6d} {data.npatt:6d}\n", f"\n\n{itree:2d}\n"]
    text = [formats.code_triplet(data, c) for c in range(data.ncodes)]
    for k, h in enumerate(reference_pattern_order(data)):
        lf = float(site_lnf[h])
        f = math.exp(lf)
        row = "".join(f"{text[int(c)]} ({formats.code_amino_acid(data, int(c))}) "
                      for c in data.z[:, h])
        out.append(f"\n{k + 1:6d} {data.fpatt[h]:6.0f} {lf:16.10f} {f:16.12f} "
                   f"{data.ls * f:12.4f}  {row}")
    return "".join(out)


def _report(tree, data, names, fit, fix_omega):
    """(out text, rst text) of one data set."""
    text = format_out(tree, data, fit, names)
    if "se" in fit.extra:
        text += "SEs for parameters:\n" + "".join(f" {v:8.6f}" for v in fit.extra["se"]) + "\n"
    rst = ""
    if not fix_omega:
        text += format_beb_out(fit.extra["beb"], data)
        rst = (format_site_table("Naive Empirical Bayes (NEB) probabilities for 4 classes",
                                 fit.extra["neb"], data, names) + "\n\n"
               + format_site_table("Bayes Empirical Bayes (BEB) probabilities for 4 classes "
                                   "(class)", fit.extra["beb"]["post"], data, names))
    return text, rst


def run(ctl_path, precision="fp64", gradient="adjoint", device=0):
    """Analyse the data sets of a control file and write ``out`` and ``rst``.
    Returns the Fit, or with ``ndata`` > 1 the list of Fits, one per data
    set, whose sections in ``out`` and ``rst`` open with "Data set k" as
    codeml's do (codeml.c:316-320).  The ML fits of several data sets run
    together on one batch engine (adjoint gradient)."""
    ctl = read_ctl(ctl_path)
    check_ctl(ctl)
    base = os.path.dirname(os.path.abspath(ctl_path))
    problems = load_problems(ctl, base)
    fixed_branches = int(_num(ctl, "fix_blength", 0)) == 2
    fix_omega = int(_num(ctl, "fix_omega", 0)) == 1
    omega = _num(ctl, "omega", 1.0)
    kappa = _num(ctl, "kappa", 2.0)
    start, fixed_vals = None, None
    inpath = os.path.join(base, "in.codeml")
    if os.path.exists(inpath):
        fixed, vals = read_in_codeml(inpath)
        if fixed:
            fixed_vals = vals
        else:
            start = vals
    def full_start(tree, m):
        # in.codeml under fix_blength = 2 holds no branch lengths
        if start is None:
            return None
        if fixed_branches:
            return np.concatenate([tree.branch_vector(), start[:m - tree.nbranch]])
        return start[:m]

    fits = [None] * len(problems)
    if len(problems) > 1 and fixed_vals is None and gradient == "adjoint":
        nps = [t.nbranch + (4 if fix_omega else 5) for t, _, _ in problems]
        fits = fit_model_a_batch([(t, d) for t, d, _ in problems], fix_omega, omega,
                                 [full_start(t, m) for (t, _, _), m in zip(problems, nps)],
                                 precision=precision, device=device, kappa=kappa,
                                 fixed_branches=fixed_branches)
    out_parts, rst_parts = [], []
    for k, (tree, data, names) in enumerate(problems):
        eng = engine.BranchSiteEngine(tree, data, precision=precision, device=device,
                                      fix_omega=fix_omega, omega_fix=omega)
        try:
            fit, xe = fits[k], None
            fixed_run = fixed_vals is not None
            if fixed_run:
                fit, xe = _fixed_fit(eng, tree, fixed_vals, fix_omega, omega, fixed_branches)
            elif fit is None:
                fit = fit_model_a(tree, data, fix_omega, omega,
                                  start=full_start(tree, eng.np), gradient=gradient,
                                  kappa=kappa, eng=eng, fixed_branches=fixed_branches)
            if xe is None:
                xe = fit.x
            # standard errors and the lnf values come from FP64 partials (the
            # reported lnL of an FP32 fit is the FP64 one, and the per-site
            # differences of 2e-6 steps would drown in FP32 rounding noise)
            e64 = eng if eng.precision == "fp64" else engine.BranchSiteEngine(
                tree, data, precision="fp64", device=device, fix_omega=fix_omega,
                omega_fix=omega)
            try:
                if int(_num(ctl, "getSE", 0)) >= 1:
                    xu = fit.x.copy()
                    nb = tree.nbranch
                    xu[nb + 1], xu[nb + 2] = fit.p0, fit.p1
                    free = range(nb if fixed_branches else 0, len(xu))
                    fit.extra["se"] = standard_errors(e64, xu, _num(ctl, "Small_Diff", 1e-6),
                                                      free)
                # the final evaluation at the estimates writes lnf
                # (codeml.c:832-834)
                e64.set_model(fix_omega, omega, p_transformed=not fixed_run)
                e64.lnl(xe)
                fit.extra["site_lnf"] = e64.fhK()[1].copy()
            finally:
                if e64 is not eng:
                    e64.close()
            eng.set_model(fix_omega, omega, p_transformed=not fixed_run)
            # site posteriors under H1, as the reference prints after the fit
            # (NEB codeml.c:5079-5129, BEB 6617-6826)
            if not fix_omega:
                fit.extra["neb"] = eng.neb(xe)
                fit.extra["beb"] = eng.beb(xe)
        finally:
            eng.close()
        fits[k] = fit
        text, rst = _report(tree, data, names, fit, fix_omega)
        if len(problems) > 1:
            text = f"\n\nData set {k + 1}\n" + text
            rst = f"\t{k + 1}\n" + rst
        out_parts.append(text)
        rst_parts.append(rst)
    if not fix_omega:
        with open(os.path.join(base, "rst"), "w") as f:
            f.write("".join(rst_parts))
    # codeml opens lnf afresh for each data set (codeml.c:596), so the file
    # holds the last one's
    tree, data, _ = problems[-1]
    with open(os.path.join(base, "lnf"), "w") as f:
        f.write(format_lnf(data, fits[-1].extra["site_lnf"]))
    with open(os.path.join(base, ctl.get("outfile", "out")), "w") as f:
        f.write("".join(out_parts))
    return fits[0] if len(fits) == 1 else fits


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("ctl", nargs="?", default="codeml.ctl")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--fd", action="store_true", help="forward-difference gradient (gradientB)")
    ap.add_argument("--device", type=int, default=0)
    a = ap.parse_args(argv)
    fits = run(a.ctl, a.precision, "fd" if a.fd else "adjoint", a.device)
    for k, fit in enumerate(fits if isinstance(fits, list) else [fits]):
        head = f"data set {k + 1}: " if isinstance(fits, list) else ""
        print(f"{head}lnL = {fit.lnL:.6f}  np = {fit.np}  iterations = {fit.iterations}  "
              f"time = {fit.seconds:.2f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
