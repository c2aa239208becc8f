"""Write the golden fixtures of the model-A likelihood path.

Each fixture is a seeded synthetic data set (our own simulator) together with
what the reference's CPU codeml (built from the reference sources by
oracle/Makefile into oracle/_ref/codeml) prints for it:

* fixed-parameter runs (``in.codeml`` starting with -1, treesub.c:3925-3927):
  lnL from ``out`` (%13.6f) and every site's log f_h from ``lnf`` (%16.10f,
  treesub.c:7332-7337), mapped from codeml's patterns back to sites;
* maximum-likelihood runs from a seeded start (``in.codeml`` without -1):
  the MLE vector and lnL printed in ``out`` (codeml.c:738-751), for the
  H1 (fix_omega=0) and H0 (fix_omega=1, omega=1) hypotheses.

Run from the repository root:  python tests/golden/make_golden.py [names...]
"""
import gzip
import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from cucodeml_amd import formats, simulate  # noqa: E402

CODEML = os.path.join(ROOT, "oracle", "_ref", "codeml")
OUT = os.path.dirname(os.path.abspath(__file__))

CTL = """seqfile = seq.phy
treefile = tree.nwk
outfile = out
noisy = 0
verbose = 0
runmode = 0
seqtype = 1
CodonFreq = 2
clock = 0
aaDist = 0
model = 2
NSsites = 2
icode = 0
Mgene = 0
fix_kappa = 0
kappa = 2
fix_omega = {fix_omega}
omega = {omega}
fix_alpha = 1
alpha = 0
Malpha = 0
getSE = 0
RateAncestor = 0
Small_Diff = 1e-6
cleandata = 0
fix_blength = 0
method = 0
"""


def run_codeml(phylip, newick, ns, xin, fix_omega, omega, fixed, fix_blength=0, raw=None):
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "seq.phy"), "w").write(phylip)
        open(os.path.join(d, "tree.nwk"), "w").write(f"{ns} 1\n{newick}\n")
        head = "-1 " if fixed else ""
        open(os.path.join(d, "in.codeml"), "w").write(
            head + " ".join(f"{v:.12f}" for v in xin) + "\n")
        ctl = CTL.format(fix_omega=fix_omega, omega=omega)
        ctl = ctl.replace("fix_blength = 0", f"fix_blength = {fix_blength}")
        open(os.path.join(d, "codeml.ctl"), "w").write(ctl)
        r = subprocess.run([CODEML, "codeml.ctl"], cwd=d, capture_output=True, text=True,
                           input="\n")
        out = open(os.path.join(d, "out")).read()
        lnf = open(os.path.join(d, "lnf")).read()
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:])
        if raw is not None:
            raw["out"], raw["lnf"], raw["stdout"] = out, lnf, r.stdout
    m = re.search(r"lnL\(ntime:\s*(\d+)\s+np:\s*(\d+)\):\s+(-?\d+\.\d+)", out)
    lnL = float(m.group(3))
    np_ = int(m.group(2))
    after = out[m.end():].splitlines()
    xs = None
    for line in after[1:6]:
        vals = line.split()
        if len(vals) == np_:
            try:
                xs = [float(v) for v in vals]
                break
            except ValueError:
                pass
    pat = {}
    for line in lnf.splitlines():
        p = line.split()
        if len(p) >= 7 and p[0].isdigit() and "(" in line:
            pat[tuple(p[5::2])] = float(p[2])
    return lnL, xs, pat


def site_values(seqs, pat):
    ls = len(seqs[0]) // 3
    out = []
    for h in range(ls):
        key = tuple(s[3 * h:3 * h + 3] for s in seqs)
        out.append(pat[key])
    return out


def inject_ambiguities(seqs, rng, frac):
    codes = ["AAR", "GAY", "NNN", "---", "TTN", "CT?", "ANN", "GGN", "YTG"]
    out = []
    for s in seqs:
        s = list(s)
        for h in range(len(s) // 3):
            if rng.random() < frac:
                s[3 * h:3 * h + 3] = list(codes[int(rng.integers(len(codes)))])
        out.append("".join(s))
    return out


def make(name, seed, ns, ncodon, shape, mean_bl, params, gaps=0.0, amb=0.0, fixed_w2=True,
         ml=False):
    rng = np.random.default_rng(seed)
    if shape == "balanced":
        tree = simulate.balanced_tree(ns, mean_bl, rng)
    else:
        tree = simulate.yule_tree(ns, mean_bl, rng)
    if params == "uniform":
        params = simulate.uniform_params(rng)
    f3, pi = simulate.dirichlet_f3x4(rng)
    while amb and f3.min() < 0.05:   # injected ambiguities need every base observed
        f3, pi = simulate.dirichlet_f3x4(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, params["kappa"], params["w0"],
                                        params["w2"], params["p0"], params["p1"], pi, rng,
                                        gap_frac=gaps)
    if amb:
        seqs = inject_ambiguities(seqs, rng, amb)
    phylip = f"{ns} {3 * ncodon}\n" + "".join(f"{n}  {s}\n" for n, s in zip(tree.names, seqs))
    newick = tree.newick(digits=10)
    t = list(tree.branch_vector())
    p = params
    fx = {"fixed_params": p, "true_branch": t, "newick": newick, "phylip": phylip,
          "seed": seed, "ns": ns, "ncodon": ncodon}
    if fixed_w2:
        xin = t + [p["kappa"], p["p0"], p["p1"], p["w0"]]
        lnL, _, pat = run_codeml(phylip, newick, ns, xin, 1, p["w2"], True)
    else:
        xin = t + [p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"]]
        lnL, _, pat = run_codeml(phylip, newick, ns, xin, 0, 1.5, True)
    fx["fixed_lnL"] = lnL
    fx["fixed_site_lnf"] = site_values(seqs, pat)
    if ml:
        # seeded start in codeml's optimisation variables (logistic p's)
        x0p = [0.4, 1.5, np.log(0.6 / 0.2), np.log(0.2 / 0.2), 0.3]
        t0 = [0.1] * len(t)
        res = {}
        for hyp, fo, om, extra in (("H1", 0, 1.5, [2.0]), ("H0", 1, 1.0, [])):
            x0 = t0 + x0p[1:4] + [x0p[4]] + extra
            lnL, xs, _ = run_codeml(phylip, newick, ns, x0, fo, om, False)
            res[hyp] = {"start": x0, "mle_x": xs, "mle_lnL": lnL}
        fx["ml"] = res
    path = os.path.join(OUT, name + ".json.gz")
    with gzip.open(path, "wt") as f:
        json.dump(fx, f)
    print(name, "lnL", fx["fixed_lnL"], "sites", len(fx["fixed_site_lnf"]),
          ("ML " + str({k: v["mle_lnL"] for k, v in fx.get("ml", {}).items()})) if ml else "")


def _logit_start(kappa=1.5, p0=0.6, p1=0.2, w0=0.3):
    return [kappa, float(np.log(p0 / (1 - p0 - p1))), float(np.log(p1 / (1 - p0 - p1))), w0]


def make_ml(name, seed, ns, ncodon, mean_bl, params):
    """An ML fixture whose H1 maximum is interior (p2 > 0, w2 > 1): the
    foreground label sits on the longest interior branch, so that the
    selected class is identifiable.  H1 and H0 are fitted by the reference
    from seeded starts (branch lengths 0.1), then refitted from their own
    printed estimates (a second ming2 run from the first one's x), which
    takes them closer to the maximum than one run's convergence rule does.
    H1 and H0 are also fitted with fix_blength = 2 (branch lengths held at
    the tree's values, treesub.c:3729-3730, codeml.c:617-627), and the
    fixed-parameter run keeps the reference's lnf text (treesub.c:7332-7337)."""
    rng = np.random.default_rng(seed)
    tree = simulate.yule_tree(ns, mean_bl, rng)
    inner = [v for v in range(ns, tree.nnode) if v != tree.root]
    fg = max(inner, key=lambda v: tree.branch[v])
    tree.label[:] = 0
    tree.label[tree.root] = -1
    tree.label[fg] = 1
    f3, pi = simulate.dirichlet_f3x4(rng)
    p = params
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng)
    phylip = f"{ns} {3 * ncodon}\n" + "".join(f"{n}  {s}\n" for n, s in zip(tree.names, seqs))
    newick = tree.newick(digits=10)
    t = list(tree.branch_vector())
    fx = {"fixed_params": p, "true_branch": t, "newick": newick, "phylip": phylip,
          "seed": seed, "ns": ns, "ncodon": ncodon}
    raw = {}
    xin = t + [p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"]]
    lnL, _, pat = run_codeml(phylip, newick, ns, xin, 0, 1.5, True, raw=raw)
    fx["fixed_lnL"] = lnL
    fx["fixed_site_lnf"] = site_values(seqs, pat)
    fx["fixed_lnf_text"] = raw["lnf"]
    res = {}
    for hyp, fo, om, extra in (("H1", 0, 1.5, [2.0]), ("H0", 1, 1.0, [])):
        x0 = [0.1] * len(t) + _logit_start() + extra
        lnL, xs, _ = run_codeml(phylip, newick, ns, x0, fo, om, False)
        # the printed x has p0, p1 untransformed; restart from it in the
        # optimisation variables
        nb = len(t)
        xr = list(xs)
        xr[nb + 1:nb + 3] = _logit_start(1, xs[nb + 1], xs[nb + 2], 1)[1:3]
        lnL2, xs2, _ = run_codeml(phylip, newick, ns, xr, fo, om, False)
        res[hyp] = {"start": x0, "mle_x": xs, "mle_lnL": lnL, "polished_start": xr,
                    "polished_x": xs2, "polished_lnL": lnL2}
        print(name, hyp, lnL, lnL2, flush=True)
    fx["ml"] = res
    fb = {}
    for hyp, fo, om, extra in (("H1", 0, 1.5, [2.0]), ("H0", 1, 1.0, [])):
        x0 = _logit_start() + extra
        lnL, xs, _ = run_codeml(phylip, newick, ns, x0, fo, om, False, fix_blength=2)
        fb[hyp] = {"start": x0, "mle_x": xs, "mle_lnL": lnL}
        print(name, "fix_blength = 2", hyp, lnL, flush=True)
    fx["ml_fix_blength2"] = fb
    path = os.path.join(OUT, name + ".json.gz")
    with gzip.open(path, "wt") as f:
        json.dump(fx, f)


def make_impossible(name, seed, ns, ncodon, mean_bl, nbad):
    """Sites no class can produce: two sister tips on branches of length 0
    (P(0) = I, in.codeml -1 values) with different sense codons at nbad
    sites, so every class's f is 0 there and the reference takes its 1e-300
    floor (fx_r treesub.c:7524-7534, lfundG 7376-7385), plus, when nodes are
    scaled (more than 15 tips below a node, SetNodeScale treesub.c:6917-6937),
    the scale logs of NodeScale (treesub.c:6940-6968), -800 where a scaled
    node's partials are all below 1e-300."""
    rng = np.random.default_rng(seed)
    tree = simulate.yule_tree(ns, mean_bl, rng)
    p = dict(kappa=2.0, w0=0.1, w2=3.0, p0=0.5, p1=0.3)
    f3, pi = simulate.dirichlet_f3x4(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng)
    cherry = next(v for v in range(ns, tree.nnode)
                  if len(tree.sons[v]) >= 2 and sum(s < ns for s in tree.sons[v]) >= 2)
    a, b = [s for s in tree.sons[cherry] if s < ns][:2]
    seqs = [list(s) for s in seqs]
    bad = sorted(rng.choice(ncodon, size=nbad, replace=False).tolist())
    for h in bad:
        seqs[a][3 * h:3 * h + 3] = list("TTT")
        seqs[b][3 * h:3 * h + 3] = list("TTC")
    seqs = ["".join(s) for s in seqs]
    # the tips were simulated on their own branches, so they differ at other
    # sites too; with both branches at length 0 every such site is impossible
    bad = [h for h in range(ncodon) if seqs[a][3 * h:3 * h + 3] != seqs[b][3 * h:3 * h + 3]]
    t = tree.branch.copy()
    t[a] = t[b] = 0.0
    tree.branch[:] = t
    tv = list(tree.branch_vector())
    phylip = f"{ns} {3 * ncodon}\n" + "".join(f"{n}  {s}\n" for n, s in zip(tree.names, seqs))
    newick = tree.newick(digits=10)
    xin = tv + [p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"]]
    raw = {}
    lnL, _, pat = run_codeml(phylip, newick, ns, xin, 0, 1.5, True, raw=raw)
    fx = {"fixed_params": p, "true_branch": tv, "newick": newick, "phylip": phylip,
          "seed": seed, "ns": ns, "ncodon": ncodon, "impossible_sites": bad,
          "fixed_lnL": lnL, "fixed_site_lnf": site_values(seqs, pat)}
    path = os.path.join(OUT, name + ".json.gz")
    with gzip.open(path, "wt") as f:
        json.dump(fx, f)
    print(name, "lnL", lnL, "impossible sites", [fx["fixed_site_lnf"][h] for h in bad][:4])


CFG0 = dict(kappa=2.0, w0=0.1, w2=3.0, p0=0.5, p1=0.3)
SETS = {
    "toy_amb": lambda: make("toy_amb", 11, 6, 120, "yule", 0.15, CFG0, amb=0.08, fixed_w2=False),
    "cfg0": lambda: make("cfg0", 100, 8, 300, "balanced", 0.1, CFG0, fixed_w2=False, ml=True),
    "cfg1": lambda: make("cfg1", 101, 32, 2000, "yule", 0.05, CFG0),
    "cfg2s": lambda: make("cfg2s", 102, 128, 400, "yule", 0.05, "uniform"),
    "cfg4s": lambda: make("cfg4s", 104, 64, 500, "yule", 0.02, "uniform", gaps=0.10),
    "ml16": lambda: make_ml("ml16", 116, 16, 600, 0.1,
                            dict(kappa=2.5, w0=0.1, w2=4.0, p0=0.5, p1=0.25)),
    "imposs8": lambda: make_impossible("imposs8", 208, 8, 150, 0.1, 5),
    "imposs40": lambda: make_impossible("imposs40", 240, 40, 150, 0.05, 5),
}

if __name__ == "__main__":
    if not os.path.exists(CODEML):
        subprocess.check_call(["make", "-f", os.path.join(ROOT, "oracle", "Makefile"),
                               CODEML])
    for n in (sys.argv[1:] or list(SETS)):
        SETS[n]()
