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


def run_codeml(phylip, newick, ns, xin, fix_omega, omega, fixed):
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "seq.phy"), "w").write(phylip)
        open(os.path.join(d, "tree.nwk"), "w").write(f"{ns} 1\n{newick}\n")
        head = "-1 " if fixed else ""
        open(os.path.join(d, "in.codeml"), "w").write(
            head + " ".join(f"{v:.12f}" for v in xin) + "\n")
        open(os.path.join(d, "codeml.ctl"), "w").write(CTL.format(fix_omega=fix_omega, omega=omega))
        r = subprocess.run([CODEML, "codeml.ctl"], cwd=d, capture_output=True, text=True,
                           input="\n")
        out = open(os.path.join(d, "out")).read()
        lnf = open(os.path.join(d, "lnf")).read()
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:])
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


CFG0 = dict(kappa=2.0, w0=0.1, w2=3.0, p0=0.5, p1=0.3)
SETS = {
    "toy_amb": lambda: make("toy_amb", 11, 6, 120, "yule", 0.15, CFG0, amb=0.08, fixed_w2=False),
    "cfg0": lambda: make("cfg0", 100, 8, 300, "balanced", 0.1, CFG0, fixed_w2=False, ml=True),
    "cfg1": lambda: make("cfg1", 101, 32, 2000, "yule", 0.05, CFG0),
    "cfg2s": lambda: make("cfg2s", 102, 128, 400, "yule", 0.05, "uniform"),
    "cfg4s": lambda: make("cfg4s", 104, 64, 500, "yule", 0.02, "uniform", gaps=0.10),
}

if __name__ == "__main__":
    if not os.path.exists(CODEML):
        subprocess.check_call(["make", "-f", os.path.join(ROOT, "oracle", "Makefile"),
                               CODEML])
    for n in (sys.argv[1:] or list(SETS)):
        SETS[n]()
