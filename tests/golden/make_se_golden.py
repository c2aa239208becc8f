"""Write the golden standard-error fixtures of model A.

For golden data sets written by make_golden.py, the reference's CPU codeml
(oracle/_ref/codeml) is run at the set's fixed parameters (in.codeml starting
with -1, treesub.c:3925-3927) with getSE = 1, under H1 and under H0
(fix_omega = 1, omega = 1).  It then computes the Hessian of Seo et al.
(2004) from per-site central differences (HessianSKT2004,
treesub.c:6979-7048) and prints "SEs for parameters" (codeml.c:754-785),
which this script keeps with the lnL of the same run.

Run from the repository root:  python tests/golden/make_se_golden.py [names...]
"""
import gzip
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_golden import CODEML, CTL, OUT  # noqa: E402


def run(name):
    with gzip.open(os.path.join(OUT, name + ".json.gz"), "rt") as f:
        fx = json.load(f)
    p = fx["fixed_params"]
    res = {"name": name}
    for hyp, fo, om, tail in (("H1", 0, 1.5, [p["w2"]]), ("H0", 1, 1.0, [])):
        xin = list(fx["true_branch"]) + [p["kappa"], p["p0"], p["p1"], p["w0"]] + tail
        with tempfile.TemporaryDirectory() as d:
            open(os.path.join(d, "seq.phy"), "w").write(fx["phylip"])
            open(os.path.join(d, "tree.nwk"), "w").write(f"{fx['ns']} 1\n{fx['newick']}\n")
            open(os.path.join(d, "in.codeml"), "w").write(
                "-1 " + " ".join(f"{v:.12f}" for v in xin) + "\n")
            ctl = CTL.format(fix_omega=fo, omega=om).replace("getSE = 0", "getSE = 1")
            open(os.path.join(d, "codeml.ctl"), "w").write(ctl)
            r = subprocess.run([CODEML, "codeml.ctl"], cwd=d, capture_output=True, text=True,
                               input="\n")
            if r.returncode != 0:
                raise RuntimeError(r.stdout[-2000:])
            out = open(os.path.join(d, "out")).read()
        m = re.search(r"SEs for parameters:\s*\n([^\n]+)", out)
        lm = re.search(r"lnL\(ntime:\s*\d+\s+np:\s*\d+\):\s+(-?\d+\.\d+)", out)
        res[hyp] = {"x": xin, "se": [float(v) for v in m.group(1).split()],
                    "lnL": float(lm.group(1))}
    path = os.path.join(OUT, f"se_{name}.json.gz")
    with gzip.open(path, "wt") as f:
        json.dump(res, f)
    print(name, {h: res[h]["se"][-3:] for h in ("H1", "H0")})


if __name__ == "__main__":
    for n in (sys.argv[1:] or ["cfg0", "toy_amb"]):
        run(n)
