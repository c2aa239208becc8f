"""Write the golden site-posterior fixtures (NEB and BEB) of model A.

For a golden data set already written by make_golden.py, the reference's CPU
codeml (oracle/_ref/codeml) is run at the set's fixed parameters (in.codeml
starting with -1, treesub.c:3925-3927) under H1.  After the evaluation it
runs the naive and the Bayes empirical Bayes analyses (codeml.c:5079-5219,
6617-6826); this script keeps what it prints:

* ``rst``: per site the NEB and the BEB posteriors of the 4 site classes
  (%5.5f, PrintProbNSsites codeml.c:4975-5016) with the amino acid of the
  first sequence;
* ``out``: the BEB list of positively selected sites (P(w>1) > 0.5, %.3f,
  codeml.c:6786-6799), the posterior on the w0 and w2 grids (%6.3f) and on
  the p0-p1 triangles (%5.3f) (codeml.c:6803-6822);
* stdout: log(fX) of the BEB grid integration (%12.6f, codeml.c:6754).

Run from the repository root:  python tests/golden/make_beb_golden.py [names...]
"""
import gzip
import json
import os
import re
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_golden import CODEML, CTL, OUT  # noqa: E402


def parse_site_table(text, title):
    """Rows (site, aa, [4 probabilities]) of the rst table after `title`."""
    i = text.index(title)
    rows = []
    for line in text[i:].splitlines()[1:]:
        m = re.match(r"\s*(\d+)\s+(\S)\s+((?:\s*[0-9.]+){4})\s+\(\s*\d+\)", line)
        if m:
            rows.append((int(m.group(1)), m.group(2), [float(v) for v in m.group(3).split()]))
        elif rows and not line.strip():
            break
    return rows


def run(name):
    with gzip.open(os.path.join(OUT, name + ".json.gz"), "rt") as f:
        fx = json.load(f)
    p = fx["fixed_params"]
    xin = list(fx["true_branch"]) + [p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"]]
    with tempfile.TemporaryDirectory() as d:
        open(os.path.join(d, "seq.phy"), "w").write(fx["phylip"])
        open(os.path.join(d, "tree.nwk"), "w").write(f"{fx['ns']} 1\n{fx['newick']}\n")
        open(os.path.join(d, "in.codeml"), "w").write("-1 " + " ".join(f"{v:.12f}" for v in xin)
                                                      + "\n")
        open(os.path.join(d, "codeml.ctl"), "w").write(CTL.format(fix_omega=0, omega=1.5))
        r = subprocess.run([CODEML, "codeml.ctl"], cwd=d, capture_output=True, text=True,
                           input="\n")
        if r.returncode != 0:
            raise RuntimeError(r.stdout[-2000:])
        out = open(os.path.join(d, "out")).read()
        rst = open(os.path.join(d, "rst")).read()
    neb = parse_site_table(rst, "Naive Empirical Bayes (NEB) probabilities")
    beb = parse_site_table(rst, "Bayes Empirical Bayes (BEB) probabilities")
    i = out.index("Bayes Empirical Bayes (BEB) analysis")
    sites = []
    for line in out[i:].splitlines()[2:]:
        m = re.match(r"\s*(\d+)\s+(\S)\s+([0-9.]+)(\**)$", line)
        if m:
            sites.append((int(m.group(1)), m.group(2), float(m.group(3)), m.group(4)))
        elif sites or line.strip().startswith("The grid"):
            break
    grid = {}
    for key in ("w0", "w2"):
        m = re.search(r"The grid.*?\n" + r"(?:.*\n)*?" + key + r":\s+([^\n]+)", out[i:])
        grid[key] = [float(v) for v in m.group(1).split()]
        m = re.search(r"Posterior on the grid.*?\n(?:.*\n)*?" + key + r":\s+([^\n]+)", out[i:])
        grid["post_" + key] = [float(v) for v in m.group(1).split()]
    j = out.index("Posterior for p0-p1", i)
    k = out.index("sum of density on p0-p1", j)
    p01 = [float(v) for v in out[j:k].split("\n", 1)[1].split()]
    m = re.search(r"log\(fX\)\s*=\s*(-?[0-9.]+)", r.stdout)
    res = {"name": name, "neb": [row[2] for row in neb], "beb": [row[2] for row in beb],
           "aa": "".join(row[1] for row in beb), "positive_sites": sites, "grid": grid,
           "post_p0p1": p01, "log_fX": float(m.group(1)) if m else None}
    path = os.path.join(OUT, f"beb_{name}.json.gz")
    with gzip.open(path, "wt") as f:
        json.dump(res, f)
    print(name, "sites", len(res["beb"]), "positive", len(sites), "log fX", res["log_fX"])


if __name__ == "__main__":
    for n in (sys.argv[1:] or ["cfg0", "toy_amb", "cfg4s"]):
        run(n)
