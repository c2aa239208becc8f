"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
import gzip
import json
import os

import numpy as np

from cucodeml_amd import formats

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["toy_amb", "cfg0", "cfg1", "cfg2s", "cfg4s", "ml16"]
# two sister tips on branches of length 0 with different codons at some
# sites: no class can produce those sites (8 taxa: no node scaling in the
# reference; 40 taxa: the reference scales nodes)
IMPOSSIBLE = ["imposs8", "imposs40"]


def load(name):
    with gzip.open(os.path.join(GOLDEN, name + ".json.gz"), "rt") as f:
        fx = json.load(f)
    names, seqs = formats.read_phylip(fx["phylip"])
    data = formats.load_codon_data(names, seqs)
    tree = formats.parse_newick(fx["newick"], names)
    tree.set_branch_vector(np.array(fx["true_branch"]))
    fx["names"], fx["seqs"] = names, seqs
    return fx, tree, data


def site_to_pattern(fx, data):
    """Reference per-site log f mapped onto our patterns (one value per pattern)."""
    seqs = fx["seqs"]
    ls = len(seqs[0]) // 3
    z, _, _ = formats.encode_codons(fx["names"], seqs)
    col = {}
    for h in range(ls):
        col[z[:, h].tobytes()] = fx["fixed_site_lnf"][h]
    return np.array([col[data.z[:, h].tobytes()] for h in range(data.npatt)])


def patterns_of_sites(fx, data, sites):
    """Our pattern indices of the given alignment sites (0-based)."""
    return sorted({int(data.pose[h]) for h in sites})


def params(fx):
    p = fx["fixed_params"]
    return p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"]


# control file of the reference codeml runs (model A; {fix_omega}, {omega})
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


def parse_out(text):
    """(lnL, np, x) from a codeml `out` file (codeml.c:738-751)."""
    import re
    m = re.search(r"lnL\(ntime:\s*(\d+)\s+np:\s*(\d+)\):\s+(-?\d+\.\d+)", text)
    lnl, npar = float(m.group(3)), int(m.group(2))
    for line in text[m.end():].splitlines()[1:6]:
        vals = line.split()
        if len(vals) == npar:
            try:
                return lnl, npar, [float(v) for v in vals]
            except ValueError:
                pass
    return lnl, npar, None
