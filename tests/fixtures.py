"""Loading of the golden fixtures written by tests/golden/make_golden.py."""
import gzip
import json
import os

import numpy as np

from cucodeml_amd import formats

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["toy_amb", "cfg0", "cfg1", "cfg2s", "cfg4s"]


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


def params(fx):
    p = fx["fixed_params"]
    return p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"]
