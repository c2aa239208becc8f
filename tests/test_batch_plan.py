"""Host-only checks of the batch engine (no GPU): the plan of a forest of
genes of different sizes, with and without ambiguity codes, renumbers every
gene's jobs into shared buffers (every index checked against the buffer it
indexes, bsa_debug_forest_plan), its sizes add up from the single-gene
plans, and bad batches are rejected before any device call."""
import ctypes as C

import numpy as np
import pytest

from cucodeml_amd import engine, formats, scan, simulate


def make_gene(rng, ns, ncodon, gap_frac=0.0):
    tree = simulate.yule_tree(ns, 0.1, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, 2.0, 0.1, 3.0, 0.5, 0.3, pi, rng,
                                        gap_frac=gap_frac)
    return tree, formats.load_codon_data(tree.names, seqs)


@pytest.fixture(scope="module")
def genes():
    rng = np.random.default_rng(21)
    spec = [(5, 30, 0.0), (12, 200, 0.1), (3, 20, 0.0), (9, 150, 0.0), (16, 90, 0.05),
            (7, 300, 0.0)]
    return [make_gene(rng, ns, nc, gf) for ns, nc, gf in spec]


def test_forest_plan_adds_up(genes):
    singles = [engine.forest_plan([g]) for g in genes]
    both = engine.forest_plan(genes)
    assert both["genes"] == len(genes)
    assert both["npad"] == max(s["npad"] for s in singles)
    assert both["levels"] == max(s["levels"] for s in singles)
    for k in ("nslot", "prune_jobs", "adj_jobs", "eig_jobs"):
        assert both[k] == sum(s[k] for s in singles), k
    # any order of the same genes plans the same sizes
    rev = engine.forest_plan(genes[::-1])
    for k in ("npad", "nslot", "levels", "prune_jobs", "adj_jobs", "eig_jobs"):
        assert rev[k] == both[k]


def test_forest_plan_single_gene_matches_engine_layout(genes):
    tree, data = genes[1]
    p = engine.forest_plan([(tree, data)])
    assert p["npad"] == (data.npatt + 127) // 128 * 128
    # one adjoint job per (branch, class), merged pairs off the foreground path
    assert tree.nbranch * 2 <= p["adj_jobs"] <= tree.nbranch * 4
    assert p["eig_jobs"] == p["adj_jobs"]


def _probs(genes, over):
    keep = []
    probs = []
    for i, (t, d) in enumerate(genes):
        pr = engine._problem(t, d, "fp64", "mfma", 0, 0, None, keep)
        for k, v in over.get(i, {}).items():
            setattr(pr, k, v)
        probs.append(pr)
    arr = (C.POINTER(engine._Problem) * len(probs))(*[C.pointer(p) for p in probs])
    return arr, keep


@pytest.mark.parametrize("over, msg", [
    ({1: {"precision": 1}}, "share device, precision"),
    ({2: {"tile": 128}}, "share device, precision"),
    ({0: {"precision": 5}}, "precision"),
    ({3: {"npatt": 0}}, "gene 3: npatt"),
    ({0: {"root": 0}}, "gene 0: root"),
])
def test_batch_create_rejects(genes, over, msg):
    arr, keep = _probs(genes, over)
    h = C.c_void_p()
    rc = engine.lib().bsa_batch_create(arr, len(genes), C.byref(h))
    assert rc == 1
    assert msg in engine.lib().bsa_last_error().decode()
    assert not h.value


def test_batch_null_arguments():
    L = engine.lib()
    assert L.bsa_batch_create(None, 0, None) == 1
    assert L.bsa_batch_destroy(None) == 0
    assert L.bsa_batch_size(None) == -1
    assert L.bsa_batch_num_params(None, 0) == -1
    assert L.bsa_batch_lnl(None, None, None) == 1
    assert L.bsa_debug_forest_plan(None, 0, None) == 1


def test_batch_groups_bound_padding():
    genes = scan.simulate_genes(40, seed=3, median=200, clip=(50, 1200), taxa=(4, 6))
    groups = scan.batch_groups(genes, batch_genes=8, max_waste=0.25)
    assert sorted(i for g in groups for i in g) == list(range(len(genes)))
    for g in groups:
        assert len(g) <= 8
        top = scan._pad(genes[g[0]].data.npatt)
        assert all(scan._pad(genes[i].data.npatt) <= top for i in g)
        assert all(scan._pad(genes[i].data.npatt) >= 0.75 * top for i in g)
