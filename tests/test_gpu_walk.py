"""Persistent pruning (BSA_PRUNE_WALK): every pruning level of a forest in
one launch, blocks taking (gene, 128-pattern tile) items in turn and walking
each gene's jobs in level order for their tile.  Per pattern the arithmetic
and the order of the children are those of the level launches, so every
result is the same bitwise: lnL, per-class values and gradients of a batch
of genes of different sizes (a multifurcating tree among them: its jobs
stage three or more interior children's P in pairs), in FP64 and FP32;
evaluations of gene subsets (rounds); graph replay; the single-gene engine;
batch fits.  The autotuner takes the pruning mode as one of its candidates."""
import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate

from .test_gpu_polytomy import gene as polytomy_gene

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def genes():
    rng = np.random.default_rng(31)
    out = []
    for ns, nc, gf in ((6, 60, 0.0), (24, 400, 0.1), (3, 20, 0.0), (11, 150, 0.0),
                       (17, 90, 0.05), (40, 700, 0.0)):
        tree = simulate.yule_tree(ns, 0.1, rng)
        _, pi = simulate.dirichlet_f3x4(rng)
        p = simulate.uniform_params(rng)
        seqs, _ = simulate.simulate_model_a(tree, nc, p["kappa"], p["w0"], p["w2"], p["p0"],
                                            p["p1"], pi, rng, gap_frac=gf)
        data = formats.load_codon_data(tree.names, seqs)
        x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"],
                                 p["w2"])
        out.append((tree, data, x))
    tree, data, x, _ = polytomy_gene()
    out.append((tree, data, x))
    return out


def batch_eval(genes, precision, mode, graph=False, subset=None):
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes], precision=precision)
    try:
        b.set_prune(mode)
        b.use_graph(graph)
        xs = [x for _, _, x in genes]
        f, g = b.lnl_grad(xs)
        fh = [b.fhK(i) for i in range(len(genes))]
        fl = b.lnl(xs)
        sub = None
        if subset is not None:
            sub = b.eval_genes(subset, [xs[i] for i in subset])
        return f, g, fh, fl, sub
    finally:
        b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_walk_equals_level_launches(genes, precision):
    say(f"case: walk vs level launches, batch of {len(genes)} genes, {precision}")
    subset = [1, 3, 6]
    lv = batch_eval(genes, precision, "levels", subset=subset)
    wk = batch_eval(genes, precision, "walk", subset=subset)
    assert np.array_equal(lv[0], wk[0])
    for a, b in zip(lv[1], wk[1]):
        assert np.array_equal(a, b)
    for (fa, sa), (fb, sb) in zip(lv[2], wk[2]):
        assert np.array_equal(fa, fb) and np.array_equal(sa, sb)
    assert np.array_equal(lv[3], wk[3])
    # a round of three genes
    assert np.array_equal(lv[4][0], wk[4][0])
    for a, b in zip(lv[4][1], wk[4][1]):
        assert np.array_equal(a, b)
    say(f"  -lnL {wk[0]}")


def test_walk_graph_replay_equals_eager(genes):
    say("case: walk, graph replay vs eager")
    eg = batch_eval(genes, "fp64", "walk")
    gr = batch_eval(genes, "fp64", "walk", graph=True)
    assert np.array_equal(eg[0], gr[0])
    for a, b in zip(eg[1], gr[1]):
        assert np.array_equal(a, b)


def test_walk_single_engine_and_fit(genes):
    say("case: walk, single-gene engine and a fit")
    tree, data, x = genes[5]
    res = {}
    for mode in ("levels", "walk"):
        e = engine.BranchSiteEngine(tree, data, precision="fp64")
        try:
            e.set_prune(mode)
            f, g = e.lnl_grad(x)
            xf, ff, info = e.fit(x, max_iter=15)
            res[mode] = (f, g, xf, ff, info["evaluations"])
        finally:
            e.close()
    a, b = res["levels"], res["walk"]
    assert a[0] == b[0] and np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2]) and a[3] == b[3] and a[4] == b[4]


def test_walk_refused_for_other_kernels(genes):
    tree, data, _ = genes[0]
    e = engine.BranchSiteEngine(tree, data, precision="fp32", kernel="fp32mc")
    try:
        with pytest.raises(engine.BsaError, match="FP64 matrix-core"):
            e.set_prune("walk")
    finally:
        e.close()


def test_tuner_chooses_a_pruning_mode(genes):
    say("case: batch autotuner with the pruning mode")
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes], precision="fp64")
    try:
        shp = b.tune([x for _, _, x in genes], reps=2)
        say(f"  {shp}")
        assert shp["prune"] in ("levels", "walk")
        assert shp["ms"] <= shp["ms_before"]
    finally:
        b.close()
