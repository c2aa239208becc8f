"""Trees with multifurcations (nodes of 3 to 8 sons; codeml's ReadTreeN
accepts any number): lnL against the oracle, the adjoint gradient against
central differences of the FP64 likelihood, FP32 partials within their error
budget, and the batch engine equal to the single-gene engine.  A node with
more sons than one job handles (8) is refused with an error that names it."""
import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate
from oracle import oracle

pytestmark = pytest.mark.gpu

# root with 4 sons; a node of 5 sons; a node of 8 sons whose branch is
# foreground, and a foreground tip branch
NEWICK = ("(((S1,S2,S3,S4,S5),S6 #1,(S7,S8,S9)),(S10,(S11,S12,S13,S14,S15,S16,S17,S18) #1),"
          "S19,S20);")


def say(msg):
    print(msg, flush=True)


def gene(newick=NEWICK, ncodon=240, seed=12):
    rng = np.random.default_rng(seed)
    ns = newick.count("S")
    names = [f"S{i + 1}" for i in range(ns)]
    tree = formats.parse_newick(newick, names)
    tree.set_branch_vector(rng.uniform(0.02, 0.3, tree.nbranch))
    _, pi = simulate.dirichlet_f3x4(rng)
    p = dict(kappa=2.3, w0=0.15, w2=3.5, p0=0.55, p1=0.3)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=0.05)
    data = formats.load_codon_data(names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"],
                             p["w2"])
    return tree, data, x, p


def test_tree_shape():
    tree, _, _, _ = gene()
    nsons = sorted(len(tree.sons[v]) for v in range(tree.ns, tree.nnode))
    assert nsons == [2, 3, 3, 4, 5, 8]
    assert len(tree.sons[tree.root]) == 4
    big = [v for v in range(tree.ns, tree.nnode) if len(tree.sons[v]) == 8][0]
    assert tree.label[big] == 1 and tree.label[5] == 1   # S6 and the 8-son clade


def test_lnl_matches_oracle():
    say("case: multifurcating tree (sons 3..8) lnL vs oracle")
    tree, data, x, p = gene()
    e = engine.BranchSiteEngine(tree, data)
    try:
        f = e.lnl(x)
    finally:
        e.close()
    want, _, _ = oracle.lnl(tree, data, p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    say(f"  lnL {-f:.10f} oracle {want:.10f}")
    assert abs(-f - want) <= 1e-9 * abs(want)


def test_adjoint_matches_central_differences():
    say("case: multifurcating tree adjoint gradient vs central differences")
    tree, data, x, _ = gene()
    e = engine.BranchSiteEngine(tree, data)
    try:
        x = x.copy()
        x[:tree.nbranch] *= 1.1
        f, g = e.lnl_grad(x)
        gfd = np.zeros(len(x))
        for i in range(len(x)):
            h = 2e-5 * (abs(x[i]) + 1)
            xp, xm = x.copy(), x.copy()
            xp[i] += h
            xm[i] -= h
            gfd[i] = (e.lnl(xp) - e.lnl(xm)) / (2 * h)
    finally:
        e.close()
    err = np.abs(g - gfd)
    tol = 2e-5 * np.maximum(1.0, np.abs(gfd)) + 1e-9 * abs(f) / 2e-5
    say(f"  max |g - fd| {err.max():.3e}  max |g| {np.abs(g).max():.3e}")
    assert np.all(err <= tol), np.nonzero(err > tol)


def test_fp32_within_budget():
    say("case: multifurcating tree fp32 vs fp64")
    tree, data, x, _ = gene(ncodon=600, seed=4)
    res = {}
    for prec in ("fp64", "fp32"):
        e = engine.BranchSiteEngine(tree, data, precision=prec)
        try:
            res[prec] = e.lnl_grad(x)
        finally:
            e.close()
    (f64, g64), (f32, g32) = res["fp64"], res["fp32"]
    rg = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    say(f"  |d lnL| {abs(f32 - f64):.2e}  grad rel {rg:.1e}")
    assert abs(f32 - f64) <= 1e-5
    assert rg <= 1e-4


def test_batch_equals_single_engine():
    say("case: batch of a multifurcating tree and a binary one vs single engines")
    a = gene()
    rng = np.random.default_rng(8)
    t2 = simulate.yule_tree(9, 0.1, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    seqs, _ = simulate.simulate_model_a(t2, 150, 2.0, 0.2, 2.5, 0.5, 0.3, pi, rng)
    d2 = formats.load_codon_data(t2.names, seqs)
    x2 = engine.x_from_params(t2.branch_vector(), 2.0, 0.5, 0.3, 0.2, 2.5)
    genes = [(a[0], a[1], a[2]), (t2, d2, x2)]
    want = []
    for t, d, x in genes:
        e = engine.BranchSiteEngine(t, d)
        try:
            want.append(e.lnl_grad(x))
        finally:
            e.close()
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes])
    try:
        f, gs = b.lnl_grad([x for _, _, x in genes])
    finally:
        b.close()
    for i, (fw, gw) in enumerate(want):
        assert f[i] == fw
        assert np.abs(gs[i] - gw).max() <= 1e-12 * np.abs(gw).max()


def test_more_sons_than_a_job_holds_is_refused():
    say("case: node with 9 sons refused")
    nw = "((S1,S2,S3,S4,S5,S6,S7,S8,S9),S10 #1,S11);"
    tree, data, _, _ = gene(nw, ncodon=30)
    with pytest.raises(engine.BsaError, match="sons"):
        engine.BranchSiteEngine(tree, data)
