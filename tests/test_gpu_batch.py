"""The batch engine on the GPU: a forest of genes of different sizes (3 to
24 taxa, 20 to 400 codons, with and without gaps, each with its own
foreground branch and model options) gives every gene the likelihood,
per-class values and gradient of that gene's own single-gene engine (FP64:
the same kernels on the same data, so lnL bitwise and gradients to the
rounding of the branch-matrix chunk sums); the likelihood of one gene
against the oracle; FP32 within its error budget; graph replay equal to
eager launches; evaluations of gene subsets equal to whole-batch ones;
batch fits (continuous batching) equal to one-gene fits; the batched genome
scan equal to the per-gene scan."""
import numpy as np
import pytest

from cucodeml_amd import codeml, engine, formats, scan, simulate
from oracle import oracle

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def make_gene(rng, ns, ncodon, gap_frac=0.0):
    tree = simulate.yule_tree(ns, 0.1, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=gap_frac)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"],
                             p["w2"])
    return tree, data, x


@pytest.fixture(scope="module")
def genes():
    rng = np.random.default_rng(77)
    spec = [(6, 60, 0.0), (24, 400, 0.1), (3, 20, 0.0), (11, 150, 0.0), (17, 90, 0.05),
            (8, 300, 0.0), (5, 45, 0.2)]
    return [make_gene(rng, ns, nc, gf) for ns, nc, gf in spec]


def singles(genes, precision="fp64", fixed=()):
    out = []
    for i, (tree, data, x) in enumerate(genes):
        e = engine.BranchSiteEngine(tree, data, precision=precision)
        try:
            xi = x
            if i in fixed:
                e.set_model(True, 1.0)
                xi = x[:-1]
            f, g = e.lnl_grad(xi)
            fh, site = e.fhK()
            out.append((f, g, fh, site))
        finally:
            e.close()
    return out


def test_batch_equals_single_engines_fp64(genes):
    say(f"case: batch of {len(genes)} genes fp64 vs single engines (genes 2, 5 under H0)")
    fixed = (2, 5)
    want = singles(genes, fixed=fixed)
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes])
    try:
        for i in fixed:
            b.set_model(i, True, 1.0)
        xs = [x[:-1] if i in fixed else x for i, (_, _, x) in enumerate(genes)]
        f = b.lnl(xs)
        fg, gs = b.lnl_grad(xs)
        for i, (fw, gw, fhw, sw) in enumerate(want):
            fh, site = b.fhK(i)
            rg = np.abs(gs[i] - gw).max() / np.abs(gw).max()
            say(f"  gene {i}: ns {genes[i][0].ns} npatt {genes[i][1].npatt} -lnL {f[i]:.10f} "
                f"single {fw:.10f} grad rel {rg:.1e}")
            assert f[i] == fw and fg[i] == fw
            assert np.array_equal(fh, fhw) and np.array_equal(site, sw)
            assert rg <= 1e-12
        t = b.timing()
        say(f"  launches per evaluation: {t['n_launch'] / t['n_eval']:.0f} pruning")
    finally:
        b.close()


def test_batch_gene_matches_oracle(genes):
    say("case: batch gene 3 vs the oracle")
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes])
    try:
        f = b.lnl([x for _, _, x in genes])
        par = b.last_params(3)
    finally:
        b.close()
    tree, data, x = genes[3]
    tree.set_branch_vector(x[:tree.nbranch])
    want, _, _ = oracle.lnl(tree, data, *par[:5])
    say(f"  lnL {-f[3]:.10f} oracle {want:.10f}")
    assert abs(-f[3] - want) <= 1e-9 * abs(want)


def test_batch_fp32_within_budget(genes):
    say("case: batch fp32 vs single fp64")
    want = singles(genes)
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes], precision="fp32")
    try:
        f, gs = b.lnl_grad([x for _, _, x in genes])
    finally:
        b.close()
    for i, (fw, gw, _, _) in enumerate(want):
        rg = np.linalg.norm(gs[i] - gw) / np.linalg.norm(gw)
        say(f"  gene {i}: |d lnL| {abs(f[i] - fw):.2e} grad rel {rg:.1e}")
        assert abs(f[i] - fw) <= 1e-5 * max(1.0, genes[i][1].npatt / 1e3)
        assert rg <= 1e-4


def test_batch_graph_equals_eager(genes):
    say("case: batch graph replay vs eager")
    xs = [x for _, _, x in genes]
    rng = np.random.default_rng(5)
    pts = [[x * (1 + 0.01 * rng.standard_normal(len(x))) for x in xs] for _ in range(3)]
    res = {}
    for graph in (False, True):
        b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes])
        b.use_graph(graph)
        try:
            res[graph] = [b.lnl_grad(p) for p in pts]
        finally:
            b.close()
    for (fa, ga), (fb, gb) in zip(res[False], res[True]):
        assert np.array_equal(fa, fb)
        for a, c in zip(ga, gb):
            assert np.array_equal(a, c)


def test_gene_subsets_equal_full_evaluation(genes):
    say("case: evaluations of gene subsets vs the whole batch")
    xs = [x for _, _, x in genes]
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes])
    try:
        f, gs = b.lnl_grad(xs)
        fh6, _ = b.fhK(6)
        for sub in ([1], [0, 2, 3], [4, 5]):
            xs2 = [xs[i] * 1.01 for i in sub]
            fs, gss = b.eval_genes(sub, xs2)
            fw, gw = b.lnl_grad([xs[i] * 1.01 if i in sub else xs[i] for i in range(len(xs))])
            for k, i in enumerate(sub):
                assert fs[k] == fw[i]
                assert np.abs(gss[k] - gw[i]).max() <= 1e-12 * np.abs(gw[i]).max()
        b.lnl_grad(xs)
        fs = b.eval_genes([2, 5], [xs[2] * 1.02, xs[5] * 1.02], grad=False)
        fh6b, _ = b.fhK(6)
        assert np.array_equal(fh6, fh6b)   # gene 6 not in the subset: kept
        assert fs[0] != f[2]
    finally:
        b.close()


def test_batch_fits_equal_single_fits(genes):
    say("case: batch fits (rounds of ready genes) vs single-engine fits")
    sub = genes[:5]
    starts = []
    for tree, _, _ in sub:
        s = codeml.default_start(tree, 2.0, 1.5, False)
        s[:tree.nbranch] = 0.1
        starts.append(s)
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in sub])
    try:
        xb, fb, ib = b.fit(starts)
    finally:
        b.close()
    for i, (tree, data, _) in enumerate(sub):
        e = engine.BranchSiteEngine(tree, data)
        try:
            xs, fs, is_ = e.fit(starts[i])
        finally:
            e.close()
        say(f"  gene {i}: batch -lnL {fb[i]:.8f} ({ib[i]['iterations']} it, {ib[i]['stop']}), "
            f"single {fs:.8f} ({is_['iterations']} it, {is_['stop']})")
        assert ib[i]["stop"] == is_["stop"] and ib[i]["stop"] != "limit"
        assert abs(fb[i] - fs) <= 1e-6 * abs(fs)
        assert np.abs(xb[i] - xs).max() <= 1e-3 * max(1.0, np.abs(xs).max())


def test_batched_scan_equals_scan():
    say("case: batched scan vs per-gene scan, 10 genes")
    genes = scan.simulate_genes(10, seed=9, median=120, clip=(50, 300), taxa=(6, 14))
    a, wa = scan.scan_batched(genes, batch_genes=4, concurrent=2)
    c, wc = scan.scan(genes, workers=4)
    say(f"  wall {wa:.2f} s batched vs {wc:.2f} s per gene")
    for r, s in zip(a, c):
        say(f"  {r['gene']}: lnL1 {r['lnL1']:.6f} / {s['lnL1']:.6f}  lnL0 {r['lnL0']:.6f} / "
            f"{s['lnL0']:.6f}  stop {r['stop']} / {s['stop']}")
        # genes of 6-14 taxa and ~100 codons: degenerate maxima (p0 -> 1,
        # w2 free) end by the stall rule; no fit may run into the limit
        assert "limit" not in r["stop"] and "limit" not in s["stop"]
        assert abs(r["lnL1"] - s["lnL1"]) <= 1e-5 and abs(r["lnL0"] - s["lnL0"]) <= 1e-5
        assert abs(r["stat"] - s["stat"]) <= 2e-5


def test_batch_fit_of_more_genes_than_fit_threads():
    say("case: batch fit of 300 small genes (more than the 256 fit threads)")
    rng = np.random.default_rng(123)
    genes = [make_gene(rng, int(rng.integers(4, 7)), int(rng.integers(20, 60))) for _ in range(300)]
    starts = []
    for tree, _, _ in genes:
        s = codeml.default_start(tree, 2.0, 1.5, False)
        s[:tree.nbranch] = 0.1
        starts.append(s)
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in genes])
    try:
        xb, fb, ib = b.fit(starts)
        t = b.timing()
    finally:
        b.close()
    stops = {}
    for i in ib:
        stops[i["stop"]] = stops.get(i["stop"], 0) + 1
    say(f"  rounds {t['n_eval']:.0f}, evaluations {sum(i['evaluations'] for i in ib)}, "
        f"stop reasons {stops}")
    # genes of 4-6 taxa and 20-60 codons: many maxima are degenerate
    # (p0 -> 1, w2 and p1 free) and end by the stall rule
    assert "limit" not in stops
    for k in (0, 150, 299):
        tree, data, _ = genes[k]
        e = engine.BranchSiteEngine(tree, data)
        try:
            xs, fs, _ = e.fit(starts[k])
        finally:
            e.close()
        assert abs(fb[k] - fs) <= 1e-6 * abs(fs)


def test_chained_lrt_fits_equal_separate_batch_fits():
    """bsa_batch_fit_lrt runs each gene's H1 -> H0 (-> H1 again) chain
    without waiting for the other genes; a gene's result does not depend on
    which genes share a round, so its H1 and H0 fits equal those of two
    whole-batch fits from the same starts bitwise (genes that need the H1
    refit are compared against a separate fit from the H0 estimates)."""
    say("case: chained H1/H0 batch fits vs separate batch fits, 12 genes")
    genes = scan.simulate_genes(12, seed=5, median=150, clip=(40, 400), taxa=(5, 16))
    starts = []
    for g in genes:
        s = codeml.default_start(g.tree, 2.0, 1.5, False)
        s[:g.tree.nbranch] = 0.1
        starts.append(s)
    b = engine.BranchSiteBatch([(g.tree, g.data) for g in genes])
    try:
        x1c, x0c, negl, info = b.fit_lrt(starts)
        par_c = [b.last_params(i) for i in range(len(genes))]
        x1, f1, i1 = b.fit(starts)
        b.set_model(-1, True, 1.0)
        x0, f0, i0 = b.fit([x[:-1] for x in x1])
        b.set_model(-1, False)
        refits = [i for i in range(len(genes)) if info[i]["refit"]]
        if refits:
            xb, fb, _ = b.fit([np.append(x0[i], 1.0) if i in refits else x1[i]
                               for i in range(len(genes))])
    finally:
        b.close()
    for i in range(len(genes)):
        say(f"  gene {i}: H1 {-negl[i, 0]:.6f} / {-f1[i]:.6f}  H0 {-negl[i, 1]:.6f} / "
            f"{-f0[i]:.6f}  refit {info[i]['refit']}")
        assert info[i]["h1_stop"] != "limit" and info[i]["h0_stop"] != "limit"
        if not info[i]["refit"]:
            assert info[i]["h1_stop"] == i1[i]["stop"]
        assert info[i]["h0_stop"] == i0[i]["stop"]
        assert np.array_equal(x0c[i], x0[i]) and negl[i, 1] == f0[i]
        assert info[i]["h0_evaluations"] == i0[i]["evaluations"] - 1
        assert info[i]["refit"] == int(-f0[i] > -f1[i] + 1e-6)
        if info[i]["refit"]:
            best = (xb[i], fb[i]) if fb[i] < f1[i] else (x1[i], f1[i])
            assert np.array_equal(x1c[i], best[0]) and negl[i, 0] == best[1]
        else:
            assert np.array_equal(x1c[i], x1[i]) and negl[i, 0] == f1[i]
            assert info[i]["h1_evaluations"] == i1[i]["evaluations"]
        # the engine is left at the H1 estimates
        assert abs(par_c[i][4] - (x1c[i][-1])) <= 1e-12 * max(1.0, abs(x1c[i][-1]))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_batch_fit_asks_sweeps_only_at_accepted_points(genes, precision):
    """A batch fit's line-search trials ask for the forward pass alone; the
    adjoint sweep of that pass follows in a later round where the optimiser
    accepts the point (other genes' rounds in between leave the kept pass
    intact).  The trajectory is that of the same minimiser given value and
    gradient at every trial, here the gene evaluated alone on the same batch
    (eval_genes), bitwise: same estimates, -lnL and iterations, one more
    evaluation (the fit's last, at its estimates).  With FP64 partials a
    stall exit is checked by Newton iterations that the host minimiser does
    not run, so there only fits that end otherwise are compared."""
    say(f"case: batch fit, sweeps only at accepted points, {precision}")
    sub = genes[:5]
    starts = []
    for tree, _, _ in sub:
        s = codeml.default_start(tree, 2.0, 1.5, False)
        s[:tree.nbranch] = 0.1
        starts.append(s)
    b = engine.BranchSiteBatch([(t, d) for t, d, _ in sub], precision=precision)
    try:
        xb, fb, ib = b.fit(starts)
        compared = 0
        for i in range(len(sub)):
            if precision == "fp64" and ib[i]["stop"] in ("stall", "nodescent", "newton"):
                continue

            def fun(x, i=i):
                f, g = b.eval_genes([i], [x])
                return f[0], g[0]
            lo, hi = b.default_bounds(i)
            xs, fs, ins = engine.minimize(fun, starts[i], lo, hi)
            say(f"  gene {i}: batch -lnL {fb[i]:.10f} ({ib[i]['iterations']} it, "
                f"{ib[i]['evaluations']} ev, {ib[i]['stop']}), every trial both "
                f"{fs:.10f} ({ins['iterations']} it, {ins['evaluations']} ev, {ins['stop']})")
            assert np.array_equal(xb[i], xs) and fb[i] == fs
            assert ib[i]["iterations"] == ins["iterations"]
            assert ib[i]["evaluations"] == ins["evaluations"] + 1
            compared += 1
        assert compared >= 3
    finally:
        b.close()
