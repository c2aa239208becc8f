"""FP32 matrix-core pruning (kernel "fp32mc", BSA_KERNEL_FP32MC, FP32
partials): the conditional likelihoods themselves are contracted on the FP32
matrix cores, C(h, j) = L(h, j) + sum_k L(h, k) (P - I)_{j,k}, with the two
largest L(h, k) of each pattern row and the identity term added in FP64 and
the rest of fl32(P - I) on 16x16x4 FP32 MFMA; the gradient sweep as
"mfma32".  The matrix core adds each product to its FP32 accumulator in turn
with round-to-nearest (tests/test_gpu_engine.py pins the lane layout), so
terms below half an ulp of the running sum are lost; without the exclusion
the loss is systematic (lnL 1.5e-3 too low at 128 taxa x 10,000 codons).

Checks, against the FP64 engine (whose lnL equals the oracle's to 1e-9):
|dlnL| <= 1e-4 at 128 taxa x 10,000 codons on two seeds and within the FP32
storage budget on every fixture and in a batch engine, the gradient within
1e-5 relative, and lnL of a plain evaluation equal to that of the gradient
evaluation."""
import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate

from . import fixtures

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", fixtures.NAMES)
def test_fp32mc_on_fixtures(name):
    print(f"case: fp32mc {name}", flush=True)
    fx, tree, data = fixtures.load(name)
    k, p0, p1, w0, w2 = fixtures.params(fx)
    x = engine.x_from_params(tree.branch_vector(), k, p0, p1, w0, w2)
    res = {}
    for prec, kern in (("fp64", "mfma"), ("fp32", "fp32mc")):
        e = engine.BranchSiteEngine(tree, data, precision=prec, kernel=kern)
        try:
            f, g = e.lnl_grad(x)
            res[kern] = (f, g, e.lnl(x))
        finally:
            e.close()
    f64, g64, _ = res["mfma"]
    f32, g32, f32p = res["fp32mc"]
    rel = np.abs(g32 - g64).max() / np.abs(g64).max()
    print(f"  dlnL {f32 - f64:+.2e}, grad rel {rel:.1e}", flush=True)
    # the FP32 storage budget of the fixtures (a few 1e-6 up to 2,000 codons)
    assert abs(f32 - f64) <= 2e-5
    assert f32p == f32
    assert rel < 1e-5


@pytest.mark.parametrize("seed", [2024, 7])
def test_fp32mc_full_size(seed):
    print(f"case: fp32mc full size 128 x 10000 seed {seed}", flush=True)
    rng = np.random.default_rng(seed)
    tree = simulate.yule_tree(128, 0.05, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, 10000, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    res = {}
    for prec, kern in (("fp64", "mfma"), ("fp32", "mfma32"), ("fp32", "fp32mc")):
        e = engine.BranchSiteEngine(tree, data, precision=prec, kernel=kern)
        try:
            e.lnl_grad(x)
            e.timing(reset=True)
            for _ in range(3):
                f, g = e.lnl_grad(x)
            t = e.timing()
        finally:
            e.close()
        res[kern] = (f, g)
        print(f"  {prec} {kern}: -lnL {f:.8f}; pruning {t['sum_prune'] / 3:.3f} ms", flush=True)
    f64, g64 = res["mfma"]
    for kern in ("mfma32", "fp32mc"):
        rel = np.abs(res[kern][1] - g64).max() / np.abs(g64).max()
        print(f"  {kern}: dlnL {res[kern][0] - f64:+.2e}, grad rel {rel:.1e}", flush=True)
    assert abs(res["fp32mc"][0] - f64) <= 1e-4
    assert np.abs(res["fp32mc"][1] - g64).max() / np.abs(g64).max() < 1e-5


def test_fp32mc_batch_within_budget():
    """A batch engine (genes padded to one pattern axis, their jobs in shared
    launches) against each gene's FP64 engine, within the FP32 budget."""
    print("case: fp32mc batch of 5 genes vs single FP64 engines", flush=True)
    rng = np.random.default_rng(11)
    genes, xs = [], []
    for ns, nc in ((6, 120), (12, 300), (20, 250), (9, 400), (16, 180)):
        tree = simulate.yule_tree(ns, 0.1, rng)
        _, pi = simulate.dirichlet_f3x4(rng)
        seqs, _ = simulate.simulate_model_a(tree, nc, 2.0, 0.2, 3.0, 0.5, 0.3, pi, rng)
        genes.append((tree, formats.load_codon_data(tree.names, seqs)))
        xs.append(engine.x_from_params(tree.branch_vector(), 2.0, 0.5, 0.3, 0.2, 3.0))
    b = engine.BranchSiteBatch(genes, precision="fp32", kernel="fp32mc")
    try:
        fb, gb = b.lnl_grad(xs)
    finally:
        b.close()
    for i, (tree, data) in enumerate(genes):
        e = engine.BranchSiteEngine(tree, data, precision="fp64")
        try:
            f, g = e.lnl_grad(xs[i])
        finally:
            e.close()
        rel = np.abs(gb[i] - g).max() / np.abs(g).max()
        print(f"  gene {i}: dlnL {fb[i] - f:+.2e} grad rel {rel:.1e}", flush=True)
        assert abs(fb[i] - f) <= 1e-5
        assert rel < 1e-5
