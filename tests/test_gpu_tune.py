"""Launch shapes and the autotuner on the GPU: every shape gives the same
-lnL bitwise (pruning tiles and child-adjoint tiles per block change no
arithmetic; the branch-matrix chunking only the order of its sums, so
gradients agree to rounding), bad shapes are rejected, and the tuner keeps
a shape no slower than the one it started from."""
import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gene():
    rng = np.random.default_rng(31)
    tree = simulate.yule_tree(40, 0.05, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, 1500, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=0.05)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"],
                             p["w2"])
    return tree, data, x


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_shapes_leave_results_unchanged(gene, precision):
    tree, data, x = gene
    print(f"case: launch shapes, 40 taxa x {data.npatt} patterns, {precision}", flush=True)
    e = engine.BranchSiteEngine(tree, data, precision=precision)
    try:
        f0, g0 = e.lnl_grad(x)
        for tile, tpb, tpc in [(128, 0, 0), (64, 1, 4), (64, 8, 32), (64, 2, 64), (128, 2, 1),
                               (64, 0, 0)]:
            e.set_shapes(tile, tpb, tpc)
            f, g = e.lnl_grad(x)
            rel = np.abs(g - g0).max() / np.abs(g0).max()
            print(f"  tile {tile} tpb {tpb} tpc {tpc}: d lnL {f - f0:.1e} grad rel {rel:.1e}",
                  flush=True)
            assert f == f0
            assert rel <= 1e-12
        for bad in [(96, 0, 0), (64, 9, 0), (64, 0, 65), (64, -1, 0)]:
            with pytest.raises(engine.BsaError):
                e.set_shapes(*bad)
    finally:
        e.close()


def test_tuner_picks_a_shape_no_slower(gene):
    tree, data, x = gene
    print("case: autotuner", flush=True)
    e = engine.BranchSiteEngine(tree, data, precision="fp32")
    try:
        f0, g0 = e.lnl_grad(x)
        t = e.tune(x, reps=3)
        print(f"  {t}", flush=True)
        assert t["tile"] in (64, 128) and 0 <= t["tpb"] <= 8 and 1 <= t["tpc"] <= 32
        assert t["ms"] <= t["ms_before"] and t["ms"] > 0
        f, g = e.lnl_grad(x)
        assert f == f0
        assert np.abs(g - g0).max() <= 1e-12 * np.abs(g0).max()
    finally:
        e.close()


def test_batch_tuner():
    print("case: batch autotuner", flush=True)
    rng = np.random.default_rng(8)
    genes, xs = [], []
    for ns, nc in [(10, 200), (14, 300), (9, 250)]:
        tree = simulate.yule_tree(ns, 0.1, rng)
        _, pi = simulate.dirichlet_f3x4(rng)
        seqs, _ = simulate.simulate_model_a(tree, nc, 2.0, 0.1, 3.0, 0.5, 0.3, pi, rng)
        genes.append((tree, formats.load_codon_data(tree.names, seqs)))
        xs.append(engine.x_from_params(tree.branch_vector(), 2.0, 0.5, 0.3, 0.1, 3.0))
    b = engine.BranchSiteBatch(genes)
    try:
        f0, _ = b.lnl_grad(xs)
        t = b.tune(xs)
        print(f"  {t}", flush=True)
        f1, _ = b.lnl_grad(xs)
        assert np.array_equal(f0, f1)
        assert t["ms"] <= t["ms_before"]
    finally:
        b.close()
