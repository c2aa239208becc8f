"""torch.autograd.gradcheck of the autograd binding (cucodeml_amd.autograd):
the adjoint gradient of lnL in codeml's variables against gradcheck's own
central differences of the engine's FP64 likelihood, under H1, H0 and with
untransformed p's."""
import numpy as np
import pytest
import torch

from cucodeml_amd import autograd, engine, formats, simulate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gene():
    rng = np.random.default_rng(17)
    tree = simulate.yule_tree(8, 0.1, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    seqs, _ = simulate.simulate_model_a(tree, 80, 2.5, 0.2, 3.0, 0.5, 0.3, pi, rng,
                                        gap_frac=0.05)
    return tree, formats.load_codon_data(tree.names, seqs)


@pytest.mark.parametrize("mode", ["h1", "h0", "untransformed"])
def test_gradcheck(gene, mode):
    tree, data = gene
    print(f"case: gradcheck {mode}", flush=True)
    eng = engine.BranchSiteEngine(tree, data)
    try:
        if mode == "h1":
            x = engine.x_from_params(tree.branch_vector(), 2.5, 0.5, 0.3, 0.2, 3.0)
        elif mode == "h0":
            eng.set_model(True, 1.0)
            x = engine.x_from_params(tree.branch_vector(), 2.5, 0.5, 0.3, 0.2)
        else:
            eng.set_model(False, 1.0, p_transformed=False)
            x = engine.x_from_params(tree.branch_vector(), 2.5, 0.5, 0.3, 0.2, 3.0,
                                     transformed=False)
        xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
        y = autograd.lnl(xt, eng)
        y.backward()
        want = -eng.lnl_grad(x)[1]
        assert torch.allclose(xt.grad, torch.from_numpy(want), rtol=0, atol=0)
        # lnL ~ -1e3: central differences with eps 1e-6 carry ~1e-13 * 1e3 /
        # 1e-6 = 1e-4 of rounding noise
        assert torch.autograd.gradcheck(lambda v: autograd.lnl(v, eng), (xt,), eps=1e-6,
                                        atol=5e-4, rtol=1e-5)
    finally:
        eng.close()
