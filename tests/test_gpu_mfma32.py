"""FP32 matrix-core products in the gradient sweep (kernel "mfma32",
BSA_KERNEL_MFMA32, FP32 partials): the child adjoints Abar_v = X P_v run on
the FP32 matrix cores with P - I in FP32 and the identity term added in
FP64, and the interior children's branch matrices M_v = sum_h X_h^T L_h in
FP32 per 16-pattern group, summed in FP64; the pruning keeps FP64 products.
Checks: lnL
and the per-class values bitwise those of the FP32 engine with FP64 products
(same forward pass), the gradient within 1e-5 relative of FP64 on the
fixtures and at 128 taxa x 10,000 codons, and the child adjoints no slower."""
import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate

from . import fixtures

pytestmark = pytest.mark.gpu


def x_fixed(tree, fx):
    k, p0, p1, w0, w2 = fixtures.params(fx)
    return engine.x_from_params(tree.branch_vector(), k, p0, p1, w0, w2)


@pytest.mark.parametrize("name", fixtures.NAMES)
def test_mfma32_gradient_on_fixtures(name):
    print(f"case: mfma32 {name}", flush=True)
    fx, tree, data = fixtures.load(name)
    x = x_fixed(tree, fx)
    res = {}
    for prec, kern in (("fp64", "mfma"), ("fp32", "mfma"), ("fp32", "mfma32")):
        e = engine.BranchSiteEngine(tree, data, precision=prec, kernel=kern)
        try:
            f, g = e.lnl_grad(x)
            fh, _ = e.fhK()
        finally:
            e.close()
        res[prec + kern] = (f, g, fh)
    f64, g64, _ = res["fp64mfma"]
    fs, gs, fhs = res["fp32mfma"]
    f32, g32, fh32 = res["fp32mfma32"]
    rel32 = np.abs(g32 - g64).max() / np.abs(g64).max()
    rels = np.abs(gs - g64).max() / np.abs(g64).max()
    print(f"  grad rel to fp64: mfma32 {rel32:.1e}, fp32 FP64 products {rels:.1e}", flush=True)
    assert f32 == fs and np.array_equal(fh32, fhs)
    assert rel32 < 1e-5


def test_mfma32_full_size():
    print("case: mfma32 full size 128 x 10000", flush=True)
    rng = np.random.default_rng(2024)
    tree = simulate.yule_tree(128, 0.05, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, 10000, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    res = {}
    for prec, kern in (("fp64", "mfma"), ("fp32", "mfma"), ("fp32", "mfma32")):
        e = engine.BranchSiteEngine(tree, data, precision=prec, kernel=kern)
        try:
            e.lnl_grad(x)
            e.timing(reset=True)
            for _ in range(3):
                f, g = e.lnl_grad(x)
            t = e.timing()
        finally:
            e.close()
        res[prec + kern] = (f, g, t["sum_adjoint"] / 3)
        print(f"  {prec} {kern}: -lnL {f:.6f}; child adjoints {t['sum_adjoint'] / 3:.3f} ms",
              flush=True)
    f64, g64, _ = res["fp64mfma"]
    fs, gs, ts = res["fp32mfma"]
    f32, g32, t32 = res["fp32mfma32"]
    rel = np.abs(g32 - g64).max() / np.abs(g64).max()
    print(f"  mfma32 grad rel to fp64 {rel:.2e}", flush=True)
    assert f32 == fs
    assert rel < 1e-5
    assert t32 < ts
