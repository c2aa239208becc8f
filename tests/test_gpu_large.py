"""The largest single-GPU workload: 512 taxa x 50,000 codons with 10 % missing
codons (gaps, an ambiguity code for all 61 codons; cleandata = 0), seeded
model-A simulation.  Too large for the oracle, so it is checked through
properties that do not depend on size: the FP64 MFMA and VALU pruning
kernels (independent implementations) agree, FP32 partials stay within the
FP32 error budget of FP64, and the adjoint gradient matches a central
difference of the FP64 likelihood along a random direction."""
import time

import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate

pytestmark = pytest.mark.gpu

NS, NCODON = 512, 50000


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def workload():
    rng = np.random.default_rng(4)
    tree = simulate.yule_tree(NS, 0.02, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, NCODON, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=0.1)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"],
                             p["w2"])
    return tree, data, x


def test_512_taxa_50k_codons(workload):
    tree, data, x = workload
    say(f"case: {NS} taxa x {NCODON} codons, {data.npatt} patterns, np {len(x)}")
    rng = np.random.default_rng(11)
    d = rng.normal(size=len(x))
    # keep the difference away from the t = 0 kink: branches shorter than
    # 1e-3 (a few dozen of 1021 at mean 0.02) are not moved
    nb = tree.nbranch
    d[:nb][x[:nb] < 1e-3] = 0.0
    d /= np.linalg.norm(d)
    h = 1e-4

    e64 = engine.BranchSiteEngine(tree, data, precision="fp64")
    try:
        f64, g64 = e64.lnl_grad(x)
        t0 = time.perf_counter()
        f64b, _ = e64.lnl_grad(x * (1 + 1e-9))
        t64 = time.perf_counter() - t0
        fp = e64.lnl(x + h * d)
        fm = e64.lnl(x - h * d)
        valu_gap = None
    finally:
        e64.close()
    ev = engine.BranchSiteEngine(tree, data, precision="fp64", kernel="valu")
    try:
        fv = ev.lnl(x)
        valu_gap = abs(fv - f64)
    finally:
        ev.close()
    e32 = engine.BranchSiteEngine(tree, data, precision="fp32")
    try:
        f32, g32 = e32.lnl_grad(x)
        t0 = time.perf_counter()
        e32.lnl_grad(x * (1 + 1e-9))
        t32 = time.perf_counter() - t0
    finally:
        e32.close()
    fd = (fp - fm) / (2 * h)
    gd = g64 @ d
    rel_g = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    say(f"  lnL fp64 {-f64:.6f}  fp32 {-f32:.6f}  |d| {abs(f32 - f64):.3e}  "
        f"valu-mfma {valu_gap:.3e}")
    say(f"  directional derivative: adjoint {gd:.8e}  central difference {fd:.8e}  "
        f"rel {abs(gd - fd) / abs(fd):.2e}")
    say(f"  fp32 gradient rel. difference {rel_g:.2e}; lnL+gradient {t64 * 1e3:.1f} ms (fp64), "
        f"{t32 * 1e3:.1f} ms (fp32)")
    assert valu_gap <= 1e-9 * abs(f64)
    # FP32 budget: the same 1e-4 per 1e5 sites as at 128 x 10k, scaled by
    # sqrt(sites x taxa / (1e4 x 128)) for independent rounding errors
    assert abs(f32 - f64) <= 1e-4 * np.sqrt(data.npatt * NS / (1e4 * 128))
    assert rel_g <= 1e-4
    # central difference error ~ h^2 f''' plus rounding |f| eps / h
    assert abs(gd - fd) <= 1e-5 * abs(fd) + 2e-16 * abs(f64) / h * 10
