"""The adjoint gradient (bsa_lnl_grad) against finite differences of the FP64
likelihood, on the golden data sets and on random trees, in both storage
precisions and with the H0 (fixed w2) and untransformed-p parameterisations."""
import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate

from . import fixtures

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def x_fixed(tree, fx):
    k, p0, p1, w0, w2 = fixtures.params(fx)
    return engine.x_from_params(tree.branch_vector(), k, p0, p1, w0, w2)


def central_fd(eng, x, rel=2e-5):
    g = np.zeros(len(x))
    for i in range(len(x)):
        h = rel * (abs(x[i]) + 1)
        xp, xm = x.copy(), x.copy()
        xp[i] += h
        xm[i] -= h
        g[i] = (eng.lnl(xp) - eng.lnl(xm)) / (2 * h)
    return g


@pytest.mark.parametrize("name", ["toy_amb", "cfg0", "cfg4s"])
def test_adjoint_matches_central_differences(name):
    say(f"case: adjoint vs fd {name}")
    fx, tree, data = fixtures.load(name)
    eng = engine.BranchSiteEngine(tree, data)
    x = x_fixed(tree, fx)
    # move off the simulation point so no component is near zero by symmetry
    x[:tree.nbranch] *= 1.1
    f, g = eng.lnl_grad(x)
    assert f == eng.lnl(x)
    gfd = central_fd(eng, x)
    err = np.abs(g - gfd)
    tol = 2e-5 * np.maximum(1.0, np.abs(gfd)) + 1e-9 * abs(f) / 2e-5
    say(f"  max |g - fd| {err.max():.3e}  max |g| {np.abs(g).max():.3e}")
    assert np.all(err <= tol), np.nonzero(err > tol)


def _random(seed, ns, ncodon, fixw=False):
    rng = np.random.default_rng(seed)
    tree = simulate.yule_tree(ns, 0.08, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=0.05)
    return tree, formats.load_codon_data(tree.names, seqs), p


@pytest.mark.parametrize("mode", ["h0", "untransformed", "fp32"])
def test_adjoint_variants(mode):
    say(f"case: adjoint variant {mode}")
    tree, data, p = _random(5, 20, 300)
    t = tree.branch_vector()
    if mode == "h0":
        e = engine.BranchSiteEngine(tree, data, fix_omega=True, omega_fix=1.0)
        x = engine.x_from_params(t, p["kappa"], p["p0"], p["p1"], p["w0"])
    elif mode == "untransformed":
        e = engine.BranchSiteEngine(tree, data, p_transformed=False)
        x = engine.x_from_params(t, p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"],
                                 transformed=False)
    else:
        e = engine.BranchSiteEngine(tree, data, precision="fp32")
        x = engine.x_from_params(t, p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    f, g = e.lnl_grad(x)
    ref = e if mode != "fp32" else engine.BranchSiteEngine(tree, data)
    gfd = central_fd(ref, x)
    err = np.abs(g - gfd)
    say(f"  max |g - fd| {err.max():.3e}")
    # central differences carry ~1e-11 relative lnL rounding / 4e-5 step
    noise = 5e-12 * abs(f) / (2e-5 * (np.abs(x) + 1))
    tol = (2e-5 if mode != "fp32" else 1e-4) * np.maximum(1.0, np.abs(gfd)) + noise
    assert np.all(err <= tol)


def test_adjoint_full_size_directional():
    """128 taxa x 10,000 codons: g . d against a central difference along a
    random direction, FP64 and FP32 storage."""
    say("case: adjoint full size")
    rng = np.random.default_rng(2024)
    tree = simulate.yule_tree(128, 0.05, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, 10000, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    e64 = engine.BranchSiteEngine(tree, data)
    f, g = e64.lnl_grad(x)
    d = rng.standard_normal(len(x)) * (np.abs(x) + 0.1)
    # Rounding of lnL (~5e5, ~1e-15 relative per evaluation) gives a
    # difference-quotient noise of ~|f| 1e-14 / eps, so the step cannot be
    # small; the h^2 truncation term of central differences (~3e-2 at
    # h = 1e-4 along this d) is removed by Richardson extrapolation over
    # h and 2h, which leaves ~1e-4 of noise.
    eps = 1e-4

    def cd(h):
        return (e64.lnl(x + h * d) - e64.lnl(x - h * d)) / (2 * h)

    d1, d2 = cd(eps), cd(2 * eps)
    dd = (4 * d1 - d2) / 3
    say(f"  g.d {g @ d:.8e}  fd(h) {d1:.8e}  fd(2h) {d2:.8e}  Richardson {dd:.8e}")
    assert abs(g @ d - dd) < 1e-5 * abs(dd) + 3e-14 * abs(f) / eps
    e32 = engine.BranchSiteEngine(tree, data, precision="fp32")
    f32, g32 = e32.lnl_grad(x)
    rel = np.abs(g32 - g) / np.maximum(1.0, np.abs(g))
    say(f"  fp32 grad max rel err {rel.max():.3e}, lnL diff {f32 - f:.3e}")
    assert rel.max() < 1e-3


def test_adjoint_full_size_every_component():
    """128 taxa x 10,000 codons (the gradient-check workload): every one of
    the 258 components of the adjoint gradient, FP64 and FP32 storage,
    against Richardson-extrapolated central differences of the FP64
    likelihood (steps h and 2h per component, h = 1e-4 (|x_i| + 1), at most
    1 % of a branch length; the h^2 term cancels, rounding noise ~|f| 1e-14
    / h)."""
    say("case: adjoint full size, every component")
    rng = np.random.default_rng(7)
    tree = simulate.yule_tree(128, 0.05, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, 10000, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng)
    data = formats.load_codon_data(tree.names, seqs)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    nb = tree.nbranch
    e64 = engine.BranchSiteEngine(tree, data)
    try:
        f, g64 = e64.lnl_grad(x)
        fd = np.zeros(len(x))
        hs = np.zeros(len(x))
        for i in range(len(x)):
            # branch steps stay well inside (0, t): 2h is at most 2 % of the
            # branch, so the h^4 term left by the extrapolation is ~1e-8
            h = 1e-4 * (abs(x[i]) + 1)
            if i < nb:
                h = min(h, 0.01 * x[i])
            hs[i] = h

            def cd(hh):
                xp, xm = x.copy(), x.copy()
                xp[i] += hh
                xm[i] -= hh
                return (e64.lnl(xp) - e64.lnl(xm)) / (2 * hh)
            fd[i] = (4 * cd(h) - cd(2 * h)) / 3
    finally:
        e64.close()
    e32 = engine.BranchSiteEngine(tree, data, precision="fp32")
    try:
        f32, g32 = e32.lnl_grad(x)
    finally:
        e32.close()
    noise = 3e-14 * abs(f) / hs
    tol64 = 5e-6 * np.maximum(1.0, np.abs(fd)) + noise
    err64 = np.abs(g64 - fd)
    err32 = np.abs(g32 - g64) / np.maximum(1.0, np.abs(g64))
    worst = int(np.argmax(err64 / tol64))
    say(f"  np {len(x)}: max |g64 - fd| {err64.max():.3e} (component {worst}, tol "
        f"{tol64[worst]:.3e}); max rel |g32 - g64| {err32.max():.3e}; lnL fp32 - fp64 "
        f"{f32 - f:.3e}")
    assert np.all(err64 <= tol64)
    assert err32.max() <= 1e-4
    assert abs(f32 - f) <= 1e-4
