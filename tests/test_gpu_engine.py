"""GPU parity of the HIP engine (through the C ABI) against the reference's
own CPU codeml output (golden fixtures) and the CPU oracle, on the same
seeded inputs.  FP64 bar: lnL within 1e-9 relative of the reference and of
the oracle; per-pattern log f within 1e-10 relative (+1e-10 print floor)."""
import numpy as np
import pytest

from cucodeml_amd import engine, formats, simulate
from oracle import oracle

from . import fixtures

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def make_engine(tree, data, fx=None, **kw):
    return engine.BranchSiteEngine(tree, data, **kw)


def x_fixed(tree, fx):
    k, p0, p1, w0, w2 = fixtures.params(fx)
    return engine.x_from_params(tree.branch_vector(), k, p0, p1, w0, w2)


def test_mfma_lane_layout():
    """One 16x16x4 MFMA in FP64 and FP32 with A[l%16][l//16], B[l//16][l%16]
    operands: element i of lane l holds D[l//16 + 4i][l%16] in FP64 and
    D[4(l//16) + i][l%16] in FP32, the layouts the pruning kernel assumes."""
    say("case: mfma probe")
    rng = np.random.default_rng(0)
    A = rng.integers(-4, 5, size=(16, 4)).astype(np.float64)
    B = rng.integers(-4, 5, size=(4, 16)).astype(np.float64)
    D = A @ B
    d64, d32 = engine.mfma_probe(A, B)
    e64 = np.array([[D[l // 16 + 4 * i, l % 16] for i in range(4)] for l in range(64)])
    e32 = np.array([[D[4 * (l // 16) + i, l % 16] for i in range(4)] for l in range(64)])
    if not np.array_equal(d64, e64):
        say("raw f64 lanes:", d64.tolist())
    if not np.array_equal(d32, e32):
        say("raw f32 lanes:", d32.tolist())
    assert np.array_equal(d64, e64)
    assert np.array_equal(d32, e32)


def test_mfma32_accumulates_in_order_with_rounding():
    """The FP32 16x16x4 MFMA adds its four products to the accumulator one
    after another (k = 0..3), rounding to nearest each time: 1 + q + q + q
    (q = 2^-25, a quarter ulp of 1) comes out as 1, q + q + q + 1 as
    1 + 2^-23, and 1 + 3q (one term) rounds up to 1 + 2^-23.  A term below
    half an ulp of the running sum is lost outright, the behaviour the FP32
    matrix-core pruning (k_prune32) is built around."""
    say("case: mfma32 accumulation order and rounding")
    u, q = 2.0 ** -23, 2.0 ** -25
    A = np.zeros((16, 4))
    A[0] = [1, q, q, q]
    A[1] = [q, q, q, 1]
    A[2] = [1, 3 * q, 0, 0]
    A[3] = [1, 2.0 ** -24, 0, 0]      # a tie: to even
    B = np.zeros((4, 16))
    B[:, 0] = 1.0
    _, d32 = engine.mfma_probe(A, B)
    got = [d32[r // 4, r % 4] for r in range(4)]   # column 0: lane 16 (r // 4), element r % 4
    say("  rows 0-3:", [(g - 1) / u for g in got])
    assert got == [1.0, 1.0 + u, 1.0 + u, 1.0]


@pytest.mark.parametrize("kernel", ["mfma", "valu"])
@pytest.mark.parametrize("name", fixtures.NAMES)
def test_fp64_matches_reference_and_oracle(name, kernel):
    say(f"case: fp64 {kernel} {name}")
    fx, tree, data = fixtures.load(name)
    eng = make_engine(tree, data, kernel=kernel)
    nl = eng.lnl(x_fixed(tree, fx))
    L = -nl
    assert abs(L - fx["fixed_lnL"]) / abs(L) < 1e-9
    assert abs(L - fx["fixed_lnL"]) <= 6e-7 + 1e-12 * abs(L)
    fh, site = eng.fhK()
    ref = fixtures.site_to_pattern(fx, data)
    assert np.all(np.abs(site - ref) <= 1e-10 * np.abs(ref) + 1e-10)
    Lo, fho, siteo = oracle.lnl(tree, data, *fixtures.params(fx))
    assert abs(L - Lo) / abs(Lo) < 1e-11
    assert np.all(np.abs(site - siteo) <= 1e-10 * np.abs(siteo) + 1e-12)
    # per-class values of improbable classes rest on tiny P(t) entries whose
    # eigen-route rounding differs between the two eigen solvers: 1e-8 relative
    assert np.all(np.abs(fh - fho) <= 1e-8 * np.abs(fho) + 1e-12)
    assert abs(np.dot(site, data.fpatt) - L) < 1e-12 * abs(L)
    pr = eng.last_params()
    fk, qf, _, _, _ = oracle.model(*fixtures.params(fx), data.pi)
    assert np.allclose(pr[5:7], qf, rtol=1e-13) and np.allclose(pr[7:11], fk, rtol=1e-14)


@pytest.mark.parametrize("kernel", ["mfma", "valu"])
@pytest.mark.parametrize("name", fixtures.NAMES)
def test_fp32_scaled_error(name, kernel):
    say(f"case: fp32 {kernel} {name}")
    fx, tree, data = fixtures.load(name)
    e32 = make_engine(tree, data, precision="fp32", kernel=kernel)
    L32 = -e32.lnl(x_fixed(tree, fx))
    Lo, _, siteo = oracle.lnl(tree, data, *fixtures.params(fx))
    _, site32 = e32.fhK()
    d = site32 - siteo
    say(f"  fp32 {kernel}: dlnL = {L32 - Lo:.3e} (npatt {data.npatt}); per-site mean "
        f"{np.mean(d):.3e} rms {np.sqrt(np.mean(d * d)):.3e} max {np.max(np.abs(d)):.3e}")
    # FP32 storage, FP64 arithmetic, stochastic rounding: unbiased per-site
    # errors of ~1e-7 that average out over patterns
    assert abs(L32 - Lo) < 2e-5
    assert abs(np.mean(d)) < 3e-8
    assert np.max(np.abs(d)) < 5e-6


def test_repeatable_and_cache_consistent():
    say("case: repeatable cfg1")
    fx, tree, data = fixtures.load("cfg1")
    eng = make_engine(tree, data)
    x = x_fixed(tree, fx)
    a = eng.lnl(x)
    x2 = x.copy()
    x2[-5] *= 1.1            # kappa: new eigen systems
    b = eng.lnl(x2)
    x3 = x.copy()
    x3[3] *= 1.5             # one branch: eigen systems reused
    c = eng.lnl(x3)
    assert eng.lnl(x) == a   # bitwise: deterministic reductions, cache restored
    fresh = make_engine(tree, data)
    assert fresh.lnl(x2) == b and fresh.lnl(x3) == c
    Lo, _, _ = oracle.lnl(tree, data, x2[-5], *fixtures.params(fx)[1:])
    assert abs(-b - Lo) / abs(Lo) < 1e-11


def test_parameterisations_agree():
    say("case: parameterisations cfg0")
    fx, tree, data = fixtures.load("cfg0")
    k, p0, p1, w0, w2 = fixtures.params(fx)
    t = tree.branch_vector()
    eng = make_engine(tree, data)
    a = eng.lnl(engine.x_from_params(t, k, p0, p1, w0, w2))
    eng.set_model(p_transformed=False)
    b = eng.lnl(engine.x_from_params(t, k, p0, p1, w0, w2, transformed=False))
    eng.set_model(fix_omega=True, omega_fix=w2)
    c = eng.lnl(engine.x_from_params(t, k, p0, p1, w0))
    assert abs(a - b) < 1e-10 * abs(a) and abs(a - c) < 1e-10 * abs(a)
    eng.set_model(fix_omega=True, omega_fix=1.0)       # H0
    d = eng.lnl(engine.x_from_params(t, k, p0, p1, w0))
    Lo, _, _ = oracle.lnl(tree, data, k, p0, p1, w0, 1.0)
    assert abs(-d - Lo) < 1e-9 * abs(Lo)


def test_fd_gradient_matches_oracle_differences():
    say("case: fd gradient cfg0")
    fx, tree, data = fixtures.load("cfg0")
    eng = make_engine(tree, data)
    x = x_fixed(tree, fx)
    f, g = eng.gradientB(x, central=True)
    assert abs(f - eng.lnl(x)) == 0
    # central differences of the oracle at the same steps
    k0 = tree.nbranch

    def onl(xx):
        tt = tree.branch.copy()
        tr = tree
        tr.set_branch_vector(xx[:k0])
        a, b = np.exp(xx[k0 + 1]), np.exp(xx[k0 + 2])
        v = -oracle.lnl(tr, data, xx[k0], a / (1 + a + b), b / (1 + a + b), xx[k0 + 3],
                        xx[k0 + 4])[0]
        tree.branch[:] = tt
        return v
    for i in range(len(x)):
        eh = (1e-6 * (abs(x[i]) + 1)) ** 0.67
        xp, xm = x.copy(), x.copy()
        xp[i] += eh
        xm[i] -= eh
        go = (onl(xp) - onl(xm)) / (2 * eh)
        assert abs(g[i] - go) < 2e-4 * max(1.0, abs(go)), (i, g[i], go)


def _random_case(seed, ns, ncodon, gaps=0.0, mean_bl=0.1):
    rng = np.random.default_rng(seed)
    tree = simulate.yule_tree(ns, mean_bl, rng)
    _, pi = simulate.dirichlet_f3x4(rng)
    p = simulate.uniform_params(rng)
    seqs, _ = simulate.simulate_model_a(tree, ncodon, p["kappa"], p["w0"], p["w2"], p["p0"],
                                        p["p1"], pi, rng, gap_frac=gaps)
    return tree, formats.load_codon_data(tree.names, seqs), p


def test_edge_cases_against_oracle():
    say("case: edge cases")
    tree, data, p = _random_case(7, 10, 150, gaps=0.2)
    args = (p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    # zero and tiny branch lengths (identity P), a long branch
    t = tree.branch_vector()
    t[0], t[1], t[2] = 0.0, 1e-120, 7.5
    tree.set_branch_vector(t)
    eng = make_engine(tree, data)
    L = -eng.lnl(engine.x_from_params(t, *args))
    Lo, _, siteo = oracle.lnl(tree, data, *args)
    assert abs(L - Lo) < 1e-10 * abs(Lo)
    # two foreground branches and a polytomy at the root
    names = [f"S{i + 1}" for i in range(6)]
    txt = "((S1 #1:0.1,S2:0.2):0.05 #1,S3:0.3,S4:0.1,(S5:0.2,S6:0.02):0.1);"
    tr = formats.parse_newick(txt, names)
    rng = np.random.default_rng(3)
    tab = [formats.codons.codon_str(i) for i in range(61)]
    seqs = ["".join(tab[k] for k in rng.integers(61, size=50)) + "---" for _ in range(6)]
    d = formats.load_codon_data(names, seqs)
    e2 = make_engine(tr, d)
    L2 = -e2.lnl(engine.x_from_params(tr.branch_vector(), *args))
    Lo2, _, so2 = oracle.lnl(tr, d, *args)
    assert abs(L2 - Lo2) < 1e-10 * abs(Lo2)
    _, s2 = e2.fhK()
    gap = [h for h in range(d.npatt) if d.z[:, h].min() >= 61]
    assert len(gap) == 1 and abs(s2[gap[0]]) < 1e-13


def test_full_size_properties():
    """128 taxa x 10,000 codons: the oracle is too slow here, so check
    MFMA = VALU (FP64), sum of sites = lnL, and the FP32-storage path within
    1e-4 of FP64 in lnL."""
    say("case: full size 128 x 10000")
    tree, data, p = _random_case(2024, 128, 10000, mean_bl=0.05)
    x = engine.x_from_params(tree.branch_vector(), p["kappa"], p["p0"], p["p1"], p["w0"], p["w2"])
    em = make_engine(tree, data)
    ev = make_engine(tree, data, kernel="valu")
    a, b = em.lnl(x), ev.lnl(x)
    say(f"  npatt {data.npatt} lnL {-a:.6f} mfma-valu {a - b:.3e} timing {em.timing()}")
    assert abs(a - b) < 1e-11 * abs(a)
    e128 = make_engine(tree, data, tile=128)
    assert abs(e128.lnl(x) - a) < 1e-11 * abs(a)
    _, site = em.fhK()
    assert abs(np.dot(site, data.fpatt) + a) < 1e-11 * abs(a)
    e32 = make_engine(tree, data, precision="fp32")
    c = e32.lnl(x)
    say(f"  fp32 - fp64 = {c - a:.3e}")
    assert abs(c - a) < 1e-4


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("name", fixtures.IMPOSSIBLE)
def test_impossible_sites(name, precision):
    """Sites no class can produce (sister tips on zero-length branches with
    different codons).  Every class's value there is log(1e-300), the
    reference's floor (fx_r treesub.c:7524-7534), and the site's mixture is
    log(1e-300) too.  Where the reference scales no node (8 taxa) that is
    what it prints.  Where it scales nodes (40 taxa) it adds NodeScale's
    scale logs to the floor, -800 for a scaled node whose partials are all
    below 1e-300 (treesub.c:6956-6959), so its value depends on where it
    scales; the engine's does not.  The difference at those sites is exactly
    the oracle's restatement of those scale logs (the oracle with and
    without the reference's node scaling), and every other site equals the
    reference's."""
    say(f"case: impossible sites {name} {precision}")
    fx, tree, data = fixtures.load(name)
    eng = make_engine(tree, data, precision=precision)
    nl = eng.lnl(x_fixed(tree, fx))
    fh, site = eng.fhK()
    ref = fixtures.site_to_pattern(fx, data)
    bad = np.zeros(data.npatt, dtype=bool)
    bad[fixtures.patterns_of_sites(fx, data, fx["impossible_sites"])] = True
    floor = np.log(1e-300)
    assert np.all(fh[:, bad] == floor)
    assert np.all(np.abs(site[bad] - floor) < 1e-12)
    tol = 1e-10 if precision == "fp64" else 5e-6
    ok = ~bad
    assert np.all(np.abs(site[ok] - ref[ok]) <= tol * np.abs(ref[ok]) + tol)
    _, _, scaled = oracle.lnl(tree, data, *fixtures.params(fx))
    _, _, plain = oracle.lnl(tree, data, *fixtures.params(fx), use_scaling=False)
    scale_logs = scaled[bad] - plain[bad]
    say(f"  reference - engine at impossible sites: {ref[bad] - site[bad]}; "
        f"oracle scale logs {scale_logs}")
    assert np.all(np.abs((ref[bad] - site[bad]) - scale_logs) < 1e-9)
    if name == "imposs8":
        assert np.all(scale_logs == 0)
    # lnL = the reference's, less the scale logs of the impossible sites
    want = fx["fixed_lnL"] - np.dot(data.fpatt[bad], scale_logs)
    assert abs(-nl - want) <= (6e-7 if precision == "fp64" else 2e-4)
